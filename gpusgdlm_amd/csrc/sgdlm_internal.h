/*
 * sgdlm_internal.h -- declarations shared by the HIP translation units of
 * libsgdlm_hip.so (kernel launchers and the device-side problem layout).
 */
#ifndef SGDLM_INTERNAL_H
#define SGDLM_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/sgdlm.h"

namespace sgdlm {

/* Simultaneous-parent structure in the form the kernels use.  The packed
 * draw layout stores only the p_j meaningful coefficients of each series:
 * theta_packed[i * P + off[j] + k], k < p_j. */
struct Structure {
  int m = 0, max_p = 0;
  int M = 0;            /* m rounded up to a multiple of 16 (LU dimension)   */
  long P = 0;           /* sum_j p_j                                          */
  int n_gam = 0;        /* number of simultaneous coefficients (Gamma nnz)   */
  const unsigned *p = nullptr;   /* device, m                                */
  const unsigned *sp = nullptr;  /* device, m * max_p                        */
  const int *off = nullptr;      /* device, m + 1 prefix sums of p           */
  /* Gamma entries sorted by (row, col): A[row][col] = -theta_packed[idx]   */
  const int *g_row = nullptr, *g_col = nullptr, *g_idx = nullptr;
  const int *g_rowptr = nullptr; /* device, m + 1 (CSR over rows)           */
  /* Sparse elimination prefix of the log-determinant LU (lu_symbolic.h):
   * host sizes and device tables. */
  int e_n1 = 0, e_n2 = 0, e_M2 = 0, e_ns = 0, e_waves = 0, e_nsg = 0, e_nblk = 0;
  int e_tab_words = 0;
  int e_o_piv = 0, e_o_mptr = 0, e_o_mult = 0, e_o_tptr = 0, e_o_tslot = 0, e_o_cptr = 0,
      e_o_contrib = 0, e_o_brounds = 0, e_o_bcbase = 0;
  uint32_t e_zero_pair = 0;
  long long e_fma = 0;
  const int *e_slot_init = nullptr, *e_tab = nullptr;
  const int *e_sg_off = nullptr, *e_sg_idx = nullptr, *e_st_off = nullptr;
  const uint32_t *e_st_pairs = nullptr;
  const int *e_wave_q = nullptr, *e_node_of = nullptr, *e_fwd_ptr = nullptr, *e_bwd_ptr = nullptr;
  const uint32_t *e_fwd = nullptr, *e_bwd = nullptr;
  /* The Schur tables regrouped by LDS staging band for the
   * register-resident LU (lu_symbolic.h BandPlan); e_rb_rows = 0: that
   * kernel does not fit this structure. */
  int e_rb_rows = 0, e_rb_stride = 0, e_rb_nband = 0, e_rb_tiles = 0;
  size_t e_rb_lds = 0, e_rb_panel_base = 0, e_rb_flag_base = 0;
  /* wavefronts of the register kernel's shape (8: one workgroup per CU; 4:
   * blocks of up to 12 tiles, two workgroups per CU, the prefix slots,
   * tables and staging band aliasing the panel region, the tables loaded
   * per draw); LDS offsets of the prefix region and the staging band */
  int e_rb_nw = 8, e_rb_tab_persist = 1;
  size_t e_rb_pre_base = 0, e_rb_stage_base = 0;
  const int *e_rb_blk_ptr = nullptr, *e_rb_st_off = nullptr, *e_rb_brounds = nullptr,
            *e_rb_bcbase = nullptr, *e_rb_sg_ptr = nullptr, *e_rb_sg_off = nullptr,
            *e_rb_sg_idx = nullptr;
  const uint32_t *e_rb_pairs = nullptr;
  /* The same for the one-wavefront LU after the prefix (k_lu_wsch): bands
   * of 16 rows, row stride e_wb_stride, dense block e_wb_M = M2 rounded up
   * to 32, 64 or 128 (0: not available). */
  int e_wb_M = 0, e_wb_stride = 0, e_wb_nband = 0;
  const int *e_wb_blk_ptr = nullptr, *e_wb_st_off = nullptr, *e_wb_brounds = nullptr,
            *e_wb_bcbase = nullptr, *e_wb_sg_ptr = nullptr, *e_wb_sg_off = nullptr,
            *e_wb_sg_idx = nullptr;
  const uint32_t *e_wb_pairs = nullptr;
};

/* Workgroup map of one launch over several independent models: workgroup b
 * works for model inst[b] as that model's workgroup b - base[inst[b]] of
 * count[inst[b]] (device arrays). */
struct GrpMap {
  const int *inst = nullptr, *base = nullptr, *count = nullptr;
};

/* A kernel's arguments, local workgroup index and workgroup count: its own
 * (GRP = false) or, in a launch over several models, those of the model
 * this workgroup serves (tab[inst]). */
template <bool GRP, typename A>
__device__ __forceinline__ A grp_args(const A &a0, const A *__restrict__ tab, const GrpMap &gm,
                                      int &bid, int &nbk) {
  if constexpr (GRP) {
    const int i = gm.inst[blockIdx.x];
    bid = (int)blockIdx.x - gm.base[i];
    nbk = gm.count[i];
    return tab[i];
  } else {
    (void)tab;
    (void)gm;
    bid = (int)blockIdx.x;
    nbk = (int)gridDim.x;
    return a0;
  }
}

/* One model of a VB step run over several independent models at once
 * (device table entry; the kernels of every stage read it by model index). */
struct VbModel {
  int m = 0, max_p = 0, chunks = 0;
  int ser0 = 0;                  /* index of its series 0 among all models' */
  long N = 0, P = 0;
  uint64_t seed = 0;
  const unsigned *p = nullptr;
  const int *off = nullptr;
  const double *m_t = nullptr, *C = nullptr, *n_t = nullptr, *s_t = nullptr;
  double *U = nullptr, *lam = nullptr, *theta = nullptr, *logdet = nullptr;
  int *lu_flags = nullptr;       /* flags[0] = flag_init before the LU      */
  int flag_init = 0;
  /* >= 0: flags[0] after the LU (the register kernel with partial
   * pivoting factors every draw with row interchanges: all N count) */
  int flag_final = -1;
  /* VbBuffers of the model */
  double *w = nullptr, *stats = nullptr, *part1 = nullptr, *part2 = nullptr;
  double *mean_lambda = nullptr, *mean_log_lambda = nullptr, *mean_m = nullptr,
         *mean_C = nullptr, *mean_n = nullptr, *mean_s = nullptr, *mean_Q = nullptr;
  int *newton_iter = nullptr;
};

/* Lets kernel fn take up to lds bytes of dynamic LDS (the device's largest
 * per workgroup, set once per kernel and process, so launches need no
 * attribute call of their own). */
hipError_t allow_dynamic_lds(const void *fn, size_t lds);

/* ---- filter (k_filter.hip) ---- */
void launch_posterior(const Structure &S, double *m_t, double *C_t,
                      double *n_t, double *s_t, const double *y,
                      const double *F, hipStream_t st);
void launch_prior(const Structure &S, double *m_t, double *C_t, double *n_t,
                  double *W_t, const double *beta, const double *delta,
                  const double *G, hipStream_t st);

/* ---- draws (k_draws.hip) ---- */
void launch_cholesky(const Structure &S, const double *C, double *U,
                     hipStream_t st);
/* epoch_dev (optional): the epoch is read from this device word instead
 * of the argument (a captured graph replayed at successive epochs). */
void launch_draw(const Structure &S, long N, const double *m_t,
                 const double *U, const double *n_t, const double *s_t,
                 uint64_t seed, uint32_t epoch, double *lam, double *theta,
                 hipStream_t st, const uint32_t *epoch_dev = nullptr);
void launch_forecast_rhs(const Structure &S, long N, const double *x,
                         const double *m_t, const double *theta,
                         const double *lam, uint64_t seed, uint32_t epoch,
                         double *rhs, hipStream_t st);
void launch_evolve(const Structure &S, long N, const double *U_W,
                   const double *r_t, const double *c_t, const double *beta,
                   const double *G, const double *m_t, uint64_t seed,
                   uint32_t epoch, double *lam, double *theta,
                   hipStream_t st);
void launch_unpack_theta(const Structure &S, long N, const double *theta,
                         const double *m_t, double *theta_full,
                         hipStream_t st);

/* ---- LU (k_lu.hip) ---- */
struct LuPlan {
  int prec = SGDLM_PREC_FP64; /* element type / update path                */
  int nb = 0;                 /* panel width, no-interchange fast kernel    */
  int fast_M = 0;             /* dense dimension of the fast kernel (the
                                 Schur complement after the sparse prefix)  */
  int pivot_nb = 0;           /* panel width, partial-pivoting kernel       */
  int grid = 0;               /* CUs (persistent workgroups per pass)       */
  int pivot_grid = 0;         /* persistent workgroups, pivoting kernel     */
  int threads = 512;          /* workgroup size of the fast kernel          */
  int fast_grid = 0;          /* persistent workgroups of the fast kernel   */
  size_t lds = 0;             /* dynamic LDS bytes, fast kernel             */
  size_t pre_base = 0;        /* LDS byte offset of the prefix slots/tables */
  int tab_persist = 0;        /* prefix tables stay in LDS across draws     */
  size_t xs_base = 0;         /* LDS offset of the solve vector (solves)    */
  size_t flag_off = 0;        /* LDS offset of the per-draw early-exit flag */
  /* solves with the prefix region aliased by the dense one: the prefix
   * slots are kept in global memory across the dense LU */
  size_t slot_save_off = 0;   /* byte offset in the workspace, 0 = unused  */
  size_t slot_stride = 0;     /* bytes per workgroup                        */
  size_t pivot_lds = 0;       /* dynamic LDS bytes, pivoting kernel         */
  /* log-determinants: the pivoting kernel re-eliminates the sparse prefix
   * and factors the pivot_schur_M x pivot_schur_M Schur complement with row
   * interchanges (0: the whole matrix); LDS offset of its prefix flag */
  int pivot_schur_M = 0;
  size_t pivot_flag_off = 0;
  /* small models: one wavefront per draw over the whole matrix in
   * registers, wave_M rows (k_lu_wave.hip); 0 = not used */
  int wave_M = 0, wave_grid = 0;
  /* dense blocks after the prefix: one wavefront per draw on the wsch_M
   * rows of the Schur complement (k_lu_wsch); 0 = not used.  Flagged draws
   * go to the partial-pivoting kernel. */
  int wsch_M = 0, wsch_grid = 0;
  size_t wsch_lds = 0;
  size_t work_bytes = 0;      /* global workspace bytes                      */
  /* register-resident LU of the dense block (k_lu_rs.hip) instead of the
   * fast kernel: rs_MT > 0 */
  int rs_MT = 0, rs_grid = 0;
  size_t rs_lds = 0, rs_panel_base = 0, rs_flag_base = 0;
  /* the same kernel with partial pivoting (SGDLM_LU_REGISTER_PIVOT, FP32
   * log-determinants): rs_piv = 1, row-exchange buffer at LDS byte
   * rs_swap_base */
  int rs_piv = 0;
  size_t rs_swap_base = 0;
  /* wavefronts per workgroup of the register kernel's shape (Structure
   * e_rb_nw); 4: two workgroups per CU */
  int rs_nw = 8;
  /* forecast solves with the register kernel: draws in chunks of
   * rs_chunk, each draw's factors (rs_sol_stride floats) in the workspace
   * from byte rs_sol_off, then one flag word per chunk draw; the
   * substitution kernel's launch shape; chunks are whole multiples of
   * rs_unit draws (whole rounds of both kernels' grids) */
  int rs_chunk = 0, subst_grid = 0, rs_unit = 0;
  size_t rs_sol_stride = 0, rs_sol_off = 0, subst_lds = 0;
};
/* Largest dense Schur-complement size (multiple of 16) for which the
 * 64-wide fast LU fits: FP32 in half a CU's LDS, FP64 in a whole CU's. */
int lu_dense_cap(size_t lds_max);
/* want_threads / want_nb: 0 = automatic, else the fast kernel's workgroup
 * size (256 or 512) and panel width; want_kind: SGDLM_LU_AUTO or the
 * kernel kind to use (SGDLM_LU_BLOCKED / _REGISTER / _PIVOTING).  An
 * infeasible request fails with SGDLM_ERR_ARG. */
int lu_plan(const Structure &S, int prec, int solve, int n_cu, size_t lds_max,
            int want_threads, int want_nb, LuPlan *plan, int want_kind = SGDLM_LU_AUTO);
/* SGDLM_LU_* kind of the kernel that factors the dense block under plan. */
int lu_plan_kind(const LuPlan &plan);
/* One-wavefront-per-draw LU (k_lu_wave.hip): its matrix dimension for a
 * model of M rows (m rounded up to 16) at precision prec (0: too large),
 * the workgroups that stay resident on n_cu CUs, and the launch. */
int lu_wave_dim(int M, int prec);
int lu_wave_grid(int wave_M, int prec, int n_cu);
int launch_lu_wave(const Structure &S, const LuPlan &plan, long N, const double *theta,
                   double *logdet, hipStream_t st);
/* The one-wavefront LU after the sparse prefix (k_lu_wave.hip k_lu_wsch):
 * its dynamic LDS for a structure at precision prec (0: it does not apply)
 * and resident workgroups per CU; the launch appends the draws whose prefix
 * fails to flags. */
size_t lu_wsch_lds(const Structure &S, int prec);
int lu_wsch_per_cu(const Structure &S, int prec, size_t lds);
int launch_lu_wsch(const Structure &S, const LuPlan &plan, long N, const double *theta,
                   double *logdet, int *flags, hipStream_t st);
/* *p = v on stream st by a one-thread kernel (capturable in a graph,
 * unlike a memset node). */
int set_device_int(int *p, int v, hipStream_t st);
/* Shrinks the persistent grids of a plan (the kernels loop over draws, so
 * fewer workgroups change only the speed) until its workspace fits in
 * work_cap bytes; SGDLM_ERR_ARG if not even one workgroup fits. */
int lu_plan_cap(const Structure &S, size_t work_cap, LuPlan *plan);
/* LDS layout of the register-resident LU for a dense block of MT 16 x 16
 * tiles: staged panel tiles (bytes, k_lu_rs.hip), and the largest number
 * of tiles per dimension it handles. */
size_t rs_panel_bytes(int MT);
/* row stride (floats) of the register-resident LU's LDS staging band, and
 * the tile count it pads a dense block of MT tiles to (identity) */
int rs_stage_stride(int MT, int nw = 8);
int rs_padded_tiles(int MT, int nw = 8);
constexpr int kRsMaxTiles = 18;
/* Register-resident LU log-determinants (plan.rs_MT > 0); flagged draws are
 * appended to flags for the partial-pivoting kernel. */
int launch_lu_rs(const Structure &S, const LuPlan &plan, long N, const double *theta,
                 double *logdet, int *flags, hipStream_t st);
/* flags: N + 1 ints of scratch (count + list of draws handed to the
 * partial-pivoting kernel). */
int launch_lu_logdet(const Structure &S, const LuPlan &plan, long N,
                     const double *theta, void *work, double *logdet,
                     int *flags, hipStream_t st);
/* One model's kernel of an LU stage in a launch over several models (its
 * group form reads the model's argument struct from a table and serves it
 * with the workgroups a GrpMap assigns). */
struct LuGrpKernel {
  const void *fn = nullptr;  /* group form of the kernel                   */
  int threads = 0;           /* workgroup size                             */
  int per_cu = 1;            /* workgroups resident per CU                 */
  size_t lds = 0;            /* dynamic LDS bytes of this model's layout   */
  size_t ws_wg = 0;          /* workspace bytes per workgroup (0: none)    */
  long units = 0;            /* most workgroups worth assigning            */
  double cost = 0.0;         /* work of the model: workgroup-ms            */
  std::vector<unsigned char> args; /* the kernel's argument struct         */
  long work_field = -1;      /* byte offset of its workspace pointer in args */
};
/* The LU stages of one model's log-determinants in a launch over several
 * models: *flag_init is stored to flags[0] first (the count the fallback
 * kernel reads), then `first` runs (one-wavefront, register, fast or
 * pivot-everything kernel) and, when second->fn is set, the
 * partial-pivoting kernel on the draws `first` flagged.  cost = N x the
 * time of one draw on one workgroup in ms: draw_wg_ms when measured (> 0),
 * else a size model. */
int lu_grp_stages(const Structure &S, const LuPlan &plan, long N, const double *theta,
                  double *logdet, int *flags, double draw_wg_ms, int *flag_init,
                  LuGrpKernel *first, LuGrpKernel *second);
int lu_grp_rs(const Structure &S, const LuPlan &plan, long N, const double *theta,
              double *logdet, int *flags, LuGrpKernel *k);
int lu_grp_wave(const Structure &S, const LuPlan &plan, long N, const double *theta,
                double *logdet, LuGrpKernel *k);
int lu_grp_wsch(const Structure &S, const LuPlan &plan, long N, const double *theta,
                double *logdet, int *flags, LuGrpKernel *k);
/* Forecast solves with the register kernel (k_lu_rs.hip): the factors of
 * a chunk of draws written out (rs_solve_stride floats per draw), then one
 * wavefront per draw substitutes (k_lu_subst). */
size_t rs_solve_stride(const Structure &S, int MT);
int lu_subst_shape(const Structure &S, int MT, int n_cu, int *grid, size_t *lds);
int launch_lu_rs_factor(const Structure &S, const LuPlan &plan, long n, const double *theta,
                        float *sol, int *sol_ok, int *flags, hipStream_t st);
int launch_lu_subst(const Structure &S, const LuPlan &plan, long n, const double *theta,
                    const double *rhs, double *y, const float *sol, const int *sol_ok,
                    int refine, hipStream_t st);
/* flags: N + 2 ints of scratch (count, list, and the total over chunks). */
int launch_lu_solve(const Structure &S, const LuPlan &plan, long N,
                    const double *theta, void *work, const double *rhs,
                    double *y, int refine, int *flags, hipStream_t st);

/* ---- VB (k_vb.hip) ---- */
struct VbBuffers {
  double *w = nullptr;        /* N normalised weights                      */
  double *stats = nullptr;    /* [0] max logdet, [1] sum exp(l - max),
                                 [2] sum w^2, [3] number of zero weights     */
  double *part1 = nullptr;    /* chunks * m: sum w log lambda               */
  double *part2 = nullptr;    /* chunks * m * (max_p + 1)^2: shifted second
                                 moments with the first ones (k_mom)        */
  double *mean_lambda = nullptr, *mean_log_lambda = nullptr; /* m each      */
  double *mean_m = nullptr;   /* m * max_p                                   */
  double *mean_C = nullptr;   /* m * max_p^2                                 */
  double *mean_n = nullptr, *mean_s = nullptr, *mean_Q = nullptr; /* m      */
  int *newton_iter = nullptr; /* m                                           */
  int chunks = 0;
};
int vb_chunks(long N);
void launch_weights(long N, const double *logdet, double *w, double *stats,
                    hipStream_t st);
void launch_moments(const Structure &S, long N, const double *lam,
                    const double *theta, const double *m_t, const double *n_t,
                    const double *s_t, const VbBuffers &B, hipStream_t st);

/* ---- a VB step over several models (one launch per stage; tab: device
 * table of n models, ser_model: model of each of the n_series series) ---- */
void launch_cholesky_grp(const VbModel *tab, const int *ser_model, int n_series, int max_p,
                         hipStream_t st);
int launch_draw_grp(const VbModel *tab, const int *ser_model, const uint32_t *epochs,
                    int n_series, long max_N, int max_p, hipStream_t st);
void launch_weights_grp(const VbModel *tab, int n, hipStream_t st);
int launch_moments_grp(const VbModel *tab, const int *ser_model, int n_series, int max_chunks,
                       int max_p, hipStream_t st);
void launch_grp_stats(const VbModel *tab, int n, double *out, hipStream_t st);

} // namespace sgdlm

#endif
