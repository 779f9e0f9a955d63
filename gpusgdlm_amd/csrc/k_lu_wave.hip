/*
 * k_lu_wave.hip -- log|det(I - Gamma_i)| of every importance draw for small
 * models, gfx950: one wavefront per draw, the whole M x M matrix (m rounded
 * up to 32, 64 or 128; identity padding) in registers, one row per lane
 * (two for M = 128), LU with partial pivoting.
 *
 * Replaces fillGammas + cublasXgetrfBatched + detFromLU of the VB loop
 * (rSGDLM/src/SGDLM/SGDLM.cu:243-298, kernel_functions.cuh:821-842,904-926)
 * where the matrices are too small for a workgroup per draw to pay: no
 * workgroup barriers, no LDS, no fallback pass (every draw is pivoted).
 *
 * Per draw, lane i builds row i of I - Gamma from the row's CSR entries
 * (compile-time column selects, so the row stays in registers).  Column k:
 *   - pivot: the largest |a_ik| over the rows not yet used (a wavefront max,
 *     then the lowest lane holding it: a ballot);
 *   - the pivot row is broadcast one entry at a time (readlane with the
 *     pivot lane in a scalar register);
 *   - every other unused row subtracts l_i times it, l_i = a_ik / a_pk.
 * The columns are taken in chunks of 16 whose lower bound is a
 * compile-time constant, so each chunk's updates touch only the columns
 * right of it (half the work of a full-width loop) while the code stays
 * small.  Pivot k stays in lane k % 64; log|det| = sum of log|pivot| in
 * FP64, one log per lane and a fixed-order wavefront sum at the end; a zero
 * pivot gives -inf (weight 0), as in k_lu_pivot.
 */
#include <hip/hip_runtime.h>

#include <math.h>

#include <type_traits>

#include "lu_prefix.h"
#include "sgdlm_internal.h"

namespace sgdlm {
namespace {

struct WaveArgs {
  long n_draws;
  int m;
  long P;
  const double *theta;
  const int *g_rowptr, *g_col, *g_idx;
  double *logdet;
};

template <typename T> __device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const T y = __shfl_xor(v, o);
    v = y > v ? y : v;
  }
  return v;
}

__device__ __forceinline__ float bcast(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double bcast(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

/* LU with partial pivoting of the M x M matrix held by one wavefront (row
 * lane + 64 s in row[s]; used[s] marks rows that take no part), returning
 * sum log|pivot| in FP64, summed over the wavefront in a fixed order;
 * singular: a zero pivot. */
template <typename T, int M>
__device__ __forceinline__ double wave_lu_logdet(T (&row)[(M + 63) / 64][M],
                                                 bool (&used)[(M + 63) / 64], int lane,
                                                 bool &singular) {
  constexpr int R = (M + 63) / 64; /* rows per lane */
  constexpr int G = 8;             /* pivot-row entries read per batch */
  /* pivot k is kept by lane k % 64 (slot k / 64); logs taken at the end */
  T pv[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    pv[s] = (T)1;
  singular = false;
#pragma unroll
  for (int kb = 0; kb < M; kb += 16) {
    for (int k = kb; k < kb + 16; ++k) {
      /* this lane's candidate |a_ik| over its unused rows */
      T best = (T)-1;
      int bslot = 0;
      T ak[R];
#pragma unroll
      for (int s = 0; s < R; ++s) {
        T x = (T)0;
#pragma unroll
        for (int c = kb; c < kb + 16; ++c)
          x = (c == k) ? row[s][c] : x;
        ak[s] = x;
        T v = fabs(x);
        if (used[s] || !(v >= (T)0))
          v = (T)-1;
        if (v > best) {
          best = v;
          bslot = s;
        }
      }
      /* pivot: largest candidate, lowest lane holding it */
      const T mx = wave_max(best);
      const unsigned long long hit = __ballot(best == mx);
      const int p = __builtin_amdgcn_readfirstlane(__builtin_ctzll(hit));
      const int ps = R == 1 ? 0 : __builtin_amdgcn_readfirstlane(__builtin_amdgcn_readlane(bslot, p));
      T piv = bcast(R == 1 ? ak[0] : (ps ? ak[R - 1] : ak[0]), p);
      if (!(fabs(piv) > (T)0)) {
        singular = true;
        piv = (T)1;
      }
#pragma unroll
      for (int s = 0; s < R; ++s)
        if (lane + 64 * s == k)
          pv[s] = piv;
#pragma unroll
      for (int s = 0; s < R; ++s)
        if (s == ps && lane == p)
          used[s] = true;
      T l[R];
#pragma unroll
      for (int s = 0; s < R; ++s)
        l[s] = used[s] ? (T)0 : ak[s] / piv;
      /* the pivot row's entries right of column k, G at a time (the
       * readlanes of a batch cover each other's latency); zero at and
       * left of k so the update leaves those columns unchanged */
#pragma unroll
      for (int c0 = kb; c0 < M; c0 += G) {
        T pc[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const int c = c0 + u;
          pc[u] = bcast(R == 1 ? row[0][c] : (ps ? row[R - 1][c] : row[0][c]), p);
        }
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const int c = c0 + u;
          const T v = (c >= kb + 16 || c > k) ? pc[u] : (T)0;
#pragma unroll
          for (int s = 0; s < R; ++s)
            row[s][c] = fma(-l[s], v, row[s][c]);
        }
      }
    }
  }
  /* log|det| = sum of log|pivot|: one log per lane and slot, then a
   * fixed-order wavefront sum */
  double ls = 0.0;
#pragma unroll
  for (int s = 0; s < R; ++s)
    ls += log(fabs((double)pv[s]));
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
    ls += __shfl_xor(ls, o);
  return ls;
}

template <typename T, int M, bool GRP>
__global__ __launch_bounds__(256, 1) void k_lu_wave(WaveArgs a0, const WaveArgs *__restrict__ tab,
                                                   GrpMap gm) {
  constexpr int R = (M + 63) / 64; /* rows per lane */
  const int lane = threadIdx.x & 63;
  int bid, nbk;
  const WaveArgs a = grp_args<GRP>(a0, tab, gm, bid, nbk);
  const long wave0 = ((long)bid * blockDim.x + threadIdx.x) >> 6;
  const long nwaves = ((long)nbk * blockDim.x) >> 6;
  for (long d = wave0; d < a.n_draws; d += nwaves) {
    const double *th = a.theta + (size_t)d * a.P;
    T row[R][M];
    bool used[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const int i = lane + 64 * s;
      used[s] = i >= M; /* lanes beyond M take no part */
#pragma unroll
      for (int c = 0; c < M; ++c)
        row[s][c] = (c == i) ? (T)1 : (T)0;
      if (i < a.m) {
        const int e1 = a.g_rowptr[i + 1];
        for (int e = a.g_rowptr[i]; e < e1; ++e) {
          const int col = a.g_col[e];
          const T v = (T)(-th[a.g_idx[e]]);
#pragma unroll
          for (int c = 0; c < M; ++c)
            row[s][c] = (c == col) ? v : row[s][c];
        }
      }
    }

    bool singular;
    const double ls = wave_lu_logdet<T, M>(row, used, lane, singular);
    if (lane == 0)
      a.logdet[d] = singular ? -INFINITY : ls;
  }
}

/* Dense blocks of up to M rows after the sparse prefix (k_lu_wsch): the
 * LDS of a workgroup holds the prefix tables once (tab_bytes) and, per
 * wavefront, its draw's prefix slots, a staging band of 16 rows x stride
 * and a flag word (wave_bytes each, from tab_bytes on). */
struct WschArgs {
  long n_draws;
  long P;
  const double *theta;
  double *logdet;
  int *flag_count, *flag_list; /* draws whose prefix failed: the fallback's */
  PrefixTabs pre;
  int tab_bytes, slot_bytes, wave_bytes, stride, n_band;
  /* the Schur tables by band of 16 rows (lu_symbolic.h BandPlan) */
  const int *blk_ptr, *st_off, *brounds, *bcbase, *sg_ptr, *sg_off, *sg_idx;
  const unsigned *st_pairs;
  unsigned zero_pair;
};

/* One wavefront per draw: the sparse prefix on the wavefront's own LDS
 * slots (multipliers up to kPrefixLooseMultiplier; beyond it, or at a zero
 * prefix pivot, the draw goes to the fallback list), the Schur complement
 * gathered band by band into registers (row lane + 64 s of the M x M block
 * in row[s]; rows past the separators are identity), then LU with partial
 * pivoting of the block (wave_lu_logdet).  Four wavefronts per workgroup,
 * persistent over draws; the prefix tables are loaded once per workgroup. */
template <typename T, int M, bool GRP>
__global__ __launch_bounds__(256, 1) void k_lu_wsch(WschArgs a0, const WschArgs *__restrict__ tab,
                                                   GrpMap gm) {
  constexpr int R = (M + 63) / 64;
  constexpr unsigned LO = 0xffffu;
  /* pair rounds loaded ahead; a block's rounds beyond its own count are
   * skipped by a wavefront-uniform test (no zero-pair LDS reads) */
  constexpr int RR = 4;
  typedef typename std::conditional<sizeof(T) == 4, float4, double2>::type V16;
  constexpr int VW = 16 / sizeof(T);
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int bid, nbk;
  const WschArgs a = grp_args<GRP>(a0, tab, gm, bid, nbk);
  int *tb = reinterpret_cast<int *>(smem);
  char *mine = smem + a.tab_bytes + (size_t)wave * a.wave_bytes;
  T *sl = reinterpret_cast<T *>(mine);
  T *stage = reinterpret_cast<T *>(mine + a.slot_bytes);
  int *eflag = reinterpret_cast<int *>(mine + a.wave_bytes - 16);
  for (int w = threadIdx.x; w < a.pre.tab_words; w += blockDim.x)
    tb[w] = a.pre.tab[w];
  __syncthreads();
  const long nwv = (long)nbk * (blockDim.x >> 6);
  for (long d = (long)bid * (blockDim.x >> 6) + wave; d < a.n_draws; d += nwv) {
    const double *th = a.theta + (size_t)d * a.P;
    int bad = 0;
    T amax = (T)0;
    double lsp = 0.0;
    prefix_eliminate<T, true>(sl, tb, th, a.pre, lane, 64, false, bad, amax, lsp, eflag,
                              (T)kPrefixLooseMultiplier);
    if (__ballot(bad != 0)) {
      if (lane == 0) {
        const int slot = atomicAdd(a.flag_count, 1);
        a.flag_list[slot] = (int)d;
        a.logdet[d] = NAN;
      }
      continue;
    }
    T row[R][M];
    bool used[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const int i = lane + 64 * s;
      used[s] = i >= M;
#pragma unroll
      for (int c = 0; c < M; ++c)
        row[s][c] = (c == i) ? (T)1 : (T)0;
    }
#pragma unroll
    for (int band = 0; band < M / 16; ++band) {
      if (band < a.n_band) {
        /* the band: zero, unit diagonal, entries of I - Gamma, then every
         * entry the prefix updated as the sum over its slot pairs */
        const int nv = 16 * a.stride / VW;
        V16 *s16 = reinterpret_cast<V16 *>(stage);
        V16 zero;
        if constexpr (sizeof(T) == 4)
          zero = make_float4(0.f, 0.f, 0.f, 0.f);
        else
          zero = make_double2(0.0, 0.0);
        for (int q = lane; q < nv; q += 64)
          s16[q] = zero;
        prefix_sync<true>();
        if (lane < 16)
          stage[lane * a.stride + 16 * band + lane] = (T)1;
        const int g1 = a.sg_ptr[band + 1];
        for (int e = a.sg_ptr[band] + lane; e < g1; e += 64)
          stage[a.sg_off[e]] = (T)(-th[a.sg_idx[e]]);
        prefix_sync<true>();
        const int b1 = a.blk_ptr[band + 1];
        for (int b = a.blk_ptr[band]; b < b1; ++b) {
          const int off = a.st_off[(size_t)b * 64 + lane];
          const int nr = a.brounds[b], cb = a.bcbase[b];
          T acc = (T)0;
          for (int r0 = 0; r0 < nr; r0 += RR) {
            unsigned pv[RR];
#pragma unroll
            for (int r = 0; r < RR; ++r)
              pv[r] = r0 + r < nr ? a.st_pairs[(size_t)cb + 64 * (r0 + r) + lane] : a.zero_pair;
#pragma unroll
            for (int r = 0; r < RR; ++r)
              if (r0 + r < nr)
                acc = fma(-sl[pv[r] & LO], sl[pv[r] >> 16], acc);
          }
          if (off >= 0)
            stage[off] = acc;
        }
        prefix_sync<true>();
        /* the band's rows into the registers of their lanes (register
         * slot s of rows 16 band ..: a compile-time index in each branch) */
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const int rl = lane + 64 * s - 16 * band;
          if (s == (16 * band) / 64 && rl >= 0 && rl < 16) {
            const V16 *src = reinterpret_cast<const V16 *>(stage + rl * a.stride);
#pragma unroll
            for (int q = 0; q < M / VW; ++q) {
              const V16 v = src[q];
              if constexpr (sizeof(T) == 4) {
                row[s][4 * q] = v.x;
                row[s][4 * q + 1] = v.y;
                row[s][4 * q + 2] = v.z;
                row[s][4 * q + 3] = v.w;
              } else {
                row[s][2 * q] = v.x;
                row[s][2 * q + 1] = v.y;
              }
            }
          }
        }
        prefix_sync<true>();
      }
    }
    bool singular;
    const double lsd = wave_lu_logdet<T, M>(row, used, lane, singular);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
      lsp += __shfl_xor(lsp, o);
    if (lane == 0)
      a.logdet[d] = singular ? -INFINITY : lsp + lsd;
  }
}

template <typename T, int M> int launch_wave(const WaveArgs &a, int grid, hipStream_t st) {
  const WaveArgs *tab = nullptr;
  hipLaunchKernelGGL((k_lu_wave<T, M, false>), dim3(grid), dim3(256), 0, st, a, tab, GrpMap());
  return hipGetLastError() == hipSuccess ? SGDLM_OK : SGDLM_ERR_DEVICE;
}

template <typename T, int M> int resident_blocks() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void *)k_lu_wave<T, M, false>, 256, 0) !=
          hipSuccess ||
      n < 1) {
    (void)hipGetLastError();
    n = 1;
  }
  return n;
}

} // namespace

int lu_wave_dim(int M, int prec) {
  if (M <= 32)
    return 32;
  if (M <= 64)
    return 64;
  if (M <= 128 && prec != SGDLM_PREC_FP64)
    return 128;
  return 0;
}

int lu_wave_grid(int wave_M, int prec, int n_cu) {
  int per_cu = 1;
  const bool f64 = prec == SGDLM_PREC_FP64;
  if (wave_M == 32)
    per_cu = f64 ? resident_blocks<double, 32>() : resident_blocks<float, 32>();
  else if (wave_M == 64)
    per_cu = f64 ? resident_blocks<double, 64>() : resident_blocks<float, 64>();
  else
    per_cu = resident_blocks<float, 128>();
  return std::max(1, n_cu) * per_cu;
}

namespace {
WaveArgs wave_args(const Structure &S, long N, const double *theta, double *logdet) {
  WaveArgs a;
  a.n_draws = N;
  a.m = S.m;
  a.P = S.P;
  a.theta = theta;
  a.g_rowptr = S.g_rowptr;
  a.g_col = S.g_col;
  a.g_idx = S.g_idx;
  a.logdet = logdet;
  return a;
}
} // namespace

int lu_grp_wave(const Structure &S, const LuPlan &plan, long N, const double *theta,
                double *logdet, LuGrpKernel *k) {
  const WaveArgs a = wave_args(S, N, theta, logdet);
  const bool f64 = plan.prec == SGDLM_PREC_FP64;
  k->fn = nullptr;
  switch (plan.wave_M) {
  case 32:
    k->fn = f64 ? (const void *)k_lu_wave<double, 32, true> : (const void *)k_lu_wave<float, 32, true>;
    break;
  case 64:
    k->fn = f64 ? (const void *)k_lu_wave<double, 64, true> : (const void *)k_lu_wave<float, 64, true>;
    break;
  case 128:
    if (!f64)
      k->fn = (const void *)k_lu_wave<float, 128, true>;
    break;
  }
  if (!k->fn)
    return SGDLM_ERR_ARG;
  k->threads = 256;
  k->per_cu = std::max(1, plan.wave_grid / std::max(1, plan.grid));
  k->lds = 0;
  k->ws_wg = 0;
  k->units = (N + 3) / 4; /* four wavefronts (draws in flight) per workgroup */
  /* ms per draw on one workgroup of four wavefronts (size model) */
  const double Md = (double)plan.wave_M;
  k->cost = (double)N * 2.5e-5 * Md * (1.0 + Md / 64.0);
  k->args.assign(reinterpret_cast<const unsigned char *>(&a),
                 reinterpret_cast<const unsigned char *>(&a) + sizeof(a));
  k->work_field = -1;
  return SGDLM_OK;
}

namespace {

/* LDS bytes of one wavefront of k_lu_wsch: its prefix slots, the staging
 * band (16 rows x stride) and a flag word. */
size_t wsch_wave_bytes(const Structure &S, size_t ts) {
  const size_t slots = ((size_t)S.e_ns * ts + 15) & ~(size_t)15;
  return slots + 16 * (size_t)S.e_wb_stride * ts + 16;
}
size_t wsch_tab_bytes(const Structure &S) { return ((size_t)S.e_tab_words * 4 + 15) & ~(size_t)15; }

WschArgs wsch_args(const Structure &S, int prec, long N, const double *theta, double *logdet,
                   int *flags) {
  const size_t ts = prec == SGDLM_PREC_FP64 ? 8 : 4;
  WschArgs a{};
  a.n_draws = N;
  a.P = S.P;
  a.theta = theta;
  a.logdet = logdet;
  a.flag_count = flags;
  a.flag_list = flags + 1;
  a.pre = PrefixTabs{S.e_ns, S.e_waves, S.e_n1, S.e_tab_words, S.e_o_piv, S.e_o_mptr, S.e_o_mult,
                     S.e_o_tptr, S.e_o_tslot, S.e_o_cptr, S.e_o_contrib, S.e_slot_init, S.e_tab};
  a.tab_bytes = (int)wsch_tab_bytes(S);
  a.slot_bytes = (int)(((size_t)S.e_ns * ts + 15) & ~(size_t)15);
  a.wave_bytes = (int)wsch_wave_bytes(S, ts);
  a.stride = S.e_wb_stride;
  a.n_band = S.e_wb_nband;
  a.blk_ptr = S.e_wb_blk_ptr;
  a.st_off = S.e_wb_st_off;
  a.brounds = S.e_wb_brounds;
  a.bcbase = S.e_wb_bcbase;
  a.sg_ptr = S.e_wb_sg_ptr;
  a.sg_off = S.e_wb_sg_off;
  a.sg_idx = S.e_wb_sg_idx;
  a.st_pairs = S.e_wb_pairs;
  a.zero_pair = S.e_zero_pair;
  return a;
}

struct WschKern {
  const void *one = nullptr, *grp = nullptr;
};
template <typename T, int M> WschKern wsch_k() {
  return {(const void *)k_lu_wsch<T, M, false>, (const void *)k_lu_wsch<T, M, true>};
}
WschKern wsch_kern(int M, int prec) {
  const bool f64 = prec == SGDLM_PREC_FP64;
  switch (M) {
  case 32:
    return f64 ? wsch_k<double, 32>() : wsch_k<float, 32>();
  case 64:
    return f64 ? wsch_k<double, 64>() : wsch_k<float, 64>();
  case 128:
    if (!f64)
      return wsch_k<float, 128>();
    break;
  }
  return WschKern();
}

} // namespace

size_t lu_wsch_lds(const Structure &S, int prec) {
  if (S.e_wb_M == 0 || (prec == SGDLM_PREC_FP64 && S.e_wb_M > 64))
    return 0;
  const size_t ts = prec == SGDLM_PREC_FP64 ? 8 : 4;
  return wsch_tab_bytes(S) + 4 * wsch_wave_bytes(S, ts);
}

int lu_wsch_per_cu(const Structure &S, int prec, size_t lds) {
  const WschKern k = wsch_kern(S.e_wb_M, prec);
  int n = 0;
  if (!k.one || allow_dynamic_lds(k.one, lds) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k.one, 256, lds) != hipSuccess || n < 1) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int launch_lu_wsch(const Structure &S, const LuPlan &plan, long N, const double *theta,
                   double *logdet, int *flags, hipStream_t st) {
  WschArgs a = wsch_args(S, plan.prec, N, theta, logdet, flags);
  const WschKern k = wsch_kern(plan.wsch_M, plan.prec);
  if (!k.one || plan.wsch_M != S.e_wb_M)
    return SGDLM_ERR_ARG;
  if (allow_dynamic_lds(k.one, plan.wsch_lds) != hipSuccess)
    return SGDLM_ERR_DEVICE;
  const long need = (N + 3) / 4; /* four wavefronts per workgroup */
  const int grid = (int)std::max(1L, std::min<long>(plan.wsch_grid, need));
  const WschArgs *tab = nullptr;
  GrpMap gm;
  void *args[] = {&a, &tab, &gm};
  if (hipLaunchKernel(k.one, dim3(grid), dim3(256), args, plan.wsch_lds, st) != hipSuccess)
    return SGDLM_ERR_DEVICE;
  return hipGetLastError() == hipSuccess ? SGDLM_OK : SGDLM_ERR_DEVICE;
}

int lu_grp_wsch(const Structure &S, const LuPlan &plan, long N, const double *theta,
                double *logdet, int *flags, LuGrpKernel *k) {
  const WschArgs a = wsch_args(S, plan.prec, N, theta, logdet, flags);
  const WschKern wk = wsch_kern(plan.wsch_M, plan.prec);
  if (!wk.grp)
    return SGDLM_ERR_ARG;
  k->fn = wk.grp;
  k->threads = 256;
  k->per_cu = std::max(1, plan.wsch_grid / std::max(1, plan.grid));
  k->lds = plan.wsch_lds;
  k->ws_wg = 0;
  k->units = (N + 3) / 4;
  /* ms per draw on one workgroup of four wavefronts (size model) */
  const double Md = (double)plan.wsch_M;
  k->cost = (double)N * (3e-5 * Md * (1.0 + Md / 64.0) + 5e-7 * (double)S.e_fma);
  k->args.assign(reinterpret_cast<const unsigned char *>(&a),
                 reinterpret_cast<const unsigned char *>(&a) + sizeof(a));
  k->work_field = -1;
  return SGDLM_OK;
}

int launch_lu_wave(const Structure &S, const LuPlan &plan, long N, const double *theta,
                   double *logdet, hipStream_t st) {
  const WaveArgs a = wave_args(S, N, theta, logdet);
  const long need = (N + 3) / 4; /* four wavefronts per workgroup */
  const int grid = (int)std::max(1L, std::min<long>(plan.wave_grid, need));
  const bool f64 = plan.prec == SGDLM_PREC_FP64;
  switch (plan.wave_M) {
  case 32:
    return f64 ? launch_wave<double, 32>(a, grid, st) : launch_wave<float, 32>(a, grid, st);
  case 64:
    return f64 ? launch_wave<double, 64>(a, grid, st) : launch_wave<float, 64>(a, grid, st);
  case 128:
    if (!f64)
      return launch_wave<float, 128>(a, grid, st);
    break;
  }
  return SGDLM_ERR_ARG;
}

} // namespace sgdlm
