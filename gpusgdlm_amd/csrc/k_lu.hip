/*
 * k_lu.hip -- batched LU of I - Gamma_i for every importance / forecast
 * draw i, on gfx950.
 *
 * Replaces fillGammas + cublasXgetrfBatched + detFromLU of the VB loop
 * (rSGDLM/src/SGDLM/SGDLM.cu:232-298, kernel_functions.cuh:821-842,904-926)
 * and fillGammas + getrf + getri + gemm of the forecast
 * (SGDLM.cu:556-620), which the reference runs on cuBLAS with the matrices
 * round-tripping HBM.
 *
 * Both kernels are persistent (one or two workgroups per CU, looping over
 * draws); the matrix of a draw lives in a per-workgroup global workspace
 * (row-major, M x M, M = m rounded up to 16 with identity padding) that
 * stays in L2 / Infinity Cache, and is assembled there from the packed
 * coefficient draws.
 *
 * k_lu_fast (VB log-determinants and forecast solves; two 256-thread
 * workgroups per CU, or one of 512 threads when its LDS needs a whole CU):
 *   Sparse prefix (tables from lu_symbolic.cpp): the low-degree series of
 *   the parental graph are eliminated first, wave by wave, on LDS slots
 *   holding only the entries their elimination touches; the Schur complement
 *   on the remaining separator series is gathered into the dense workspace
 *   (M = separator count rounded up to 16) and factored by the blocked LU
 *   below.  At m = 400 with up to 20 parents per series this removes a third
 *   of the series from the dense LU (about 3x fewer flops).
 *   Threshold partial pivoting keeps the diagonal pivot while every
 *   multiplier satisfies |l_ij| <= 1/tau (LAPACK's partial pivoting is
 *   tau = 1 and then performs no interchange exactly when all |l_ij| <= 1).
 *   For draws of I - Gamma the multipliers stay near or below 1, so this
 *   kernel factors without interchanges (no pivot searches, one barrier per
 *   column: the panel rows live in registers in sub-panels of 32 columns,
 *   the pivot row is broadcast through LDS; FP32 multipliers through the
 *   hardware reciprocal), tracks max |l_ij|, and hands any draw with
 *   max |l| > 1/tau = 2 or a zero pivot to k_lu_pivot through a device-side
 *   list.
 *     - U12 = L11^{-1} A12 per 16-column tile on MFMA from the inverses of
 *       the 16 x 16 diagonal blocks of L11 (written back in place of A12);
 *     - A22 -= L21 U12 on MFMA: f64 16x16x4 (FP64), f32 16x16x4 (FP32), or
 *       bf16 16x16x32 hi*hi + hi*lo + lo*hi with FP32 accumulation (BF16x3);
 *       work items (column-tile pair, row tile) are split into one
 *       contiguous range per wavefront and their loads are issued one batch
 *       ahead.
 *   The contraction index of every MFMA is permuted so that an accumulator
 *   tile can be reused directly as a B operand (rho() below).
 * k_lu_pivot (fallback; 256 threads, two workgroups per CU when the panel
 *   fits half a CU's LDS):
 *   the same blocked algorithm with partial pivoting in the panel (largest
 *   |a|, lowest row on ties, as LAPACK getf2); interchanges permute a
 *   logical->physical row map in LDS instead of moving rows.  Forecast mode
 *   keeps L and U and solves L U y = P rhs with FP64 accumulation, then
 *   optionally refines once with the FP64 residual rhs - (I - Gamma) y on
 *   the sparse Gamma.
 * log|det| = sum log|u_kk| is accumulated in FP64; a zero or non-finite
 * pivot after pivoting gives log|det| = -inf (weight 0, as detFromLU's
 * NaN/inf -> 0).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <utility>
#include <type_traits>
#include <map>
#include <math.h>
#include <mutex>

#include "lu_prefix.h"
#include "lu_symbolic.h"
#include "sgdlm_internal.h"

namespace sgdlm {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

typedef double f64x2 __attribute__((ext_vector_type(2)));

template <typename T> struct Vec4;
template <> struct Vec4<float> { typedef f32x4 type; };
template <> struct Vec4<double> { typedef f64x4 type; };
/* 16-byte vector of T */
template <typename T> struct Vec16;
template <> struct Vec16<float> { typedef f32x4 type; static constexpr int n = 4; };
template <> struct Vec16<double> { typedef f64x2 type; static constexpr int n = 2; };

__device__ __forceinline__ f32x4 mfma16(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4 mfma16(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma_bf16(u16x8 a, u16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b),
                                                 c, 0, 0, 0);
}

/* Row (within a 16 x 16 tile) of accumulator register r in lane group
 * g = lane / 16; the column is lane % 16.  v_mfma_f32_16x16x4f32 and
 * v_mfma_f32_16x16x32_bf16: row = 4g + r.  v_mfma_f64_16x16x4f64:
 * row = 4r + g. */
template <typename T> __device__ __forceinline__ int rho(int g, int r);
template <> __device__ __forceinline__ int rho<float>(int g, int r) { return 4 * g + r; }
template <> __device__ __forceinline__ int rho<double>(int g, int r) { return 4 * r + g; }

/* float -> bf16 bits, round to nearest even (finite inputs). */
__device__ __forceinline__ unsigned short f2bf(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) {
  return __uint_as_float((unsigned)h << 16);
}

enum { MODE_NATIVE = 0, MODE_BF16X3 = 1 };

/* LDS row stride of the panel: odd for conflict-free column access; 16-byte
 * aligned rows for the bf16 operand loads. */
constexpr int panel_stride(int NB, int MODE) { return MODE == MODE_BF16X3 ? NB + 4 : NB + 1; }

struct LuArgs {
  int m, M;
  long n_draws;
  const double *theta;
  long P;
  int n_gam;
  const int *g_row, *g_col, *g_idx, *g_rowptr;
  void *work;
  double *logdet;
  const double *rhs;
  double *y;
  int refine;
  int *flag_count; /* fast: appends flagged draws; pivot: number to process */
  int *flag_list;
  /* sparse prefix (fast kernel) */
  int e_n1, e_ns, e_waves, e_nsg, e_nblk, e_tab_words;
  int e_o_piv, e_o_mptr, e_o_mult, e_o_tptr, e_o_tslot, e_o_cptr, e_o_contrib, e_o_brounds,
      e_o_bcbase;
  unsigned e_zero_pair;
  int e_pre_base, e_slot_bytes, e_tab_persist;
  int e_flag_off; /* LDS byte offset of the early-exit flag (fast kernel) */
  /* solves with an aliased prefix region: per-workgroup global copy of the
   * prefix slots (nullptr: slots stay in LDS) */
  void *e_slot_save;
  long e_slot_stride;
  const int *e_slot_init, *e_tab, *e_sg_off, *e_sg_idx, *e_st_off;
  const unsigned *e_st_pairs;
  /* partial-pivoting kernel, log-determinants: pv_schur_M > 0 = re-eliminate
   * the sparse prefix (LDS at e_pre_base, bound kPrefixLooseMultiplier) and
   * factor only the pv_schur_M x pv_schur_M Schur complement with row
   * interchanges; pv_flag_off = LDS byte offset of the prefix flag, outside
   * the prefix region */
  int pv_schur_M, pv_flag_off;
  /* forecast solves through the prefix (fast kernel, SOLVE) */
  int e_n2, e_xs_base;
  const int *e_wave_q, *e_node_of, *e_fwd_ptr, *e_bwd_ptr;
  const unsigned *e_fwd, *e_bwd;
};

template <typename T>
__device__ void assemble(T *A, int M, const double *th, const LuArgs &a, int tid,
                         int nthreads) {
  const size_t nvec = (size_t)M * M * sizeof(T) / 16;
  float4 *A4 = reinterpret_cast<float4 *>(A);
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  for (size_t q = tid; q < nvec; q += nthreads)
    A4[q] = z;
  __syncthreads();
  for (int r = tid; r < M; r += nthreads)
    A[(size_t)r * M + r] = (T)1;
  for (int e = tid; e < a.n_gam; e += nthreads)
    A[(size_t)a.g_row[e] * M + a.g_col[e]] = (T)(-th[a.g_idx[e]]);
  __syncthreads();
}

/* Sparse prefix + Schur-complement assembly of one draw (fast kernel).
 * LDS at a.e_pre_base: slots sl[] (the entries the prefix elimination
 * touches, plus constant and base slots), then the prefix tables tb[]
 * (copied from a.e_tab on every draw unless they persist).  Per wave the
 * multipliers l_ik = a_ik / a_kk replace a_ik, then every target slot
 * subtracts its gathered sum of l_ik * a_kj (fixed order).  The dense
 * workspace S (M x M, identity padding) receives the separator block of
 * I - Gamma; every entry the prefix updates is then overwritten with
 * -sum x * y over its slot pairs, read in ELL blocks of 64 (one per lane),
 * several blocks in flight per wavefront.  Accumulates the prefix pivots'
 * log|a_kk| (FP64), the zero-pivot / multiplier-bound (lbound) flag and
 * max |l|. */
template <typename T>
__device__ void prefix_assemble(T *A, int M, char *smem, const double *th, const LuArgs &a,
                                int tid, int nthreads, bool load_tab, int &bad, T &amax,
                                double &ls, int *eflag, T lbound) {
  constexpr unsigned LO = 0xffffu;
  const int NWV = nthreads >> 6; /* wavefronts of the workgroup */
  constexpr int UB = 4;   /* Schur blocks in flight per wavefront */
  constexpr int RR = 8;   /* rounds loaded ahead per block */
  constexpr int GB = 8;   /* global reads in flight per thread */
  const int lane = tid & 63, wave = tid >> 6;
  T *sl = reinterpret_cast<T *>(smem + a.e_pre_base);
  int *tb = reinterpret_cast<int *>(smem + a.e_pre_base + a.e_slot_bytes);
  const PrefixTabs pt{a.e_ns, a.e_waves, a.e_n1, a.e_tab_words, a.e_o_piv, a.e_o_mptr,
                      a.e_o_mult, a.e_o_tptr, a.e_o_tslot, a.e_o_cptr, a.e_o_contrib,
                      a.e_slot_init, a.e_tab};
  /* zero fill of the workspace, ordered before the writes below by the
   * prefix elimination's barriers */
  const size_t nvec = (size_t)M * M * sizeof(T) / 16;
  float4 *A4 = reinterpret_cast<float4 *>(A);
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  for (size_t q = tid; q < nvec; q += nthreads)
    A4[q] = z;
  prefix_eliminate<T>(sl, tb, th, pt, tid, nthreads, load_tab, bad, amax, ls, eflag, lbound);
  for (int r = tid; r < M; r += nthreads)
    A[(size_t)r * M + r] = (T)1;
  for (int e0 = tid; e0 < a.e_nsg; e0 += GB * nthreads) {
    int off[GB], ix[GB];
#pragma unroll
    for (int u = 0; u < GB; ++u) {
      const int e = e0 + u * nthreads;
      off[u] = e < a.e_nsg ? a.e_sg_off[e] : -1;
      ix[u] = e < a.e_nsg ? a.e_sg_idx[e] : 0;
    }
    T x[GB];
#pragma unroll
    for (int u = 0; u < GB; ++u)
      x[u] = off[u] >= 0 ? (T)(-th[ix[u]]) : (T)0;
#pragma unroll
    for (int u = 0; u < GB; ++u)
      if (off[u] >= 0)
        A[off[u]] = x[u];
  }
  __syncthreads();
  const int *brounds = tb + a.e_o_brounds, *bcbase = tb + a.e_o_bcbase;
  const unsigned *pairs = a.e_st_pairs;
  for (int b0 = wave; b0 < a.e_nblk; b0 += NWV * UB) {
    int off[UB], R[UB], cb[UB];
    unsigned pv[UB][RR];
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      const int b = b0 + NWV * u;
      const bool ok = b < a.e_nblk;
      R[u] = ok ? brounds[b] : 0;
      cb[u] = ok ? bcbase[b] : 0;
      off[u] = ok ? a.e_st_off[(size_t)b * 64 + lane] : -1;
#pragma unroll
      for (int r = 0; r < RR; ++r)
        pv[u][r] = r < R[u] ? pairs[(size_t)cb[u] + 64 * r + lane] : a.e_zero_pair;
    }
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      T acc = (T)0;
#pragma unroll
      for (int r = 0; r < RR; ++r)
        acc = fma(-sl[pv[u][r] & LO], sl[pv[u][r] >> 16], acc);
      /* longer targets: further rounds RR at a time (one round trip each) */
      for (int rb = RR; rb < R[u]; rb += RR) {
        unsigned c[RR];
#pragma unroll
        for (int r = 0; r < RR; ++r)
          c[r] = rb + r < R[u] ? pairs[(size_t)cb[u] + 64 * (rb + r) + lane] : a.e_zero_pair;
#pragma unroll
        for (int r = 0; r < RR; ++r)
          acc = fma(-sl[c[r] & LO], sl[c[r] >> 16], acc);
      }
      if (off[u] >= 0)
        A[off[u]] = acc;
    }
  }
  __syncthreads();
}

/* Rows of the workspace one block step of dense_substitute reads: the
 * block column's 16-entry segments of the thread's rows (RS rows per
 * thread, M <= RS NT) and, for lanes 0..15 of wavefront 0, the row of
 * the diagonal block.  Step st >= M/16: zeros. */
template <bool FWD, int RS, typename T, int NT>
__device__ __forceinline__ void substitute_fetch(const T *A, int M, int st, int tid,
                                                 T (&dst)[RS + 1][16]) {
  typedef typename Vec16<T>::type VV;
  constexpr int VW = Vec16<T>::n;
  const int NBK = M / 16;
  const int lane = tid & 63, wave = tid >> 6;
  const int kb = FWD ? st : NBK - 1 - st, b0 = 16 * kb;
  const bool live = st < NBK;
#pragma unroll
  for (int rr = 0; rr <= RS; ++rr) {
    int i;
    bool ok;
    if (rr < RS) {
      i = FWD ? b0 + 16 + tid + rr * NT : tid + rr * NT;
      ok = live && (FWD ? i < M : i < b0);
    } else {
      i = b0 + lane;
      ok = live && wave == 0 && lane < 16;
    }
    const VV *src = reinterpret_cast<const VV *>(A + (ok ? (size_t)i * M + b0 : 0));
#pragma unroll
    for (int c4 = 0; c4 < 16 / VW; ++c4) {
      VV x = {};
      if (ok)
        x = src[c4];
#pragma unroll
      for (int e = 0; e < VW; ++e)
        dst[rr][c4 * VW + e] = x[e];
    }
  }
}

/* One block step st of dense_substitute with its fetched rows buf; then
 * refills buf with those of step st + DP. */
template <bool FWD, typename T, int NT, int RS, int DP>
__device__ __forceinline__ void substitute_step(const T *A, int M, double *xd, int tid, int st,
                                                T (&buf)[RS + 1][16]) {
  const int NBK = M / 16;
  const int lane = tid & 63, wave = tid >> 6;
  const int b0 = 16 * (FWD ? st : NBK - 1 - st);
  if (wave == 0 && lane < 16) {
    /* the diagonal block: 16-step substitution across lanes 0..15 */
    double x = xd[b0 + lane];
    if (FWD) {
#pragma unroll
      for (int t = 0; t < 15; ++t) {
        const double xt = __shfl(x, t, 16);
        if (lane > t)
          x -= (double)buf[RS][t] * xt;
      }
    } else {
      double dgv = 0.0;
#pragma unroll
      for (int t = 0; t < 16; ++t)
        dgv = (t == lane) ? (double)buf[RS][t] : dgv;
#pragma unroll
      for (int t = 15; t >= 0; --t) {
        if (lane == t)
          x = x / dgv;
        const double xt = __shfl(x, t, 16);
        if (lane < t)
          x -= (double)buf[RS][t] * xt;
      }
    }
    xd[b0 + lane] = x;
  }
  __syncthreads();
#pragma unroll
  for (int rr = 0; rr < RS; ++rr) {
    const int i = FWD ? b0 + 16 + tid + rr * NT : tid + rr * NT;
    if (FWD ? i < M : i < b0) {
      double acc = xd[i];
#pragma unroll
      for (int t = 0; t < 16; ++t)
        acc -= (double)buf[rr][t] * xd[b0 + t];
      xd[i] = acc;
    }
  }
  substitute_fetch<FWD, RS, T, NT>(A, M, st + DP, tid, buf);
  __syncthreads();
}

/* Dense forward (FWD: unit lower L22) or backward (U22) substitution on
 * xd (LDS, FP64), in block steps of 16 (block kb = st forward,
 * M/16 - 1 - st backward).  The rows a step reads are fetched at the end
 * of the previous step, so their latency overlaps its barrier and the
 * diagonal block's substitution. */
template <bool FWD, typename T, int NT, int RS>
__device__ void dense_substitute(const T *A, int M, double *xd, int tid) {
  const int NBK = M / 16;
  T buf[RS + 1][16];
  substitute_fetch<FWD, RS, T, NT>(A, M, 0, tid, buf);
  for (int st = 0; st < NBK; ++st)
    substitute_step<FWD, T, NT, RS, 1>(A, M, xd, tid, st, buf);
}

/* y = (I - Gamma)^{-1} rhs of draw d from the factors the fast kernel
 * leaves: prefix L and U in the LDS slots, dense L22 / U22 of the Schur
 * complement in the workspace A (M x M).  x (n1 + M doubles of LDS) is
 * indexed by prefix position then dense position.  FP64 accumulation;
 * with a.refine one more pass on the FP64 residual rhs - (I - Gamma) y,
 * formed with the sparse Gamma in series order. */
template <typename T, int NT>
__device__ void prefix_dense_solve(const T *A, int M, char *smem, const double *th,
                                   const LuArgs &a, long d, int tid) {
  constexpr unsigned LO = 0xffffu;
  const int lane = tid & 63, wave = tid >> 6;
  const T *sl = reinterpret_cast<const T *>(smem + a.e_pre_base);
  const int *tb = reinterpret_cast<const int *>(smem + a.e_pre_base + a.e_slot_bytes);
  const int *piv = tb + a.e_o_piv;
  double *xs = reinterpret_cast<double *>(smem + a.e_xs_base);
  double *xd = xs + a.e_n1;
  const int n1 = a.e_n1, n2 = a.e_n2;
  const double *rhs = a.rhs + (size_t)d * a.m;
  double *y = a.y + (size_t)d * a.m;
  for (int pass = 0; pass < (a.refine ? 2 : 1); ++pass) {
    auto bval = [&](int node) -> double {
      if (pass == 0)
        return rhs[node];
      double acc = rhs[node] - y[node];
      for (int e = a.g_rowptr[node]; e < a.g_rowptr[node + 1]; ++e)
        acc += th[a.g_idx[e]] * y[a.g_col[e]];
      return acc;
    };
    /* forward through the prefix, wave by wave */
    for (int w = 0; w < a.e_waves; ++w) {
      for (int q = a.e_wave_q[w] + tid; q < a.e_wave_q[w + 1]; q += NT) {
        double acc = bval(a.e_node_of[q]);
        for (int e = a.e_fwd_ptr[q]; e < a.e_fwd_ptr[q + 1]; ++e) {
          const unsigned c = a.e_fwd[e];
          acc -= (double)sl[c & LO] * xs[c >> 16];
        }
        xs[q] = acc;
      }
      __syncthreads();
    }
    /* right-hand side of the dense part (identity padding: 0) */
    for (int s2 = tid; s2 < M; s2 += NT) {
      double acc = 0.0;
      if (s2 < n2) {
        acc = bval(a.e_node_of[n1 + s2]);
        for (int e = a.e_fwd_ptr[n1 + s2]; e < a.e_fwd_ptr[n1 + s2 + 1]; ++e) {
          const unsigned c = a.e_fwd[e];
          acc -= (double)sl[c & LO] * xs[c >> 16];
        }
      }
      xd[s2] = acc;
    }
    __syncthreads();
    /* dense: L22 (unit lower) then U22 */
    if constexpr (sizeof(T) == 8) {
      /* FP64 factors: plain block steps (the prefetch buffers would cost
       * the FP64 kernel registers in its factorization phases) */
      for (int kb = 0; kb < M / 16; ++kb) {
        const int b0 = 16 * kb;
        if (wave == 0 && lane < 16) {
          double x = xd[b0 + lane];
          const size_t rowoff = (size_t)(b0 + lane) * M;
          for (int t = 0; t < 15; ++t) {
            const double xt = __shfl(x, t, 16);
            if (lane > t)
              x -= (double)A[rowoff + b0 + t] * xt;
          }
          xd[b0 + lane] = x;
        }
        __syncthreads();
        for (int i = b0 + 16 + tid; i < M; i += NT) {
          const T *Lr = A + (size_t)i * M + b0;
          double acc = xd[i];
#pragma unroll
          for (int t = 0; t < 16; ++t)
            acc -= (double)Lr[t] * xd[b0 + t];
          xd[i] = acc;
        }
        __syncthreads();
      }
      for (int kb = M / 16 - 1; kb >= 0; --kb) {
        const int b0 = 16 * kb;
        if (wave == 0 && lane < 16) {
          const size_t rowoff = (size_t)(b0 + lane) * M;
          double x = xd[b0 + lane];
          const double dgv = (double)A[rowoff + b0 + lane];
          for (int t = 15; t >= 0; --t) {
            if (lane == t)
              x = x / dgv;
            const double xt = __shfl(x, t, 16);
            if (lane < t)
              x -= (double)A[rowoff + b0 + t] * xt;
          }
          xd[b0 + lane] = x;
        }
        __syncthreads();
        for (int i = tid; i < b0; i += NT) {
          const T *Ur = A + (size_t)i * M + b0;
          double acc = xd[i];
#pragma unroll
          for (int t = 0; t < 16; ++t)
            acc -= (double)Ur[t] * xd[b0 + t];
          xd[i] = acc;
        }
        __syncthreads();
      }
    } else if (M <= NT) {
      dense_substitute<true, T, NT, 1>(A, M, xd, tid);
      dense_substitute<false, T, NT, 1>(A, M, xd, tid);
    } else {
      dense_substitute<true, T, NT, 2>(A, M, xd, tid);
      dense_substitute<false, T, NT, 2>(A, M, xd, tid);
    }
    /* backward through the prefix, waves in reverse order */
    for (int w = a.e_waves - 1; w >= 0; --w) {
      for (int q = a.e_wave_q[w] + tid; q < a.e_wave_q[w + 1]; q += NT) {
        double acc = xs[q];
        for (int e = a.e_bwd_ptr[q]; e < a.e_bwd_ptr[q + 1]; ++e) {
          const unsigned c = a.e_bwd[e];
          acc -= (double)sl[c & LO] * xs[c >> 16];
        }
        xs[q] = acc / (double)sl[piv[q]];
      }
      __syncthreads();
    }
    for (int i = tid; i < n1 + n2; i += NT) {
      const int node = a.e_node_of[i];
      y[node] = (pass == 0) ? xs[i] : y[node] + xs[i];
    }
    __syncthreads();
  }
}

/* f(std::integral_constant<int, i>) for i = 0 .. N - 1, each a separate
 * inlined body (constant register indices however large the loop). */
template <typename F, int... I>
__device__ __forceinline__ void static_for_seq(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F> __device__ __forceinline__ void static_for(F &&f) {
  static_for_seq(f, std::make_integer_sequence<int, N>{});
}

/* Value of v in lane l, for every lane (an SGPR). */
template <typename T> __device__ __forceinline__ T bcast_lane(T v, int l) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
  } else {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
  }
}

/* Partial-pivoting LU of the tall panel Pn (nr <= 64 RS rows, jb columns,
 * LDS row stride PS) by one wavefront, no workgroup barrier: row lane + 64 s
 * in registers.  Rows are not exchanged while factoring: the pivot of
 * column j is the largest |a_rj| over the rows not chosen yet (the
 * candidates after LAPACK's exchanges; lowest row on ties), broadcast by
 * readlane, then eliminated from the remaining rows with the workgroup
 * loop's element operations (l = a_rj / u_jj, a_rc = fma(-l, u_jc, a_rc)).
 * At the end pivot row j goes to position j, the rows it displaced from
 * positions below jb to the positions the pivots left, and rowmap[K + .]
 * is permuted the same way.  perm: jb ints of LDS scratch. */
template <typename T, int NB, int PS, int RS>
__device__ void panel_gepp_wave(T *Pn, int *rowmap, int *perm, int K, int nr, int jb, int lane,
                                double &logdet, int &singular) {
  T v[RS][NB];
  bool live[RS];
#pragma unroll
  for (int s = 0; s < RS; ++s) {
    const int r = lane + 64 * s;
    live[s] = r < nr;
#pragma unroll
    for (int c = 0; c < NB; ++c)
      v[s][c] = live[s] ? Pn[r * PS + c] : (T)0;
  }
  static_for<NB>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j < jb) {
      /* |a| ordered with NaN above "no row" and below every number */
      T bv = (T)-2;
      int bi = 0x7fffffff;
#pragma unroll
      for (int s = 0; s < RS; ++s) {
        T x = fabs(v[s][j]);
        if (x != x)
          x = (T)-1;
        if (live[s] && x > bv) {
          bv = x;
          bi = lane + 64 * s;
        }
      }
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        const T ov = __shfl_xor(bv, o);
        const int oi = __shfl_xor(bi, o);
        if (ov > bv || (ov == bv && oi < bi)) {
          bv = ov;
          bi = oi;
        }
      }
      const int p = __builtin_amdgcn_readfirstlane(bi);
      const int lp = p & 63, sp = p >> 6;
      T u[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c)
        u[c] = (T)0;
#pragma unroll
      for (int s = 0; s < RS; ++s)
        if (sp == s) {
#pragma unroll
          for (int c = j; c < NB; ++c)
            u[c] = bcast_lane(v[s][c], lp);
        }
#pragma unroll
      for (int s = 0; s < RS; ++s)
        if (s == sp && lane == lp)
          live[s] = false;
      T dj = u[j];
      if (!(fabs(dj) > (T)0)) {
        singular = 1;
        dj = (T)1;
      } else if (lane == 0) {
        logdet += log(fabs((double)dj));
      }
      if (lane == 0)
        perm[j] = p;
#pragma unroll
      for (int s = 0; s < RS; ++s) {
        if (live[s]) {
          const T l = v[s][j] / dj;
          v[s][j] = l;
#pragma unroll
          for (int c = j + 1; c < NB; ++c)
            if (c < jb)
              v[s][c] = fma(-l, u[c], v[s][c]);
        }
      }
    }
  });
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  /* rows below jb chosen as pivots (jb <= 32) */
  unsigned pm = 0;
  for (int j = 0; j < jb; ++j) {
    const int pj = perm[j];
    if (pj < jb)
      pm |= 1u << pj;
  }
  int np[RS], old[RS];
#pragma unroll
  for (int s = 0; s < RS; ++s) {
    const int r = lane + 64 * s;
    np[s] = r;
    old[s] = 0;
    if (r < nr) {
      old[s] = rowmap[K + r];
      if (!live[s]) {
        for (int j = 0; j < jb; ++j)
          if (perm[j] == r)
            np[s] = j;
      } else if (r < jb) {
        /* the t-th displaced row takes the t-th position a pivot left */
        const int t = __popc(~pm & ((1u << r) - 1u));
        int cnt = 0;
        for (int j = 0; j < jb; ++j) {
          const int pj = perm[j];
          if (pj >= jb) {
            if (cnt == t)
              np[s] = pj;
            ++cnt;
          }
        }
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int s = 0; s < RS; ++s) {
    if (lane + 64 * s < nr) {
#pragma unroll
      for (int c = 0; c < NB; ++c)
        Pn[np[s] * PS + c] = v[s][c];
      rowmap[K + np[s]] = old[s];
    }
  }
}

/* Inverses of the nq 16x16 unit-lower diagonal blocks of the panel rows
 * 0..16nq (LDS, stride PS) into Linv[b][i][c]; wave b, lanes 0..15. */
template <typename T, int PS>
__device__ void diag_block_inverses(const T *Pn, T *Linv, int nq, int wave,
                                    int nwaves, int lane) {
  for (int b = wave; b < nq; b += nwaves) {
    if (lane < 16) {
      const int c = lane;
      T x[16];
#pragma unroll
      for (int i = 0; i < 16; ++i)
        x[i] = (i == c) ? (T)1 : (T)0;
#pragma unroll
      for (int i = 1; i < 16; ++i) {
        if (i > c) {
          T acc = (T)0;
#pragma unroll
          for (int t = 0; t < i; ++t)
            acc = fma(Pn[(16 * b + i) * PS + 16 * b + t], x[t], acc);
          x[i] = -acc;
        }
      }
#pragma unroll
      for (int i = 0; i < 16; ++i)
        Linv[b * 256 + i * 16 + c] = x[i];
    }
  }
}

/* ========================================================================
 * Fast path: no row interchanges, verified afterwards.
 * ====================================================================== */
template <typename T, int NB, int MODE, int RPT, int NT, int NBI, bool SOLVE, bool GRP>
__global__ __launch_bounds__(NT, 2) void k_lu_fast(LuArgs a0, const LuArgs *__restrict__ tab,
                                                   GrpMap gm) {
  static_assert(NB % NBI == 0 && (MODE == MODE_NATIVE || NBI == NB), "sub-panel width");
  constexpr int NS = NB / NBI; /* register sub-panels per panel */
  typedef typename Vec4<T>::type V4;
  constexpr int NW = NT / 64;
  constexpr int PS = panel_stride(NB, MODE);
  constexpr int NQ = NB / 16;
  constexpr int NKB = (NQ + 1) / 2; /* bf16 k-blocks of 32 */
  constexpr int BATCH = sizeof(T) == 4 ? 4 : 2; /* trailing items per prefetch */
  typedef typename Vec16<T>::type VV;
  constexpr int VW = Vec16<T>::n;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  int bid, nbk;
  const LuArgs a = grp_args<GRP>(a0, tab, gm, bid, nbk);

  const int M = a.M;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int l16 = lane & 15;
  const int g = lane >> 4;

  T *Pn = reinterpret_cast<T *>(smem);
  T *Linv = Pn + (size_t)M * PS;
  T *prow = Linv + NB * 16; /* [2][NB] pivot-row broadcast, double buffered */
  T *dg = prow + 2 * NB;    /* [M] pivots u_kk */
  double *redd = reinterpret_cast<double *>(dg + M);
  float *redf = reinterpret_cast<float *>(redd + NW);
  auto row_of = [](int i) -> int { return i; };

  T *A = reinterpret_cast<T *>(a.work) + (size_t)bid * M * M;

  for (long d = bid; d < a.n_draws; d += nbk) {
    int bad = 0;
    T amax = (T)0;
    double ls = 0.0;
    /* early exit: any thread that sees |l| > kMaxMultiplier or a zero
     * pivot sets the LDS flag; it is read right after barriers, so every
     * thread takes the same branch, and a flagged draw skips the rest of
     * its factorization (the partial-pivoting kernel redoes it) */
    int *eflag = reinterpret_cast<int *>(smem + a.e_flag_off);
    if (tid == 0)
      *eflag = 0;
    prefix_assemble<T>(A, M, smem, a.theta + (size_t)d * a.P, a, tid, NT,
                       !a.e_tab_persist || d == (long)bid, bad, amax, ls, eflag,
                       (T)kMaxMultiplier);
    bool stop = *eflag != 0;
    if constexpr (SOLVE) {
      if (a.e_slot_save) {
        /* the dense LU reuses the prefix slots' LDS: keep them for the
         * substitutions */
        T *save = reinterpret_cast<T *>(reinterpret_cast<char *>(a.e_slot_save) +
                                        (size_t)bid * a.e_slot_stride);
        const T *sl = reinterpret_cast<const T *>(smem + a.e_pre_base);
        for (int q = tid; q < a.e_ns; q += NT)
          save[q] = sl[q];
        __syncthreads();
      }
    }

    /* ---- phase A: U12 = L11^{-1} A12 for the column tiles
     * [c_first, c_first + 16 ntc), L11 = the nqp 16-row blocks of the panel
     * (Pn rows 0..16 nqp, Linv), written in place of A12 ---- */
    /* Two column tiles per wavefront at a time (h = 0, 1): their block
     * forward substitutions are independent MFMA chains that share the A
     * operands, so the second hides the first's MFMA latency. */
    auto phase_a = [&](int Kp, int nqp, int c_first, int ntc) {
      for (int ct = wave; ct < ntc; ct += 2 * NW) {
        const bool two = ct + NW < ntc;
        V4 u[2][NQ];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int c0 = c_first + 16 * (ct + h * NW);
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            V4 acc = {0, 0, 0, 0};
            if (q < nqp && (h == 0 || two)) {
#pragma unroll
              for (int r = 0; r < 4; ++r)
                acc[r] = A[(size_t)row_of(Kp + 16 * q + rho<T>(g, r)) * M + c0 + l16];
            }
            u[h][q] = acc;
          }
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          if (q < nqp) {
            V4 acc0 = u[0][q], acc1 = u[1][q];
#pragma unroll
            for (int aa = 0; aa < NQ; ++aa)
              if (aa < q)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                  const T av = -Pn[(16 * q + l16) * PS + 16 * aa + rho<T>(g, r)];
                  acc0 = mfma16(av, u[0][aa][r], acc0);
                  acc1 = mfma16(av, u[1][aa][r], acc1);
                }
            V4 uq0 = {0, 0, 0, 0}, uq1 = {0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const T lv = Linv[q * 256 + l16 * 16 + rho<T>(g, r)];
              uq0 = mfma16(lv, acc0[r], uq0);
              uq1 = mfma16(lv, acc1[r], uq1);
            }
            u[0][q] = uq0;
            u[1][q] = uq1;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              if (h == 0 || two) {
                const int c0 = c_first + 16 * (ct + h * NW);
#pragma unroll
                for (int r = 0; r < 4; ++r)
                  A[(size_t)row_of(Kp + 16 * q + rho<T>(g, r)) * M + c0 + l16] = u[h][q][r];
              }
            }
          }
        }
      }
    };

    /* ---- phase B: A22 -= L21 U12 over the row tiles starting at
     * Kp + jbp (ntr of them) and the column tiles [c_first, + 16 ntc);
     * L21 = Pn rows jbp.., K-width 16 nqp.  Items are (pair of column tiles
     * = 32 columns, row tile): full 128-byte rows for FP32, shared A
     * operands; 8 contiguous item ranges; loads one batch ahead. ---- */
    auto phase_b = [&](int Kp, int jbp, int nqp, int c_first, int ntc, int ntr, bool bf16) {
      const int ntp = (ntc + 1) >> 1;
      const long tot = (long)ntp * ntr;
      const long lo_it = tot * wave / NW, hi_it = tot * (wave + 1) / NW;
      int cur_cp = -1;
      V4 ub[2][NQ];
      u16x8 bhi[2][NKB], blo[2][NKB];
      V4 cc[BATCH][2], cn[BATCH][2];
      auto load_batch = [&](long it0, V4 (*dst)[2]) {
#pragma unroll
        for (int b = 0; b < BATCH; ++b) {
          const long it = it0 + b;
          if (it < hi_it) {
            const int cp = (int)(it / ntr);
            const int rt = (int)(it - (long)cp * ntr);
            const int col = c_first + 32 * cp + l16;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const T *src = A + (size_t)row_of(Kp + jbp + 16 * rt + rho<T>(g, r)) * M + col;
#pragma unroll
              for (int h = 0; h < 2; ++h)
                if (2 * cp + h < ntc)
                  dst[b][h][r] = src[16 * h];
            }
          }
        }
      };
      if (lo_it < hi_it)
        load_batch(lo_it, cc);
      for (long it0 = lo_it; it0 < hi_it; it0 += BATCH) {
        if (it0 + BATCH < hi_it)
          load_batch(it0 + BATCH, cn);
#pragma unroll
        for (int b = 0; b < BATCH; ++b) {
          const long it = it0 + b;
          if (it < hi_it) {
            const int cp = (int)(it / ntr);
            const int rt = (int)(it - (long)cp * ntr);
            const bool two = 2 * cp + 1 < ntc;
            if (cp != cur_cp) {
              cur_cp = cp;
#pragma unroll
              for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                  V4 t = {0, 0, 0, 0};
                  if (q < nqp && (h == 0 || two)) {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                      t[r] = A[(size_t)row_of(Kp + 16 * q + rho<T>(g, r)) * M + c_first + 32 * cp + 16 * h + l16];
                  }
                  ub[h][q] = t;
                }
              if constexpr (MODE == MODE_BF16X3) {
                if (bf16) {
#pragma unroll
                  for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
                      for (int e = 0; e < 8; ++e) {
                        const int q = 2 * kb + (e >> 2);
                        const float x = (q < NQ) ? (float)ub[h][q < NQ ? q : 0][e & 3] : 0.f;
                        const unsigned short hh = f2bf(x);
                        bhi[h][kb][e] = hh;
                        blo[h][kb][e] = f2bf(x - bf2f(hh));
                      }
                }
              }
            }
            V4 c0v = cc[b][0], c1v = cc[b][1];
            const int prw = jbp + 16 * rt + l16;
            bool done = false;
            if constexpr (MODE == MODE_BF16X3) {
              if (bf16) {
                const char *rowp = reinterpret_cast<const char *>(Pn + prw * PS);
#pragma unroll
                for (int kb = 0; kb < NKB; ++kb) {
                  if (2 * kb < nqp) {
                    const u16x8 ahi = *reinterpret_cast<const u16x8 *>(rowp + (kb * 4 + g) * 32);
                    const u16x8 alo = *reinterpret_cast<const u16x8 *>(rowp + (kb * 4 + g) * 32 + 16);
                    c0v = mfma_bf16(alo, bhi[0][kb], c0v);
                    c0v = mfma_bf16(ahi, blo[0][kb], c0v);
                    c0v = mfma_bf16(ahi, bhi[0][kb], c0v);
                    if (two) {
                      c1v = mfma_bf16(alo, bhi[1][kb], c1v);
                      c1v = mfma_bf16(ahi, blo[1][kb], c1v);
                      c1v = mfma_bf16(ahi, bhi[1][kb], c1v);
                    }
                  }
                }
                done = true;
              }
            }
            if (!done) {
#pragma unroll
              for (int aa = 0; aa < NQ; ++aa)
                if (aa < nqp)
#pragma unroll
                  for (int r = 0; r < 4; ++r) {
                    const T av = -Pn[prw * PS + 16 * aa + rho<T>(g, r)];
                    c0v = mfma16(av, ub[0][aa][r], c0v);
                    if (two)
                      c1v = mfma16(av, ub[1][aa][r], c1v);
                  }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              T *dst = A + (size_t)row_of(Kp + jbp + 16 * rt + rho<T>(g, r)) * M + c_first + 32 * cp + l16;
              dst[0] = (T)c0v[r];
              if (two)
                dst[16] = (T)c1v[r];
            }
          }
        }
#pragma unroll
        for (int b = 0; b < BATCH; ++b) {
          cc[b][0] = cn[b][0];
          cc[b][1] = cn[b][1];
        }
      }
    };

    for (int K = 0; K < M && !stop; K += NB) {
      const int jb = min(NB, M - K);
      const int nq = jb >> 4;

      /* ---- the panel in sub-panels of NBI columns: rows in registers,
       * factored without interchanges, one barrier per column; between
       * sub-panels the next sub-panel's columns get the previous one's
       * update (phases A and B restricted to those columns, native MFMA) */
      /* One loop runs every trailing update of this panel through a single
       * call site of phases A and B (the kernel stays small enough for the
       * instruction cache): upd < NS - 1 updates the next sub-panel's
       * columns with the sub-panels factored so far, upd = NS - 1 is the
       * full trailing update. */
      for (int upd = 0; upd < NS; ++upd) {
        const int so = upd * NBI;            /* columns factored so far */
        const bool last = upd == NS - 1 || so + NBI >= jb;
        const int Ks = K + so;               /* first column of this sub-panel */
        const int jbs = min(NBI, K + jb - Ks);
        const int nr = M - Ks;
        {

        /* ---- sub-panel rows into registers (16-byte loads) ---- */
        T v[RPT][NBI];
#pragma unroll
        for (int rr = 0; rr < RPT; ++rr) {
          const int r = tid + NT * rr;
          const VV *src = reinterpret_cast<const VV *>(A + (size_t)row_of(Ks + r < M ? Ks + r : 0) * M + Ks);
#pragma unroll
          for (int c4 = 0; c4 < NBI / VW; ++c4) {
            VV x = {};
            if (r < nr && c4 * VW < jbs)
              x = src[c4];
#pragma unroll
            for (int e = 0; e < VW; ++e)
              v[rr][c4 * VW + e] = x[e];
          }
        }

        /* ---- factor without interchanges: one barrier per column ---- */
#pragma unroll
        for (int j = 0; j < NBI; ++j) {
          if (j < jbs) {
            T *pr = prow + (j & 1) * NB;
            const int c4s = j / VW; /* first 16-byte chunk holding column j */
            if (tid == j) {
#pragma unroll
              for (int c4 = c4s; c4 < NBI / VW; ++c4) {
                VV x;
#pragma unroll
                for (int e = 0; e < VW; ++e)
                  x[e] = v[0][c4 * VW + e];
                reinterpret_cast<VV *>(pr)[c4] = x;
              }
              dg[Ks + j] = v[0][j];
            }
            __syncthreads();
            T dj = pr[j];
            if (!(fabs(dj) > (T)0)) {
              bad = 1;
              dj = (T)1;
            }
            /* FP32: multipliers through the hardware reciprocal (1 ulp),
             * off the IEEE division's longer dependent sequence */
            T rdj = (T)0;
            if constexpr (sizeof(T) == 4)
              rdj = __builtin_amdgcn_rcpf(dj);
#pragma unroll
            for (int rr = 0; rr < RPT; ++rr) {
              const int r = tid + NT * rr;
              if (r > j && r < nr) {
                T l;
                if constexpr (sizeof(T) == 4)
                  l = v[rr][j] * rdj;
                else
                  l = v[rr][j] / dj;
                v[rr][j] = l;
                amax = fmax(amax, fabs(l));
                if (__builtin_isnan(l))
                  bad = 1; /* fmax drops NaN */
#pragma unroll
                for (int c4 = c4s; c4 < NBI / VW; ++c4) {
                  const VV x = reinterpret_cast<const VV *>(pr)[c4];
#pragma unroll
                  for (int e = 0; e < VW; ++e)
                    if (c4 * VW + e > j)
                      v[rr][c4 * VW + e] = fma(-l, x[e], v[rr][c4 * VW + e]);
                }
              }
            }
          }
        }

        /* a multiplier beyond the bound or a zero pivot: flag the draw
         * (read after the barrier below) */
        if (bad || !(amax <= (T)kMaxMultiplier))
          *eflag = 1;

        /* ---- sub-panel to LDS (Pn rows relative to K; bf16 mode: rows
         * >= jb of the whole panel as -L21 hi/lo, single sub-panel only) */
#pragma unroll
        for (int rr = 0; rr < RPT; ++rr) {
          const int r = tid + NT * rr;
          if (r < nr) {
            const int pr_ = so + r;
            if (MODE == MODE_BF16X3 && pr_ >= jb) {
              unsigned short *row = reinterpret_cast<unsigned short *>(Pn + pr_ * PS);
#pragma unroll
              for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
                for (int gg = 0; gg < 4; ++gg)
#pragma unroll
                  for (int e = 0; e < 8; ++e) {
                    const int col = 32 * kb + 16 * (e >> 2) + 4 * gg + (e & 3);
                    const float x = (col < NBI) ? -(float)v[rr][col < NBI ? col : 0] : 0.f;
                    const unsigned short hi = f2bf(x);
                    row[(kb * 4 + gg) * 16 + e] = hi;
                    row[(kb * 4 + gg) * 16 + 8 + e] = f2bf(x - bf2f(hi));
                  }
            } else {
#pragma unroll
              for (int c = 0; c < NBI; ++c)
                Pn[pr_ * PS + so + c] = v[rr][c];
            }
            if constexpr (SOLVE) {
              /* keep L and U of the panel for the triangular solves */
              T *dst = A + (size_t)row_of(Ks + r) * M + Ks;
#pragma unroll
              for (int c = 0; c < NBI; ++c)
                if (c < jbs)
                  dst[c] = v[rr][c];
            }
          }
        }
        __syncthreads();
        }
        if (*eflag) {
          stop = true;
          break;
        }
        /* trailing update: the next sub-panel's columns, or everything right
         * of the panel after its last sub-panel */
        const int kw = last ? jb : so + NBI;          /* factored width */
        diag_block_inverses<T, PS>(Pn, Linv, kw >> 4, wave, NW, lane);
        __syncthreads();
        const int c_first = K + kw;
        const int ntc = last ? (M - K - jb) >> 4 : min(NBI, jb - kw) >> 4;
        const int ntr = (M - c_first) >> 4;
        if (ntc > 0) {
          phase_a(K, kw >> 4, c_first, ntc);
          __syncthreads();
          phase_b(K, kw, kw >> 4, c_first, ntc, ntr, last && MODE == MODE_BF16X3);
        }
        __syncthreads();
        if (last)
          break;
      }
    }

    /* ---- log|det| = sum log|u_kk| (FP64, fixed reduction order) and the
     * check: any |l| > 1 or zero pivot -> partial-pivoting fallback ---- */
    for (int i = tid; i < M; i += NT)
      ls += log(fabs((double)dg[i]));
    /* a zero pivot seen by any thread flags the draw through the max */
    float wm = bad ? INFINITY : (float)amax;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      wm = fmaxf(wm, __shfl_xor(wm, o));
      ls += __shfl_xor(ls, o);
    }
    if (lane == 0) {
      redf[wave] = wm;
      redd[wave] = ls;
    }
    __syncthreads();
    float mx = redf[0];
    for (int w = 1; w < NW; ++w)
      mx = fmaxf(mx, redf[w]);
    const bool flagged = !(mx <= kMaxMultiplier);
    if (tid == 0) {
      double logdet = redd[0];
      for (int w = 1; w < NW; ++w)
        logdet += redd[w];
      if (flagged) {
        const int slot = atomicAdd(a.flag_count, 1);
        a.flag_list[slot] = (int)d;
      }
      if (a.logdet)
        a.logdet[d] = flagged ? NAN : logdet;
    }
    __syncthreads();
    if constexpr (SOLVE) {
      if (!flagged) {
        if (a.e_slot_save) {
          /* prefix tables and slots back into LDS (the dense region is free) */
          int *tb = reinterpret_cast<int *>(smem + a.e_pre_base + a.e_slot_bytes);
          T *sl = reinterpret_cast<T *>(smem + a.e_pre_base);
          const T *save = reinterpret_cast<const T *>(reinterpret_cast<char *>(a.e_slot_save) +
                                                      (size_t)bid * a.e_slot_stride);
          for (int w = tid; w < a.e_tab_words; w += NT)
            tb[w] = a.e_tab[w];
          for (int q = tid; q < a.e_ns; q += NT)
            sl[q] = save[q];
          __syncthreads();
        }
        prefix_dense_solve<T, NT>(A, M, smem, a.theta + (size_t)d * a.P, a, d, tid);
      }
    }
  }
}

/* ========================================================================
 * Partial-pivoting path (fallback for flagged draws, forecast solves).
 * ====================================================================== */
template <typename T, int NB, int MODE, bool SOLVE, bool GRP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_lu_pivot(
    LuArgs a0, const LuArgs *__restrict__ tab,
                                                  GrpMap gm) {
  typedef typename Vec4<T>::type V4;
  constexpr int WG = 256;
  constexpr int NW = 4;
  constexpr int PS = panel_stride(NB, MODE_NATIVE);
  constexpr int NQ = NB / 16;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  int bid, nbk;
  const LuArgs a = grp_args<GRP>(a0, tab, gm, bid, nbk);

  const int Mfull = a.M;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int l16 = lane & 15;
  const int g = lane >> 4;

  /* LDS laid out for the full dimension; the Schur mode uses a prefix of it */
  T *Pn = reinterpret_cast<T *>(smem);
  T *Linv = Pn + (size_t)Mfull * PS;
  int *rowmap = reinterpret_cast<int *>(Linv + NB * 16);
  size_t off = (size_t)Mfull * PS * sizeof(T) + (size_t)NB * 16 * sizeof(T) +
               (size_t)Mfull * sizeof(int);
  off = (off + 15) & ~(size_t)15;
  double *redv = reinterpret_cast<double *>(smem + off);
  int *redi = reinterpret_cast<int *>(redv + 4);
  off += 4 * (sizeof(double) + sizeof(int));
  off = (off + 15) & ~(size_t)15;
  double *sv = reinterpret_cast<double *>(smem + off); /* SOLVE only */

  T *A = reinterpret_cast<T *>(a.work) + (size_t)bid * Mfull * Mfull;
  const long count = a.flag_list ? (long)*a.flag_count : a.n_draws;

  for (long q = bid; q < count; q += nbk) {
    const long d = a.flag_list ? (long)a.flag_list[q] : q;
    const double *th = a.theta + (size_t)d * a.P;
    /* the dimension factored for this draw: the Schur complement after the
     * sparse prefix (log-determinants, prefix within the loose bound) or
     * the whole matrix */
    int M = Mfull;
    double logdet = 0.0;
    bool schur = false;
    if constexpr (!SOLVE) {
      if (a.pv_schur_M > 0) {
        int *pflag = reinterpret_cast<int *>(smem + a.pv_flag_off);
        if (tid == 0)
          *pflag = 0;
        __syncthreads();
        int pbad = 0;
        T pmax = (T)0;
        double ls = 0.0;
        prefix_assemble<T>(A, a.pv_schur_M, smem, th, a, tid, WG, true, pbad, pmax, ls, pflag,
                           (T)kPrefixLooseMultiplier);
        schur = *pflag == 0;
        if (schur) {
          /* prefix log|a_kk| summed over the workgroup in a fixed order */
#pragma unroll
          for (int o = 32; o >= 1; o >>= 1)
            ls += __shfl_xor(ls, o);
          if (lane == 0)
            redv[wave] = ls;
          __syncthreads();
          logdet = (redv[0] + redv[1]) + (redv[2] + redv[3]);
          M = a.pv_schur_M;
        }
        __syncthreads();
      }
    }
    if (!schur)
      assemble<T>(A, M, th, a, tid, WG);
    for (int r = tid; r < M; r += WG)
      rowmap[r] = r;
    __syncthreads();

    int singular = 0;

    for (int K = 0; K < M; K += NB) {
      const int jb = min(NB, M - K);
      const int nr = M - K;
      const int nq = jb >> 4;

      for (int e = tid; e < nr * NB; e += WG) {
        const int r = e / NB;
        const int c = e - r * NB;
        Pn[r * PS + c] = (c < jb) ? A[(size_t)rowmap[K + r] * M + K + c] : (T)0;
      }
      __syncthreads();

      /* FP32 panels up to 32 wide and 320 rows: one wavefront, rows in
       * registers, no workgroup barrier per column */
      if (sizeof(T) == 4 && NB <= 32 && nr <= 320) {
        if (wave == 0) {
          panel_gepp_wave<T, NB, PS, 5>(Pn, rowmap, reinterpret_cast<int *>(Linv), K, nr, jb,
                                        lane, logdet, singular);
          if (lane == 0)
            redi[0] = singular;
        }
        __syncthreads();
        singular = redi[0];
      } else
      for (int j = 0; j < jb; ++j) {
        T bv = (T)-1;
        int bi = 0x7fffffff;
        for (int r = j + tid; r < nr; r += WG) {
          const T v = fabs(Pn[r * PS + j]);
          if (v > bv) {
            bv = v;
            bi = r;
          }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
          const T ov = __shfl_xor(bv, o);
          const int oi = __shfl_xor(bi, o);
          if (ov > bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
          }
        }
        if (lane == 0) {
          redv[wave] = (double)bv;
          redi[wave] = bi;
        }
        __syncthreads();
        double best = redv[0];
        int piv = redi[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) {
          const double ov = redv[w];
          const int oi = redi[w];
          if (ov > best || (ov == best && oi < piv)) {
            best = ov;
            piv = oi;
          }
        }
        if (piv >= nr)
          piv = j;
        if (piv != j) {
          for (int c = tid; c < jb; c += WG) {
            const T t0 = Pn[j * PS + c];
            Pn[j * PS + c] = Pn[piv * PS + c];
            Pn[piv * PS + c] = t0;
          }
          if (tid == 0) {
            const int t1 = rowmap[K + j];
            rowmap[K + j] = rowmap[K + piv];
            rowmap[K + piv] = t1;
          }
        }
        __syncthreads();
        T dj = Pn[j * PS + j];
        if (!(fabs(dj) > (T)0)) {
          singular = 1;
          dj = (T)1;
        } else if (tid == 0) {
          logdet += log(fabs((double)dj));
        }
        for (int r = j + 1 + tid; r < nr; r += WG) {
          const T l = Pn[r * PS + j] / dj;
          Pn[r * PS + j] = l;
          for (int c = j + 1; c < jb; ++c)
            Pn[r * PS + c] = fma(-l, Pn[j * PS + c], Pn[r * PS + c]);
        }
        __syncthreads();
      }

      diag_block_inverses<T, PS>(Pn, Linv, nq, wave, NW, lane);
      if constexpr (SOLVE) {
        for (int e = tid; e < nr * jb; e += WG) {
          const int r = e / jb;
          const int c = e - r * jb;
          A[(size_t)rowmap[K + r] * M + K + c] = Pn[r * PS + c];
        }
      }
      __syncthreads();

      /* U12 and the trailing update; each column tile is owned by one wave */
      const int ntc = (M - K - jb) >> 4;
      for (int ct = wave; ct < ntc; ct += NW) {
        const int c0 = K + jb + 16 * ct;
        V4 u[NQ];
#pragma unroll
        for (int qq = 0; qq < NQ; ++qq) {
          V4 acc = {0, 0, 0, 0};
          if (qq < nq) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              acc[r] = A[(size_t)rowmap[K + 16 * qq + rho<T>(g, r)] * M + c0 + l16];
#pragma unroll
            for (int aa = 0; aa < NQ; ++aa)
              if (aa < qq)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                  acc = mfma16(-Pn[(16 * qq + l16) * PS + 16 * aa + rho<T>(g, r)], u[aa][r], acc);
            V4 uq = {0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < 4; ++r)
              uq = mfma16(Linv[qq * 256 + l16 * 16 + rho<T>(g, r)], acc[r], uq);
            acc = uq;
            if constexpr (SOLVE) {
#pragma unroll
              for (int r = 0; r < 4; ++r)
                A[(size_t)rowmap[K + 16 * qq + rho<T>(g, r)] * M + c0 + l16] = acc[r];
            }
          }
          u[qq] = acc;
        }
        for (int rt = 0; rt < ntc; ++rt) {
          const int row0 = K + jb + 16 * rt;
          int pr[4];
          V4 c;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            pr[r] = rowmap[row0 + rho<T>(g, r)];
            c[r] = A[(size_t)pr[r] * M + c0 + l16];
          }
          const int prw = jb + 16 * rt + l16;
#pragma unroll
          for (int aa = 0; aa < NQ; ++aa)
            if (aa < nq)
#pragma unroll
              for (int r = 0; r < 4; ++r)
                c = mfma16(-Pn[prw * PS + 16 * aa + rho<T>(g, r)], u[aa][r], c);
#pragma unroll
          for (int r = 0; r < 4; ++r)
            A[(size_t)pr[r] * M + c0 + l16] = c[r];
        }
      }
      __syncthreads();
    }

    if constexpr (!SOLVE) {
      if (tid == 0)
        a.logdet[d] = singular ? -INFINITY : logdet;
      __syncthreads();
      continue;
    }

    /* ---- forecast: solve L U y = P rhs, optional refinement ---- */
    const double *rhs = a.rhs + (size_t)d * a.m;
    double *yout = a.y + (size_t)d * a.m;
    for (int pass = 0; pass < (a.refine ? 2 : 1); ++pass) {
      for (int i = tid; i < M; i += WG) {
        const int src = rowmap[i];
        double b = 0.0;
        if (src < a.m) {
          if (pass == 0) {
            b = rhs[src];
          } else {
            double acc = rhs[src] - yout[src];
            for (int e = a.g_rowptr[src]; e < a.g_rowptr[src + 1]; ++e)
              acc += th[a.g_idx[e]] * yout[a.g_col[e]];
            b = acc;
          }
        }
        sv[i] = b;
      }
      __syncthreads();
      for (int kb = 0; kb < M / 16; ++kb) {
        const int b0 = 16 * kb;
        if (wave == 0 && lane < 16) {
          double x = sv[b0 + lane];
          const size_t rowoff = (size_t)rowmap[b0 + lane] * M;
          for (int t = 0; t < 15; ++t) {
            const double xt = __shfl(x, t, 16);
            if (lane > t)
              x -= (double)A[rowoff + b0 + t] * xt;
          }
          sv[b0 + lane] = x;
        }
        __syncthreads();
        for (int i = b0 + 16 + tid; i < M; i += WG) {
          const T *Lr = A + (size_t)rowmap[i] * M + b0;
          double acc = sv[i];
#pragma unroll
          for (int t = 0; t < 16; ++t)
            acc -= (double)Lr[t] * sv[b0 + t];
          sv[i] = acc;
        }
        __syncthreads();
      }
      for (int kb = M / 16 - 1; kb >= 0; --kb) {
        const int b0 = 16 * kb;
        if (wave == 0 && lane < 16) {
          const size_t rowoff = (size_t)rowmap[b0 + lane] * M;
          double x = sv[b0 + lane];
          const double diag = (double)A[rowoff + b0 + lane];
          for (int t = 15; t >= 0; --t) {
            if (lane == t)
              x = x / diag;
            const double xt = __shfl(x, t, 16);
            if (lane < t)
              x -= (double)A[rowoff + b0 + t] * xt;
          }
          sv[b0 + lane] = x;
        }
        __syncthreads();
        for (int i = tid; i < b0; i += WG) {
          const T *Ur = A + (size_t)rowmap[i] * M + b0;
          double acc = sv[i];
#pragma unroll
          for (int t = 0; t < 16; ++t)
            acc -= (double)Ur[t] * sv[b0 + t];
          sv[i] = acc;
        }
        __syncthreads();
      }
      for (int i = tid; i < a.m; i += WG) {
        const double vv = singular ? NAN : sv[i];
        yout[i] = (pass == 0) ? vv : yout[i] + vv;
      }
      __syncthreads();
    }
  }
}

size_t fast_lds_bytes(int M, int NB, int MODE, size_t ts) {
  size_t b = (size_t)M * panel_stride(NB, MODE) * ts + (size_t)NB * 16 * ts + 2 * (size_t)NB * ts +
             (size_t)M * ts;
  b = (b + 15) & ~(size_t)15;
  /* reductions (8 doubles + 8 floats), then 16 bytes for the early-exit
   * flag at the end */
  return b + 8 * sizeof(double) + 8 * sizeof(float) + 16;
}

size_t pivot_lds_bytes(int M, int NB, size_t ts, bool solve) {
  size_t b = (size_t)M * panel_stride(NB, MODE_NATIVE) * ts + (size_t)NB * 16 * ts + (size_t)M * sizeof(int);
  b = (b + 15) & ~(size_t)15;
  b += 4 * (sizeof(double) + sizeof(int));
  b = (b + 15) & ~(size_t)15;
  if (solve)
    b += (size_t)M * sizeof(double);
  return b;
}

/* One int set on the device: the fallback counter before an LU launch (a
 * kernel rather than a memset node, so that a captured graph replays it
 * like every other launch). */
__global__ void k_set_int(int *p, int v) {
  if (threadIdx.x == 0)
    *p = v;
}

/* *a = 0 (op 0), *a += *b (op 1), *a = *b (op 2): the fallback count over
 * the chunks of a solve, kept by the device. */
__global__ void k_int_op(int *a, const int *b, int op) {
  if (threadIdx.x == 0)
    *a = op == 0 ? 0 : op == 1 ? *a + *b : *b;
}

int int_op(int *a, const int *b, int op, hipStream_t st) {
  k_int_op<<<1, 64, 0, st>>>(a, b, op);
  return hipGetLastError() == hipSuccess ? SGDLM_OK : SGDLM_ERR_DEVICE;
}

int set_int(int *p, int v, hipStream_t st) {
  k_set_int<<<1, 64, 0, st>>>(p, v);
  return hipGetLastError() == hipSuccess ? SGDLM_OK : SGDLM_ERR_DEVICE;
}

/* An LU kernel: its single-model form, its form for a launch over several
 * models (log-determinants only) and its workgroup size. */
struct LuKern {
  const void *one = nullptr, *grp = nullptr;
  int threads = 0;
};

template <typename T, int NB, int MODE, int RPT, int NT, int NBI, bool SOLVE>
LuKern fast_k() {
  LuKern k;
  k.one = (const void *)k_lu_fast<T, NB, MODE, RPT, NT, NBI, SOLVE, false>;
  if constexpr (!SOLVE)
    k.grp = (const void *)k_lu_fast<T, NB, MODE, RPT, NT, NBI, false, true>;
  k.threads = NT;
  return k;
}

template <typename T, int NB, bool SOLVE> LuKern pivot_k() {
  LuKern k;
  k.one = (const void *)k_lu_pivot<T, NB, MODE_NATIVE, SOLVE, false>;
  if constexpr (!SOLVE)
    k.grp = (const void *)k_lu_pivot<T, NB, MODE_NATIVE, false, true>;
  k.threads = 256;
  return k;
}

/* One model's launch of kernel k (its single-model form). */
int launch_one(const LuKern &k, size_t lds, int grid, const LuArgs &a, hipStream_t st) {
  if (!k.one)
    return SGDLM_ERR_ARG;
  if (allow_dynamic_lds(k.one, lds) != hipSuccess)
    return SGDLM_ERR_DEVICE;
  LuArgs a0 = a;
  const LuArgs *tab = nullptr;
  GrpMap gm;
  void *args[] = {&a0, &tab, &gm};
  if (hipLaunchKernel(k.one, dim3(grid), dim3(k.threads), args, lds, st) != hipSuccess)
    return SGDLM_ERR_DEVICE;
  return hipGetLastError() == hipSuccess ? SGDLM_OK : SGDLM_ERR_DEVICE;
}

/* The fast kernel for plan at dense dimension M (empty: none fits). */
template <typename T, int NB, int MODE, bool SOLVE> LuKern fast_sel(const LuPlan &plan, int M) {
  /* register sub-panels of 32 columns for 64-wide native panels */
  constexpr int NBI = (MODE == MODE_NATIVE && NB > 32) ? 32 : NB;
  if (plan.threads == 256) {
    /* two workgroups per CU; rows per thread up to 2 (M <= 512); solves
     * only with 32-wide panels */
    if constexpr (NBI <= 32 && (!SOLVE || NB == 32 || sizeof(T) == 4)) {
      if (M <= 256)
        return fast_k<T, NB, MODE, 1, 256, NBI, SOLVE>();
      if constexpr (!(sizeof(T) == 8 && NB == 16)) {
        if (M <= 512)
          return fast_k<T, NB, MODE, 2, 256, NBI, SOLVE>();
      }
    }
    return LuKern();
  }
  if (plan.threads != 512)
    return LuKern();
  if (M <= 512)
    return fast_k<T, NB, MODE, 1, 512, NBI, SOLVE>();
  /* FP64 with 16-wide panels and two rows per thread is never planned (see
   * lu_plan); not instantiating it also avoids a register-coalescer crash of
   * the ROCm 7.2 compiler on that instantiation. */
  if constexpr (!(sizeof(T) == 8 && NB == 16)) {
    if (M <= 1024)
      return fast_k<T, NB, MODE, 2, 512, NBI, SOLVE>();
  }
  return LuKern();
}

template <bool SOLVE> LuKern select_fast(const LuPlan &plan, int M) {
  switch (plan.prec) {
  case SGDLM_PREC_FP64:
    if (plan.nb == 64)
      return fast_sel<double, 64, MODE_NATIVE, SOLVE>(plan, M);
    if (plan.nb == 32)
      return fast_sel<double, 32, MODE_NATIVE, SOLVE>(plan, M);
    if constexpr (!SOLVE)
      return fast_sel<double, 16, MODE_NATIVE, false>(plan, M);
    return LuKern();
  case SGDLM_PREC_FP32:
    if (plan.nb == 64)
      return fast_sel<float, 64, MODE_NATIVE, SOLVE>(plan, M);
    if (plan.nb == 32)
      return fast_sel<float, 32, MODE_NATIVE, SOLVE>(plan, M);
    if constexpr (!SOLVE)
      return fast_sel<float, 16, MODE_NATIVE, false>(plan, M);
    return LuKern();
  case SGDLM_PREC_BF16X3:
    if constexpr (!SOLVE) {
      if (plan.nb == 64)
        return fast_sel<float, 64, MODE_BF16X3, false>(plan, M);
      return fast_sel<float, 32, MODE_BF16X3, false>(plan, M);
    }
    return LuKern();
  }
  return LuKern();
}

template <bool SOLVE> LuKern select_pivot(const LuPlan &plan) {
  /* BF16x3 falls back to FP32 partial pivoting */
  if (plan.prec == SGDLM_PREC_FP64)
    return plan.pivot_nb == 32 ? pivot_k<double, 32, SOLVE>() : pivot_k<double, 16, SOLVE>();
  if (plan.pivot_nb == 64)
    return pivot_k<float, 64, SOLVE>();
  if (plan.pivot_nb == 32)
    return pivot_k<float, 32, SOLVE>();
  return pivot_k<float, 16, SOLVE>();
}

template <bool SOLVE>
int dispatch_fast(const LuArgs &a, const LuPlan &plan, hipStream_t st) {
  return launch_one(select_fast<SOLVE>(plan, a.M), plan.lds, plan.fast_grid, a, st);
}

template <bool SOLVE>
int dispatch_pivot(const LuArgs &a, const LuPlan &plan, hipStream_t st) {
  return launch_one(select_pivot<SOLVE>(plan), plan.pivot_lds, plan.pivot_grid, a, st);
}

LuArgs base_args(const Structure &S, long N, const double *theta, void *work) {
  LuArgs a{};
  a.m = S.m;
  a.M = S.M;
  a.n_draws = N;
  a.theta = theta;
  a.P = S.P;
  a.n_gam = S.n_gam;
  a.g_row = S.g_row;
  a.g_col = S.g_col;
  a.g_idx = S.g_idx;
  a.g_rowptr = S.g_rowptr;
  a.work = work;
  a.e_n1 = S.e_n1;
  a.e_ns = S.e_ns;
  a.e_waves = S.e_waves;
  a.e_nsg = S.e_nsg;
  a.e_nblk = S.e_nblk;
  a.e_tab_words = S.e_tab_words;
  a.e_o_piv = S.e_o_piv;
  a.e_o_mptr = S.e_o_mptr;
  a.e_o_mult = S.e_o_mult;
  a.e_o_tptr = S.e_o_tptr;
  a.e_o_tslot = S.e_o_tslot;
  a.e_o_cptr = S.e_o_cptr;
  a.e_o_contrib = S.e_o_contrib;
  a.e_o_brounds = S.e_o_brounds;
  a.e_o_bcbase = S.e_o_bcbase;
  a.e_zero_pair = S.e_zero_pair;
  a.e_slot_init = S.e_slot_init;
  a.e_tab = S.e_tab;
  a.e_sg_off = S.e_sg_off;
  a.e_sg_idx = S.e_sg_idx;
  a.e_st_off = S.e_st_off;
  a.e_st_pairs = S.e_st_pairs;
  a.e_n2 = S.e_n2;
  a.e_wave_q = S.e_wave_q;
  a.e_node_of = S.e_node_of;
  a.e_fwd_ptr = S.e_fwd_ptr;
  a.e_fwd = S.e_fwd;
  a.e_bwd_ptr = S.e_bwd_ptr;
  a.e_bwd = S.e_bwd;
  return a;
}

/* The partial-pivoting kernel's view of the arguments (log-determinants):
 * the whole matrix, and in the Schur mode the prefix region at LDS offset 0
 * with its tables loaded on every draw. */
void set_pivot_layout(LuArgs &a, const Structure &S, const LuPlan &plan) {
  a.M = S.M;
  a.pv_schur_M = plan.pivot_schur_M;
  a.pv_flag_off = (int)plan.pivot_flag_off;
  if (plan.pivot_schur_M > 0) {
    a.e_pre_base = 0;
    a.e_tab_persist = 0;
    a.e_slot_bytes =
        (int)((((size_t)S.e_ns * (plan.prec == SGDLM_PREC_FP64 ? 8 : 4)) + 15) & ~(size_t)15);
  }
}

void set_fast_layout(LuArgs &a, const Structure &S, const LuPlan &plan) {
  a.M = plan.fast_M;
  a.e_pre_base = (int)plan.pre_base;
  a.e_tab_persist = plan.tab_persist;
  a.e_slot_bytes = (int)((((size_t)S.e_ns * (plan.prec == SGDLM_PREC_FP64 ? 8 : 4)) + 15) & ~(size_t)15);
  a.e_xs_base = (int)plan.xs_base;
  a.e_flag_off = (int)plan.flag_off;
  a.e_slot_save = plan.slot_stride ? static_cast<char *>(a.work) + plan.slot_save_off : nullptr;
  a.e_slot_stride = (long)plan.slot_stride;
}

} // namespace

int lu_dense_cap(size_t lds_max) {
  int cap = 0;
  for (int M2 = 16; M2 <= 1024; M2 += 16)
    if (fast_lds_bytes(M2, 64, MODE_NATIVE, 4) <= lds_max / 2 &&
        fast_lds_bytes(M2, 64, MODE_NATIVE, 8) <= lds_max)
      cap = M2;
  return cap;
}

int lu_plan(const Structure &S, int prec, int solve, int n_cu, size_t lds_max,
            int want_threads, int want_nb, LuPlan *plan, int want_kind) {
  const int M = S.M;
  const int M2 = S.e_M2;
  const size_t ts = (prec == SGDLM_PREC_FP64) ? 8 : 4;
  const int cands[3] = {64, 32, 16};
  plan->prec = prec;
  plan->nb = 0;
  plan->pivot_nb = 0;
  plan->slot_stride = 0;
  plan->slot_save_off = 0;
  plan->fast_M = M2;
  plan->grid = n_cu > 0 ? n_cu : 256;
  const size_t slot_bytes = ((size_t)S.e_ns * ts + 15) & ~(size_t)15;
  const size_t pre_bytes = slot_bytes + (size_t)S.e_tab_words * 4;
  /* Fast-kernel configurations in order of preference: 256 threads with two
   * workgroups per CU (LDS within half a CU, M <= 512), then 512 threads
   * with one workgroup per CU. */
  for (int pass = 0; pass < 2 && !plan->nb && !solve && want_kind != SGDLM_LU_PIVOTING; ++pass) {
    const int threads = pass == 0 ? 256 : 512;
    const int per_cu = pass == 0 ? 2 : 1;
    if (want_threads && want_threads != threads)
      continue;
    if (threads == 256 && M2 > 512)
      continue;
    const size_t cap = lds_max / per_cu;
    for (int i = 0; i < 3 && !plan->nb; ++i) {
      const int nb = cands[i];
      if (want_nb && want_nb != nb)
        continue;
      if (prec == SGDLM_PREC_BF16X3 && nb < 32)
        continue;
      if (threads == 256 && nb > 32 && prec == SGDLM_PREC_BF16X3)
        continue;
      const bool fp64_narrow_wide =
          prec == SGDLM_PREC_FP64 && nb == 16 && M2 > (threads == 256 ? 256 : 512);
      if (M2 > 1024 || fp64_narrow_wide)
        continue;
      const size_t dense =
          fast_lds_bytes(M2, nb, prec == SGDLM_PREC_BF16X3 ? MODE_BF16X3 : MODE_NATIVE, ts);
      if (dense + pre_bytes <= cap) {
        /* prefix region after the dense one: tables loaded once per workgroup */
        plan->lds = dense + pre_bytes;
        plan->pre_base = dense;
        plan->flag_off = dense - 16;
        plan->tab_persist = 1;
      } else {
        /* prefix region aliased with the dense one; the early-exit flag
         * moves past the prefix region when that reaches the flag slot */
        size_t flag = dense - 16, lds = std::max(dense, pre_bytes);
        if (pre_bytes > dense - 16) {
          flag = (pre_bytes + 15) & ~(size_t)15;
          lds = std::max(dense, flag + 16);
        }
        if (lds > cap)
          continue;
        plan->lds = lds;
        plan->pre_base = 0;
        plan->flag_off = flag;
        plan->tab_persist = 0;
      }
      plan->nb = nb;
      plan->threads = threads;
      plan->fast_grid = plan->grid * per_cu;
    }
  }
  if (solve && prec != SGDLM_PREC_BF16X3) {
    /* forecast solves: two 256-thread workgroups per CU (M2 <= 512) when
     * that fits in half a CU, else one 512-thread workgroup.  Either the
     * prefix slots and tables stay in LDS after the dense region, then x
     * (n1 + M2 doubles), or (FP32 with 64-wide panels) both alias the dense
     * region and the slots wait in the workspace; with FP64 factors the
     * aliased layout measured slower and is not used */
    const size_t xs_bytes = (((size_t)S.e_n1 + M2) * sizeof(double) + 15) & ~(size_t)15;
    for (int pass = 0; pass < 2 && !plan->nb; ++pass) {
      const int threads = pass == 0 ? 256 : 512;
      const int per_cu = pass == 0 ? 2 : 1;
      for (int i = 0; i < 2 && !plan->nb; ++i) {
        const int nb = cands[i];
        if (M2 > 1024 || (threads == 256 && M2 > 512))
          continue;
        const size_t dense = fast_lds_bytes(M2, nb, MODE_NATIVE, ts);
        if (dense + pre_bytes + xs_bytes <= lds_max / per_cu &&
            (threads == 512 || nb == 32)) {
          plan->nb = nb;
          plan->threads = threads;
          plan->fast_grid = plan->grid * per_cu;
          plan->lds = dense + pre_bytes + xs_bytes;
          plan->pre_base = dense;
          plan->flag_off = dense - 16;
          plan->tab_persist = 1;
          plan->xs_base = dense + pre_bytes;
        } else if (threads == 256 && prec == SGDLM_PREC_FP32 &&
                   ((pre_bytes + 15) & ~(size_t)15) + xs_bytes + 16 <= dense &&
                   dense <= lds_max / per_cu) {
          /* the prefix region and x aliased with the dense one, the prefix
           * slots kept in global memory across the dense LU (64-wide
           * panels) */
          plan->nb = nb;
          plan->threads = threads;
          plan->fast_grid = plan->grid * per_cu;
          plan->lds = dense;
          plan->pre_base = 0;
          plan->flag_off = dense - 16;
          plan->tab_persist = 0;
          plan->xs_base = (pre_bytes + 15) & ~(size_t)15;
          plan->slot_stride = ((size_t)S.e_ns * ts + 255) & ~(size_t)255;
        }
      }
    }
  }
  if (!solve && (want_threads || want_nb || want_kind == SGDLM_LU_BLOCKED) && !plan->nb)
    return SGDLM_ERR_ARG;
  /* partial pivoting: two workgroups per CU when a 32- or 16-wide panel
   * fits in half a CU, else one with the widest panel that fits */
  plan->pivot_grid = plan->grid;
  /* log-determinants with a sparse prefix: the pivoting kernel re-eliminates
   * the prefix in LDS (aliasing its panel region; the prefix flag after both)
   * and factors only the Schur complement */
  const bool pv_schur = !solve && S.e_n1 > 0 && M2 < M;
  plan->pivot_schur_M = 0;
  plan->pivot_flag_off = 0;
  for (int per_cu = 2; per_cu >= 1 && !plan->pivot_nb; --per_cu) {
    for (int i = 0; i < 3 && !plan->pivot_nb; ++i) {
      const int nb = cands[i];
      if ((prec == SGDLM_PREC_FP64 || per_cu == 2) && nb > 32)
        continue;
      size_t lds = pivot_lds_bytes(M, nb, ts, solve != 0);
      if (pv_schur)
        lds = ((std::max(lds, pre_bytes) + 15) & ~(size_t)15) + 16;
      if (lds <= lds_max / per_cu) {
        plan->pivot_nb = nb;
        plan->pivot_lds = lds;
        plan->pivot_grid = plan->grid * per_cu;
      }
    }
  }
  if (plan->pivot_nb == 0)
    return SGDLM_ERR_ARG;
  if (pv_schur) {
    plan->pivot_schur_M = M2;
    plan->pivot_flag_off = plan->pivot_lds - 16;
  }
  /* FP32 / BF16x3 log-determinants of dense blocks up to 288: the
   * register-resident kernel, unless another kind or a fast-kernel shape
   * was asked for */
  plan->rs_MT = 0;
  plan->rs_chunk = 0;
  if ((solve ? prec == SGDLM_PREC_FP32 : prec != SGDLM_PREC_FP64) && !want_threads && !want_nb &&
      (want_kind == SGDLM_LU_AUTO || want_kind == SGDLM_LU_REGISTER) && S.e_rb_rows > 0) {
    plan->rs_MT = S.e_rb_tiles;
    plan->rs_nw = S.e_rb_nw;
    plan->rs_grid = plan->grid * (S.e_rb_nw == 4 ? 2 : 1);
    plan->rs_lds = S.e_rb_lds;
    plan->rs_panel_base = S.e_rb_panel_base;
    plan->rs_flag_base = S.e_rb_flag_base;
    if (solve) {
      /* forecast solves: factors of up to rs_chunk draws in the workspace
       * (at most 8 GB), a whole number of units that are whole rounds of
       * both the factor kernel's grid and the substitution kernel's
       * wavefronts (one draw each), so that no chunk but the last leaves
       * either launch a partial round; 4 GB and the factor kernel's round
       * alone when such a unit is larger */
      plan->rs_sol_stride = rs_solve_stride(S, plan->rs_MT);
      if (lu_subst_shape(S, plan->rs_MT, n_cu, &plan->subst_grid, &plan->subst_lds) ||
          plan->subst_lds > lds_max) {
        plan->rs_MT = 0;
        plan->rs_chunk = 0;
      } else {
        const size_t per_draw = plan->rs_sol_stride * 4, g = (size_t)plan->rs_grid;
        const size_t sw = (size_t)plan->subst_grid * 4;
        size_t unit = g;
        while (unit % sw)
          unit += g;
        size_t n_units = ((size_t)8 << 30) / per_draw / unit;
        if (n_units == 0) {
          unit = g;
          n_units = std::max<size_t>(1, std::min<size_t>(((size_t)4 << 30) / per_draw / g, 64));
        }
        plan->rs_unit = (int)unit;
        plan->rs_chunk = (int)(n_units * unit);
      }
    }
  }
  if (solve && want_kind == SGDLM_LU_PIVOTING)
    plan->nb = 0;
  if (want_kind == SGDLM_LU_REGISTER && !plan->rs_MT)
    return SGDLM_ERR_ARG;
  /* the register kernel with partial pivoting (FP32 log-determinants): its
   * row-exchange buffer (32 rows of the trailing column tiles) after the
   * staged panel tiles */
  plan->rs_piv = 0;
  plan->rs_swap_base = 0;
  if (want_kind == SGDLM_LU_REGISTER_PIVOT) {
    if (solve || prec != SGDLM_PREC_FP32 || S.e_rb_rows <= 0)
      return SGDLM_ERR_ARG;
    const int MT = S.e_rb_tiles;
    const size_t swap_base = (S.e_rb_panel_base + rs_panel_bytes(MT) + 15) & ~(size_t)15;
    const size_t need = swap_base + (size_t)32 * (MT - 1) * 16 * sizeof(float);
    if (need > lds_max / (S.e_rb_nw == 4 ? 2 : 1))
      return SGDLM_ERR_ARG;
    plan->rs_MT = MT;
    plan->rs_nw = S.e_rb_nw;
    plan->rs_grid = plan->grid * (S.e_rb_nw == 4 ? 2 : 1);
    plan->rs_lds = std::max(S.e_rb_lds, need);
    plan->rs_panel_base = S.e_rb_panel_base;
    plan->rs_flag_base = S.e_rb_flag_base;
    plan->rs_piv = 1;
    plan->rs_swap_base = swap_base;
  }
  /* small models: one wavefront per draw on the whole matrix */
  plan->wave_M = 0;
  plan->wave_grid = 0;
  if (want_kind == SGDLM_LU_WAVE) {
    if (solve || !(plan->wave_M = lu_wave_dim(S.M, prec)))
      return SGDLM_ERR_ARG;
    plan->wave_grid = lu_wave_grid(plan->wave_M, prec, n_cu);
    plan->nb = 0;
    plan->rs_MT = 0;
  }
  /* dense blocks after the prefix up to 128 rows: one wavefront per draw on
   * the Schur complement */
  plan->wsch_M = 0;
  plan->wsch_grid = 0;
  plan->wsch_lds = 0;
  if (want_kind == SGDLM_LU_WAVE_PREFIX) {
    const size_t lds = solve ? 0 : lu_wsch_lds(S, prec);
    if (!lds || lds > lds_max)
      return SGDLM_ERR_ARG;
    const int per_cu = lu_wsch_per_cu(S, prec, lds);
    if (per_cu < 1)
      return SGDLM_ERR_ARG;
    plan->wsch_M = S.e_wb_M;
    plan->wsch_lds = lds;
    plan->wsch_grid = std::max(1, n_cu) * per_cu;
    plan->nb = 0;
    plan->rs_MT = 0;
  }
  plan->work_bytes =
      (size_t)std::max(plan->pivot_grid, std::max(plan->grid, plan->fast_grid)) * M * M * ts;
  if (plan->slot_stride) {
    plan->slot_save_off = (plan->work_bytes + 255) & ~(size_t)255;
    plan->work_bytes = plan->slot_save_off + (size_t)plan->fast_grid * plan->slot_stride;
  }
  if (plan->rs_chunk) {
    plan->rs_sol_off = (plan->work_bytes + 255) & ~(size_t)255;
    plan->work_bytes = plan->rs_sol_off + (size_t)plan->rs_chunk * plan->rs_sol_stride * 4 +
                       (size_t)plan->rs_chunk * 4;
  }
  return SGDLM_OK;
}

int set_device_int(int *p, int v, hipStream_t st) { return set_int(p, v, st); }

hipError_t allow_dynamic_lds(const void *fn, size_t lds) {
  static std::mutex mu;
  static std::map<const void *, size_t> allowed;
  std::lock_guard<std::mutex> lk(mu);
  auto it = allowed.find(fn);
  if (it != allowed.end() && it->second >= lds)
    return hipSuccess;
  int dev = 0, mx = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&mx, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    mx = 0;
  const size_t want = std::max(lds, (size_t)std::max(mx, 0));
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)want) ==
      hipSuccess) {
    allowed[fn] = want;
    return hipSuccess;
  }
  (void)hipGetLastError();
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess)
    allowed[fn] = lds;
  return e;
}

int lu_plan_kind(const LuPlan &plan) {
  if (plan.wsch_M > 0)
    return SGDLM_LU_WAVE_PREFIX;
  if (plan.wave_M > 0)
    return SGDLM_LU_WAVE;
  if (plan.rs_MT > 0)
    return plan.rs_piv ? SGDLM_LU_REGISTER_PIVOT : SGDLM_LU_REGISTER;
  return plan.nb ? SGDLM_LU_BLOCKED : SGDLM_LU_PIVOTING;
}

int lu_plan_cap(const Structure &S, size_t work_cap, LuPlan *plan) {
  const size_t ts = (plan->prec == SGDLM_PREC_FP64) ? 8 : 4;
  const size_t wg_full = (size_t)S.M * S.M * ts;   /* partial-pivoting kernel */
  const size_t wg_fast = (size_t)plan->fast_M * plan->fast_M * ts + plan->slot_stride;
  const size_t wg = std::max(wg_full, plan->nb ? wg_fast : (size_t)0) + 256;
  if (plan->work_bytes <= work_cap)
    return SGDLM_OK;
  const int max_wg = (int)std::min<size_t>(work_cap / wg, 1 << 20);
  if (max_wg < 1)
    return SGDLM_ERR_ARG;
  plan->grid = std::min(plan->grid, max_wg);
  plan->pivot_grid = std::min(plan->pivot_grid, max_wg);
  if (plan->nb)
    plan->fast_grid = std::min(plan->fast_grid, max_wg);
  plan->work_bytes =
      (size_t)std::max(plan->pivot_grid, std::max(plan->grid, plan->fast_grid)) * S.M * S.M * ts;
  if (plan->slot_stride) {
    plan->slot_save_off = (plan->work_bytes + 255) & ~(size_t)255;
    plan->work_bytes = plan->slot_save_off + (size_t)plan->fast_grid * plan->slot_stride;
  }
  return plan->work_bytes <= work_cap ? SGDLM_OK : SGDLM_ERR_ARG;
}

int launch_lu_logdet(const Structure &S, const LuPlan &plan, long N,
                     const double *theta, void *work, double *logdet,
                     int *flags, hipStream_t st) {
  LuArgs a = base_args(S, N, theta, work);
  a.logdet = logdet;
  a.flag_count = flags;
  a.flag_list = flags + 1;
  if (plan.wave_M > 0) {
    /* every draw pivoted by its wavefront: no fallback draws */
    if (set_int(flags, 0, st))
      return SGDLM_ERR_DEVICE;
    return launch_lu_wave(S, plan, N, theta, logdet, st);
  }
  if (plan.wsch_M > 0) {
    /* one wavefront per draw after the prefix; draws whose prefix fails go
     * to the partial-pivoting kernel */
    if (set_int(flags, 0, st))
      return SGDLM_ERR_DEVICE;
    const int rc = launch_lu_wsch(S, plan, N, theta, logdet, flags, st);
    if (rc)
      return rc;
    set_pivot_layout(a, S, plan);
    return dispatch_pivot<false>(a, plan, st);
  }
  if (plan.nb == 0 && plan.rs_MT == 0) {
    /* no fast configuration: partial pivoting for every draw (all of them
     * count as fallback draws) */
    a.flag_count = nullptr;
    a.flag_list = nullptr;
    if (set_int(flags, (int)N, st))
      return SGDLM_ERR_DEVICE;
    set_pivot_layout(a, S, plan);
    return dispatch_pivot<false>(a, plan, st);
  }
  if (set_int(flags, 0, st))
    return SGDLM_ERR_DEVICE;
  int rc;
  if (plan.rs_MT > 0) {
    rc = launch_lu_rs(S, plan, N, theta, logdet, flags, st);
  } else {
    set_fast_layout(a, S, plan);
    rc = dispatch_fast<false>(a, plan, st);
  }
  if (rc)
    return rc;
  set_pivot_layout(a, S, plan);
  if ((rc = dispatch_pivot<false>(a, plan, st)))
    return rc;
  /* the register kernel with partial pivoting factored every draw with
   * row interchanges: all of them count as fallback draws, as for the
   * partial-pivoting kernel alone (after the launch above has read the
   * list of the draws it could not factor) */
  if (plan.rs_MT > 0 && plan.rs_piv && set_int(flags, (int)N, st))
    return SGDLM_ERR_DEVICE;
  return SGDLM_OK;
}

int launch_lu_solve(const Structure &S, const LuPlan &plan, long N,
                    const double *theta, void *work, const double *rhs,
                    double *y, int refine, int *flags, hipStream_t st) {
  if (plan.rs_chunk > 0) {
    /* register kernel: per chunk, factors written out, one wavefront per
     * draw substitutes, flagged draws solved with partial pivoting; the
     * fallback count summed over the chunks in flags[N + 1] */
    float *sol = reinterpret_cast<float *>(static_cast<char *>(work) + plan.rs_sol_off);
    int *ok = reinterpret_cast<int *>(sol + (size_t)plan.rs_chunk * plan.rs_sol_stride);
    int *total = flags + N + 1;
    if (int_op(total, nullptr, 0, st))
      return SGDLM_ERR_DEVICE;
    /* equal chunks, each a whole number of units */
    const long nch = (N + plan.rs_chunk - 1) / plan.rs_chunk;
    const long g = std::max(1, plan.rs_unit);
    const long per = std::min<long>(plan.rs_chunk, ((N + nch - 1) / nch + g - 1) / g * g);
    for (long c0 = 0; c0 < N; c0 += per) {
      const long n = std::min<long>(per, N - c0);
      const double *th = theta + (size_t)c0 * S.P;
      int rc;
      if (set_int(flags, 0, st))
        return SGDLM_ERR_DEVICE;
      if ((rc = launch_lu_rs_factor(S, plan, n, th, sol, ok, flags, st)) ||
          (rc = launch_lu_subst(S, plan, n, th, rhs + (size_t)c0 * S.m, y + (size_t)c0 * S.m, sol,
                                ok, refine, st)))
        return rc;
      LuArgs a = base_args(S, n, th, work);
      a.rhs = rhs + (size_t)c0 * S.m;
      a.y = y + (size_t)c0 * S.m;
      a.refine = refine;
      a.flag_count = flags;
      a.flag_list = flags + 1;
      if ((rc = dispatch_pivot<true>(a, plan, st)) || int_op(total, flags, 1, st))
        return rc ? rc : SGDLM_ERR_DEVICE;
    }
    return int_op(flags, total, 2, st);
  }
  LuArgs a = base_args(S, N, theta, work);
  a.rhs = rhs;
  a.y = y;
  a.refine = refine;
  if (plan.nb == 0) {
    if (flags && set_int(flags, (int)N, st))
      return SGDLM_ERR_DEVICE;
    return dispatch_pivot<true>(a, plan, st);
  }
  /* no-interchange fast path with the sparse prefix; draws it flags are
   * solved by the partial-pivoting kernel */
  a.flag_count = flags;
  a.flag_list = flags + 1;
  if (set_int(flags, 0, st))
    return SGDLM_ERR_DEVICE;
  set_fast_layout(a, S, plan);
  int rc = dispatch_fast<true>(a, plan, st);
  if (rc)
    return rc;
  a.M = S.M;
  return dispatch_pivot<true>(a, plan, st);
}

int lu_grp_stages_shape(const Structure &S, const LuPlan &plan, long N, const double *theta,
                        double *logdet, int *flags, int *flag_init, LuGrpKernel *first,
                        LuGrpKernel *second);

int lu_grp_stages(const Structure &S, const LuPlan &plan, long N, const double *theta,
                  double *logdet, int *flags, double draw_wg_ms, int *flag_init,
                  LuGrpKernel *first, LuGrpKernel *second) {
  *first = LuGrpKernel();
  *second = LuGrpKernel();
  const int rc = lu_grp_stages_shape(S, plan, N, theta, logdet, flags, flag_init, first, second);
  /* the measured time of a draw on one workgroup of the first stage's
   * kernel replaces the size model */
  if (!rc && draw_wg_ms > 0.0)
    first->cost = (double)N * draw_wg_ms;
  return rc;
}

int lu_grp_stages_shape(const Structure &S, const LuPlan &plan, long N, const double *theta,
                        double *logdet, int *flags, int *flag_init, LuGrpKernel *first,
                        LuGrpKernel *second) {
  const size_t ts = plan.prec == SGDLM_PREC_FP64 ? 8 : 4;
  /* the fallback: partial pivoting of the flagged draws (or of every draw) */
  auto pivot_stage = [&](bool all, LuGrpKernel *k) {
    LuArgs a = base_args(S, N, theta, nullptr);
    a.logdet = logdet;
    a.flag_count = all ? nullptr : flags;
    a.flag_list = all ? nullptr : flags + 1;
    set_pivot_layout(a, S, plan);
    const LuKern lk = select_pivot<false>(plan);
    if (!lk.grp)
      return SGDLM_ERR_ARG;
    k->fn = lk.grp;
    k->threads = lk.threads;
    k->per_cu = std::max(1, plan.pivot_grid / std::max(1, plan.grid));
    k->lds = plan.pivot_lds;
    k->ws_wg = (size_t)S.M * S.M * ts;
    k->units = N;
    const double Md = (double)(plan.pivot_schur_M > 0 ? plan.pivot_schur_M : S.M);
    /* ms per draw on one workgroup (size model: a latency-bound column
     * chain plus the cubic updates); every draw, or (fallback) a share of
     * them that is not known yet */
    k->cost = (double)N * (1e-3 * Md + 1e-8 * Md * Md * Md) * (all ? 1.0 : 0.05);
    k->args.assign(reinterpret_cast<const unsigned char *>(&a),
                   reinterpret_cast<const unsigned char *>(&a) + sizeof(a));
    k->work_field = (long)offsetof(LuArgs, work);
    return SGDLM_OK;
  };
  if (plan.wave_M > 0) {
    *flag_init = 0;
    return lu_grp_wave(S, plan, N, theta, logdet, first);
  }
  if (plan.wsch_M > 0) {
    *flag_init = 0;
    const int rc = lu_grp_wsch(S, plan, N, theta, logdet, flags, first);
    return rc ? rc : pivot_stage(false, second);
  }
  if (plan.nb == 0 && plan.rs_MT == 0) {
    *flag_init = (int)N;
    return pivot_stage(true, first);
  }
  *flag_init = 0;
  int rc;
  if (plan.rs_MT > 0) {
    rc = lu_grp_rs(S, plan, N, theta, logdet, flags, first);
  } else {
    LuArgs a = base_args(S, N, theta, nullptr);
    a.logdet = logdet;
    a.flag_count = flags;
    a.flag_list = flags + 1;
    set_fast_layout(a, S, plan);
    const LuKern lk = select_fast<false>(plan, a.M);
    rc = lk.grp ? SGDLM_OK : SGDLM_ERR_ARG;
    if (!rc) {
      first->fn = lk.grp;
      first->threads = lk.threads;
      first->per_cu = std::max(1, plan.fast_grid / std::max(1, plan.grid));
      first->lds = plan.lds;
      first->ws_wg = (size_t)plan.fast_M * plan.fast_M * ts;
      first->units = N;
      const double Md = (double)plan.fast_M;
      first->cost = (double)N * (2e-4 * Md + 6e-9 * Md * Md * Md + 5e-7 * (double)S.e_fma);
      first->args.assign(reinterpret_cast<const unsigned char *>(&a),
                         reinterpret_cast<const unsigned char *>(&a) + sizeof(a));
      first->work_field = (long)offsetof(LuArgs, work);
    }
  }
  if (rc)
    return rc;
  return pivot_stage(false, second);
}

} // namespace sgdlm
