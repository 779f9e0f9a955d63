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
 * Design (one persistent workgroup of 4 wavefronts per CU, looping over
 * draws):
 *   - The matrix of a draw lives in a per-workgroup global workspace
 *     (row-major, M x M, M = m rounded up to 16, identity padding), hot in
 *     L2 / Infinity Cache.  Row interchanges are never applied to memory: a
 *     logical->physical row map in LDS is permuted instead.
 *   - Blocked right-looking LU with panel width NB.  The panel
 *     (rows K..M, columns K..K+NB) is factored in LDS with partial pivoting
 *     (the pivot search picks the largest |a|, lowest row on ties, like
 *     LAPACK's getf2).
 *   - U12 = L11^{-1} A12 is formed per 16-column tile on MFMA from the
 *     inverses of the 16 x 16 diagonal blocks of L11, and kept in registers
 *     in the accumulator layout; the contraction index of the trailing
 *     update is permuted so that this accumulator layout is directly the
 *     MFMA B-operand layout (no shuffles).
 *   - A22 -= L21 U12 on MFMA: f64 16x16x4 (FP64), f32 16x16x4 (FP32) or
 *     three bf16 16x16x32 products hi*hi + hi*lo + lo*hi with FP32
 *     accumulation (BF16x3).  Work items are (column tile, row tile) pairs,
 *     split into 4 contiguous ranges, so the waves stay balanced.
 *   - log|det| = sum log|u_kk| is accumulated in FP64 during the panel
 *     factorisation; a zero or non-finite pivot marks the draw singular
 *     (log|det| = -inf, weight 0, as detFromLU's NaN/inf -> 0).
 *   - Forecast mode additionally keeps L and U in the workspace, solves
 *     L U y = P rhs with FP64 accumulation and optionally applies one step
 *     of iterative refinement with the FP64 residual rhs - (I - Gamma) y
 *     evaluated on the sparse Gamma.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <math.h>

#include "sgdlm_internal.h"

namespace sgdlm {
namespace {

constexpr int WG = 256; /* 4 wavefronts */

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

template <typename T> struct Vec4;
template <> struct Vec4<float> { typedef f32x4 type; };
template <> struct Vec4<double> { typedef f64x4 type; };

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

template <typename T, int NB, int MODE> struct LuShape {
  /* LDS row stride of the panel: odd for conflict-free column access in the
   * panel factorisation; 16-byte aligned rows for the bf16 operand loads. */
  static constexpr int PS = (MODE == MODE_BF16X3) ? NB + 4 : NB + 1;
  static constexpr int NQ = NB / 16;
};

__host__ __device__ inline size_t lu_lds_bytes(int M, int NB, int PS,
                                               size_t tsize, bool solve) {
  size_t b = (size_t)M * PS * tsize;       /* panel                        */
  b += (size_t)NB * 16 * tsize;            /* diagonal-block inverses      */
  b += (size_t)M * sizeof(int);            /* row map                      */
  b = (b + 15) & ~(size_t)15;
  b += 4 * (sizeof(double) + sizeof(int)); /* pivot reduction              */
  b = (b + 15) & ~(size_t)15;
  if (solve)
    b += (size_t)M * sizeof(double);       /* solve vector                 */
  return b;
}

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
};

template <typename T, int NB, int MODE, bool SOLVE>
__global__ __launch_bounds__(WG) void k_lu(LuArgs a) {
  typedef LuShape<T, NB, MODE> S;
  typedef typename Vec4<T>::type V4;
  constexpr int PS = S::PS;
  constexpr int NQ = S::NQ;
  extern __shared__ __attribute__((aligned(16))) char smem[];

  const int M = a.M;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int l16 = lane & 15;
  const int g = lane >> 4;

  T *Pn = reinterpret_cast<T *>(smem);
  T *Linv = Pn + (size_t)M * PS;
  int *rowmap = reinterpret_cast<int *>(Linv + NB * 16);
  size_t off = (size_t)M * PS * sizeof(T) + (size_t)NB * 16 * sizeof(T) +
               (size_t)M * sizeof(int);
  off = (off + 15) & ~(size_t)15;
  double *redv = reinterpret_cast<double *>(smem + off);
  int *redi = reinterpret_cast<int *>(redv + 4);
  off += 4 * (sizeof(double) + sizeof(int));
  off = (off + 15) & ~(size_t)15;
  double *sv = reinterpret_cast<double *>(smem + off); /* SOLVE only */

  T *A = reinterpret_cast<T *>(a.work) + (size_t)blockIdx.x * M * M;

  for (long d = blockIdx.x; d < a.n_draws; d += gridDim.x) {
    const double *th = a.theta + (size_t)d * a.P;

    /* ---- assemble A = I - Gamma_d (row-major, identity padding) ---- */
    {
      const size_t nvec = (size_t)M * M * sizeof(T) / 16;
      float4 *A4 = reinterpret_cast<float4 *>(A);
      const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
      for (size_t q = tid; q < nvec; q += WG)
        A4[q] = z;
    }
    __syncthreads();
    for (int r = tid; r < M; r += WG) {
      A[(size_t)r * M + r] = (T)1;
      rowmap[r] = r;
    }
    for (int e = tid; e < a.n_gam; e += WG)
      A[(size_t)a.g_row[e] * M + a.g_col[e]] = (T)(-th[a.g_idx[e]]);
    __syncthreads();

    double logdet = 0.0;
    int singular = 0;

    for (int K = 0; K < M; K += NB) {
      const int jb = min(NB, M - K);
      const int nr = M - K;
      const int nq = jb >> 4;

      /* ---- 1. load panel (logical rows K..M, columns K..K+NB) ---- */
      for (int q = tid; q < nr * NB; q += WG) {
        const int r = q / NB;
        const int c = q - r * NB;
        Pn[r * PS + c] =
            (c < jb) ? A[(size_t)rowmap[K + r] * M + K + c] : (T)0;
      }
      __syncthreads();

      /* ---- 2. panel factorisation with partial pivoting ---- */
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
        for (int w = 1; w < 4; ++w) {
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

      /* ---- 3. inverses of the 16x16 unit-lower diagonal blocks of L11 ---- */
      for (int b = wave; b < nq; b += 4) {
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
      if constexpr (SOLVE) {
        /* keep L (below the diagonal) and U11 of the panel in the workspace */
        for (int q = tid; q < nr * jb; q += WG) {
          const int r = q / jb;
          const int c = q - r * jb;
          A[(size_t)rowmap[K + r] * M + K + c] = Pn[r * PS + c];
        }
      }
      if constexpr (MODE == MODE_BF16X3) {
        /* Split -L21 into bf16 hi/lo in place, in the permuted contraction
         * order used below: chunk (kb, g) holds columns
         * 32kb + 4g + {0..3} and 32kb + 16 + 4g + {0..3}. */
        __syncthreads(); /* Linv reads rows < jb only, but keep it simple */
        for (int r = jb + tid; r < nr; r += WG) {
          float v[NB];
#pragma unroll
          for (int c = 0; c < NB; ++c)
            v[c] = (float)Pn[r * PS + c];
          unsigned short *row = reinterpret_cast<unsigned short *>(Pn + r * PS);
#pragma unroll
          for (int kb = 0; kb < NB / 32; ++kb)
#pragma unroll
            for (int gg = 0; gg < 4; ++gg) {
#pragma unroll
              for (int e = 0; e < 8; ++e) {
                const int col = 32 * kb + 16 * (e >> 2) + 4 * gg + (e & 3);
                const float x = -v[col];
                const unsigned short hi = f2bf(x);
                const unsigned short lo = f2bf(x - bf2f(hi));
                row[(kb * 4 + gg) * 16 + e] = hi;
                row[(kb * 4 + gg) * 16 + 8 + e] = lo;
              }
            }
        }
      }
      __syncthreads();

      /* ---- 4. U12 and trailing update A22 -= L21 U12 on MFMA ---- */
      const int ntc = (M - K - jb) >> 4;
      if (ntc > 0) {
        V4 uC[NQ];
        u16x8 bhi[(NQ + 1) / 2], blo[(NQ + 1) / 2];
        int cur_ct = -1;
        long lo_it, hi_it;
        if constexpr (SOLVE) {
          lo_it = 0;
          hi_it = 0; /* unused: SOLVE assigns whole column tiles per wave */
        } else {
          const long tot = (long)ntc * ntc;
          lo_it = tot * wave / 4;
          hi_it = tot * (wave + 1) / 4;
        }
        auto compute_u12 = [&](int ct) {
          const int c0 = K + jb + 16 * ct;
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            V4 u = {0, 0, 0, 0};
            if (q < nq) {
              V4 acc;
#pragma unroll
              for (int r = 0; r < 4; ++r)
                acc[r] = A[(size_t)rowmap[K + 16 * q + 4 * g + r] * M + c0 + l16];
#pragma unroll
              for (int aa = 0; aa < NQ; ++aa)
                if (aa < q)
#pragma unroll
                  for (int r = 0; r < 4; ++r)
                    acc = mfma16(-Pn[(16 * q + l16) * PS + 16 * aa + 4 * g + r],
                                 uC[aa][r], acc);
#pragma unroll
              for (int r = 0; r < 4; ++r)
                u = mfma16(Linv[q * 256 + l16 * 16 + 4 * g + r], acc[r], u);
              if constexpr (SOLVE) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                  A[(size_t)rowmap[K + 16 * q + 4 * g + r] * M + c0 + l16] = u[r];
              }
            }
            uC[q] = u;
          }
          if constexpr (MODE == MODE_BF16X3) {
#pragma unroll
            for (int kb = 0; kb < (NQ + 1) / 2; ++kb)
#pragma unroll
              for (int e = 0; e < 8; ++e) {
                const int q = 2 * kb + (e >> 2);
                const float x = (q < NQ) ? (float)uC[q < NQ ? q : 0][e & 3] : 0.f;
                const unsigned short h = f2bf(x);
                bhi[kb][e] = h;
                blo[kb][e] = f2bf(x - bf2f(h));
              }
          }
        };
        auto update_tile = [&](int ct, int rt) {
          const int c0 = K + jb + 16 * ct;
          const int row0 = K + jb + 16 * rt;
          int pr[4];
          V4 c;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            pr[r] = rowmap[row0 + 4 * g + r];
            c[r] = A[(size_t)pr[r] * M + c0 + l16];
          }
          const int prow = jb + 16 * rt + l16; /* panel row of this lane */
          if constexpr (MODE == MODE_BF16X3) {
            const char *rowp = reinterpret_cast<const char *>(Pn + prow * PS);
#pragma unroll
            for (int kb = 0; kb < (NQ + 1) / 2; ++kb) {
              if (2 * kb < nq) {
                const u16x8 ahi =
                    *reinterpret_cast<const u16x8 *>(rowp + (kb * 4 + g) * 32);
                const u16x8 alo = *reinterpret_cast<const u16x8 *>(
                    rowp + (kb * 4 + g) * 32 + 16);
                c = mfma_bf16(alo, bhi[kb], c);
                c = mfma_bf16(ahi, blo[kb], c);
                c = mfma_bf16(ahi, bhi[kb], c);
              }
            }
          } else {
#pragma unroll
            for (int aa = 0; aa < NQ; ++aa)
              if (aa < nq)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                  c = mfma16(-Pn[prow * PS + 16 * aa + 4 * g + r], uC[aa][r], c);
          }
#pragma unroll
          for (int r = 0; r < 4; ++r)
            A[(size_t)pr[r] * M + c0 + l16] = (T)c[r];
        };

        if constexpr (SOLVE) {
          for (int ct = wave; ct < ntc; ct += 4) {
            compute_u12(ct);
            for (int rt = 0; rt < ntc; ++rt)
              update_tile(ct, rt);
          }
        } else {
          for (long it = lo_it; it < hi_it; ++it) {
            const int ct = (int)(it / ntc);
            const int rt = (int)(it - (long)ct * ntc);
            if (ct != cur_ct) {
              compute_u12(ct);
              cur_ct = ct;
            }
            update_tile(ct, rt);
          }
        }
      }
      __syncthreads();
    }

    if constexpr (!SOLVE) {
      if (tid == 0)
        a.logdet[d] = singular ? -INFINITY : logdet;
      continue;
    }

    /* ---- forecast: solve L U y = P rhs, optional refinement ---- */
    const double *rhs = a.rhs + (size_t)d * a.m;
    double *yout = a.y + (size_t)d * a.m;
    for (int pass = 0; pass < (a.refine ? 2 : 1); ++pass) {
      /* permuted right-hand side */
      for (int i = tid; i < M; i += WG) {
        const int src = rowmap[i];
        double b = 0.0;
        if (src < a.m) {
          if (pass == 0) {
            b = rhs[src];
          } else {
            /* residual of row src: rhs - y + sum_e theta_e y_col(e) */
            double acc = rhs[src] - yout[src];
            for (int e = a.g_rowptr[src]; e < a.g_rowptr[src + 1]; ++e)
              acc += th[a.g_idx[e]] * yout[a.g_col[e]];
            b = acc;
          }
        }
        sv[i] = b;
      }
      __syncthreads();
      /* forward substitution with unit-lower L (logical rows) */
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
      /* back substitution with U */
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
        const double v = singular ? NAN : sv[i];
        yout[i] = (pass == 0) ? v : yout[i] + v;
      }
      __syncthreads();
    }
  }
}

template <typename T, int NB, int MODE, bool SOLVE>
int launch_one(const LuArgs &a, const LuPlan &plan, hipStream_t st) {
  auto kern = k_lu<T, NB, MODE, SOLVE>;
  if (hipFuncSetAttribute((const void *)kern,
                          hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)plan.lds) != hipSuccess)
    return SGDLM_ERR_DEVICE;
  hipLaunchKernelGGL(kern, dim3(plan.grid), dim3(WG), plan.lds, st, a);
  return hipGetLastError() == hipSuccess ? SGDLM_OK : SGDLM_ERR_DEVICE;
}

template <bool SOLVE>
int dispatch(const LuArgs &a, const LuPlan &plan, hipStream_t st) {
  switch (plan.prec) {
  case SGDLM_PREC_FP64:
    if (plan.nb == 32)
      return launch_one<double, 32, MODE_NATIVE, SOLVE>(a, plan, st);
    return launch_one<double, 16, MODE_NATIVE, SOLVE>(a, plan, st);
  case SGDLM_PREC_FP32:
    if (plan.nb == 64)
      return launch_one<float, 64, MODE_NATIVE, SOLVE>(a, plan, st);
    if (plan.nb == 32)
      return launch_one<float, 32, MODE_NATIVE, SOLVE>(a, plan, st);
    return launch_one<float, 16, MODE_NATIVE, SOLVE>(a, plan, st);
  case SGDLM_PREC_BF16X3:
    if (SOLVE)
      return SGDLM_ERR_ARG;
    if (plan.nb == 64)
      return launch_one<float, 64, MODE_BF16X3, false>(a, plan, st);
    return launch_one<float, 32, MODE_BF16X3, false>(a, plan, st);
  }
  return SGDLM_ERR_ARG;
}

} // namespace

int lu_plan(const Structure &S, int prec, int solve, int n_cu, size_t lds_max,
            LuPlan *plan) {
  const int M = S.M;
  const size_t ts = (prec == SGDLM_PREC_FP64) ? 8 : 4;
  int cands[3] = {64, 32, 16};
  plan->prec = prec;
  plan->nb = 0;
  for (int i = 0; i < 3; ++i) {
    const int nb = cands[i];
    if (prec == SGDLM_PREC_FP64 && nb > 32)
      continue;
    if (prec == SGDLM_PREC_BF16X3 && nb < 32)
      continue;
    const int ps = (prec == SGDLM_PREC_BF16X3) ? nb + 4 : nb + 1;
    const size_t lds = lu_lds_bytes(M, nb, ps, ts, solve != 0);
    /* narrower panels only when the wide one does not fit */
    if (lds <= lds_max) {
      plan->nb = nb;
      plan->lds = lds;
      break;
    }
  }
  if (plan->nb == 0)
    return SGDLM_ERR_ARG;
  plan->grid = n_cu > 0 ? n_cu : 256;
  plan->work_bytes = (size_t)plan->grid * M * M * ts;
  return SGDLM_OK;
}

int launch_lu_logdet(const Structure &S, const LuPlan &plan, long N,
                     const double *theta, void *work, double *logdet,
                     hipStream_t st) {
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
  a.logdet = logdet;
  return dispatch<false>(a, plan, st);
}

int launch_lu_solve(const Structure &S, const LuPlan &plan, long N,
                    const double *theta, void *work, const double *rhs,
                    double *y, int refine, hipStream_t st) {
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
  a.rhs = rhs;
  a.y = y;
  a.refine = refine;
  return dispatch<true>(a, plan, st);
}

} // namespace sgdlm
