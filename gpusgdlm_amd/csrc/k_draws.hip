/*
 * k_draws.hip -- naive-posterior draws, forecast right-hand sides, the
 * multi-step evolution of forecast samples, and the packed <-> reference
 * theta layout conversion.
 *
 * Replaces sampleGamma2/make_gamma2 + curandGenerateNormal + cholesky +
 * gemmBatched(U^T z) + makeVB_MVN (rSGDLM/src/SGDLM/SGDLM.cu:623-677,
 * curand_manager.cuh:111-141, kernel_functions.cuh:71-83,568-634,844-902),
 * batchedComputeNuPlusMu (kernel_functions.cuh:1155-1178) and
 * evolve_lambdas_and_thetas (SGDLM.cu:679-762).
 *
 * Thread mapping: blockIdx.y = series j, consecutive threads = consecutive
 * draws of that series, so a wavefront shares p_j, n_j, s_j and reads the
 * same Cholesky factor (broadcast loads) with no loop divergence except the
 * gamma sampler's rare second attempt.  Per-thread accumulators live in LDS
 * with the thread index as the fastest dimension.
 * Random numbers come from counter-based Philox, so every (draw, series)
 * value is independent of launch shape and identical to the host oracle.
 */
#include <hip/hip_runtime.h>

#include "sgdlm_internal.h"
#include "sgdlm_numerics.h"

namespace sgdlm {
namespace {

constexpr int TB = 128;
constexpr int TS = TB + 1; /* LDS stride of per-thread accumulators */

/* Grids of the per-(draw, series) kernels: blockIdx.x = blocks of TB
 * draws (fastest: the workgroups in flight share a series' Cholesky factor
 * and parameters), blockIdx.y = series j (m < 46341 by the m^2 < 2^31
 * bound of sgdlm_init_memory, within the 65535 limit of a grid dimension). */
dim3 grid_nm(long N, int m) {
  const long nblk = (N + TB - 1) / TB;
  return dim3((unsigned)nblk, (unsigned)m);
}

/* One workgroup (one wavefront) per series: the right-looking Cholesky of
 * sg_cholesky_upper with one element per lane and a barrier per column.
 * Every element sees the same operations on the same operands in the same
 * order as the serial loop, so U is bit-identical to it. */
__device__ __forceinline__ void cholesky_body(int j, int max_p, const unsigned *p,
                                              const double *C, double *U) {
  extern __shared__ double a[]; /* [max_p][max_p], a[k * max_p + n] */
  const int t = threadIdx.x;
  const int pj = (int)p[j], ld = max_p, nn = max_p * max_p;
  const size_t o = (size_t)j * nn;
  for (int e = t; e < nn; e += 64) {
    const int k = e / ld, n = e - k * ld;
    a[e] = (n < pj && k >= n && k < pj) ? C[o + e] : 0.0;
  }
  __syncthreads();
  for (int n = 0; n < pj; ++n) {
    const double dn = a[n * ld + n];
    __syncthreads();
    if (!(dn > 0.0)) {
      /* zero or invalid pivot of a semi-definite block: the column is zero */
      for (int k = n + t; k < pj; k += 64)
        a[k * ld + n] = 0.0;
      __syncthreads();
      continue;
    }
    const double diag = __builtin_sqrt(dn);
    const double inv = 1.0 / diag;
    if (t == 0)
      a[n * ld + n] = diag;
    for (int k = n + 1 + t; k < pj; k += 64)
      a[k * ld + n] = a[k * ld + n] * inv;
    __syncthreads();
    const int nt = pj - n - 1;
    for (int e = t; e < nt * nt; e += 64) {
      const int k = n + 1 + e / nt, l = n + 1 + e % nt;
      if (l >= k)
        a[l * ld + k] = a[l * ld + k] - a[k * ld + n] * a[l * ld + n];
    }
    __syncthreads();
  }
  for (int e = t; e < nn; e += 64)
    U[o + e] = a[e];
}

__global__ __launch_bounds__(64) void k_cholesky(int m, int max_p, const unsigned *p,
                                                 const double *C, double *U) {
  (void)m;
  cholesky_body(blockIdx.x, max_p, p, C, U);
}

/* k_cholesky over the series of several models (one workgroup per series
 * of every model; dynamic LDS for the largest max_p). */
/* Also stores each model's LU fallback counter (flag_init) from the block
 * of its first series: the first launch of a group step, so that no
 * separate launch (which can wait for a free CU behind other parts' LU
 * workgroups) is needed before the LU. */
__global__ __launch_bounds__(64) void k_cholesky_grp(const VbModel *__restrict__ tab,
                                                     const int *__restrict__ ser_model) {
  const VbModel &md = tab[ser_model[blockIdx.x]];
  if ((int)blockIdx.x == md.ser0 && threadIdx.x == 0)
    md.lu_flags[0] = md.flag_init;
  cholesky_body((int)blockIdx.x - md.ser0, md.max_p, md.p, md.C, md.U);
}

__global__ __launch_bounds__(TB) void k_draw(long N, int m, int max_p,
                                             const unsigned *p, const int *off,
                                             long P, const double *m_t,
                                             const double *U, const double *n_t,
                                             const double *s_t, uint64_t seed,
                                             uint32_t epoch_arg, const uint32_t *epoch_dev,
                                             double *lam, double *theta) {
  extern __shared__ double sacc[]; /* [max_p][TS] */
  const uint32_t epoch = epoch_dev ? *epoch_dev : epoch_arg;
  const int j = blockIdx.y;
  const int pj = (int)p[j];
  for (long i = (long)blockIdx.x * TB + threadIdx.x; i < N; i += (long)gridDim.x * TB) {
    const double l = sg_draw_lambda_theta(
        pj, max_p, m_t + (size_t)j * max_p, U + (size_t)j * max_p * max_p, n_t[j],
        s_t[j], seed, epoch, (uint32_t)i, (uint32_t)j, sacc + threadIdx.x, TS);
    lam[(size_t)i * m + j] = l;
    double *out = theta + (size_t)i * P + off[j];
    for (int k = 0; k < pj; ++k)
      out[k] = sacc[k * TS + threadIdx.x];
  }
}

/* k_draw with the p_j accumulators in registers (p_j <= PMAX): the same
 * operations in the same order as sg_draw_lambda_theta, without the LDS
 * that bounds k_draw's occupancy.  p_j is uniform over a workgroup, so the
 * unrolled loops' bounds checks are uniform branches. */
template <int PMAX>
__device__ __forceinline__ void draw_reg_body(long N, int m, int max_p, const unsigned *p,
                                              const int *off, long P, const double *m_t,
                                              const double *U, const double *n_t,
                                              const double *s_t, uint64_t seed, uint32_t epoch,
                                              double *lam, double *theta, int j, long bx,
                                              long gx) {
  const int pj = (int)p[j];
  const double *Uj = U + (size_t)j * max_p * max_p, *mj = m_t + (size_t)j * max_p;
  const double nj = n_t[j], sj = s_t[j];
  for (long i = bx * TB + threadIdx.x; i < N; i += gx * TB) {
    const double l = sg_gamma(0.5 * nj, 0.5 * nj * sj, seed, epoch, SG_PURPOSE_LAMBDA,
                              (uint32_t)i, (uint32_t)j);
    const double scale = __builtin_sqrt(sj * l);
    double acc[PMAX];
#pragma unroll
    for (int k = 0; k < PMAX; ++k)
      acc[k] = 0.0;
#pragma unroll
    for (int q = 0; q < PMAX; q += 2) {
      if (q < pj) {
        double z0, z1;
        sg_normal_pair(sg_rng_block(seed, epoch, SG_PURPOSE_THETA, (uint32_t)i, (uint32_t)j,
                                    (uint32_t)(q >> 1)),
                       &z0, &z1);
#pragma unroll
        for (int k = q; k < PMAX; ++k) {
          if (k < pj) {
            double t = acc[k] + Uj[k * max_p + q] * z0;
            if (q + 1 <= k)
              t = t + Uj[k * max_p + q + 1] * z1;
            acc[k] = t;
          }
        }
      }
    }
    lam[(size_t)i * m + j] = l;
    double *out = theta + (size_t)i * P + off[j];
#pragma unroll
    for (int k = 0; k < PMAX; ++k)
      if (k < pj)
        out[k] = mj[k] + acc[k] / scale;
  }
}

template <int PMAX>
__global__ __launch_bounds__(TB) void k_draw_reg(long N, int m, int max_p,
                                                 const unsigned *p, const int *off,
                                                 long P, const double *m_t,
                                                 const double *U, const double *n_t,
                                                 const double *s_t, uint64_t seed,
                                                 uint32_t epoch_arg, const uint32_t *epoch_dev,
                                                 double *lam, double *theta) {
  const uint32_t epoch = epoch_dev ? *epoch_dev : epoch_arg;
  draw_reg_body<PMAX>(N, m, max_p, p, off, P, m_t, U, n_t, s_t, seed, epoch, lam, theta,
                      (int)blockIdx.y, (long)blockIdx.x, (long)gridDim.x);
}

/* k_draw_reg over the series of several models: blockIdx.y = series of
 * all models (ser_model: its model), blockIdx.x = blocks of TB draws
 * (models with fewer draws leave the high blocks idle); the epoch of model
 * i is epochs[i].  PMAX covers the largest max_p; the draws are
 * bit-identical to the single-model launch (same operations per value). */
template <int PMAX>
__global__ __launch_bounds__(TB) void k_draw_reg_grp(const VbModel *__restrict__ tab,
                                                     const int *__restrict__ ser_model,
                                                     const uint32_t *__restrict__ epochs) {
  const int mi = ser_model[blockIdx.y];
  const VbModel &md = tab[mi];
  draw_reg_body<PMAX>(md.N, md.m, md.max_p, md.p, md.off, md.P, md.m_t, md.U, md.n_t, md.s_t,
                      md.seed, epochs[mi], md.lam, md.theta, (int)blockIdx.y - md.ser0,
                      (long)blockIdx.x, (long)gridDim.x);
}

/* rhs_ij = nu_ij / sqrt(lambda_ij) + sum_{k: sp_jk external} x_jk theta_ijk
 * (kernel_functions.cuh:1155-1178; theta_k = m_k beyond p_j). */
__global__ __launch_bounds__(TB) void k_forecast_rhs(
    long N, int m, int max_p, const unsigned *p, const unsigned *sp,
    const int *off, long P, const double *x, const double *m_t,
    const double *theta, const double *lam, uint64_t seed, uint32_t epoch,
    double *rhs) {
  const int j = blockIdx.y;
  const int pj = (int)p[j];
  const unsigned mm = (unsigned)m * (unsigned)m;
  for (long i = (long)blockIdx.x * TB + threadIdx.x; i < N; i += (long)gridDim.x * TB) {
    const double *th = theta + (size_t)i * P + off[j];
    double nu = sg_normal_k(seed, epoch, SG_PURPOSE_NU, (uint32_t)i, (uint32_t)j, 0u);
    nu = nu / __builtin_sqrt(lam[(size_t)i * m + j]);
    for (int k = 0; k < max_p; ++k) {
      if (sp[(size_t)j * max_p + k] >= mm) {
        const double t = (k < pj) ? th[k] : m_t[(size_t)j * max_p + k];
        nu = nu + x[(size_t)j * max_p + k] * t;
      }
    }
    rhs[(size_t)i * m + j] = nu;
  }
}

/* one (draw, series) of k_evolve: lambda *= eta (beta draw),
 * omega_k = z_k / sqrt(c lambda) for k < p_j, theta = G theta (optional),
 * theta += chol(W)^T omega (SGDLM.cu:679-762 with the intended
 * scale_omegas indexing). */
__device__ void evolve_one(long i, int j, int m, int max_p, const unsigned *p, const int *off,
                           long P, const double *U_W, const double *r_t, const double *c_t,
                           const double *beta, const double *G, const double *m_t,
                           uint64_t seed, uint32_t epoch, double *lam, double *theta,
                           double *om, double *ot) {
  const int pj = (int)p[j];
  double *th = theta + (size_t)i * P + off[j];
  const double *U = U_W + (size_t)j * max_p * max_p;
  double *lp = lam + (size_t)i * m + j;
  const double eta = sg_beta_eta(r_t[j], beta[j], seed, epoch, (uint32_t)i, (uint32_t)j);
  const double l = *lp * eta;
  *lp = l;
  const double sc = __builtin_sqrt(c_t[j] * l);
  for (int q = 0; q < pj; q += 2) {
    double z0, z1;
    sg_normal_pair(sg_rng_block(seed, epoch, SG_PURPOSE_OMEGA, (uint32_t)i,
                                (uint32_t)j, (uint32_t)(q >> 1)),
                   &z0, &z1);
    om[q * TB] = z0 / sc;
    if (q + 1 < pj)
      om[(q + 1) * TB] = z1 / sc;
  }
  if (G != nullptr) {
    const double *Gj = G + (size_t)j * max_p * max_p;
    const double *mj = m_t + (size_t)j * max_p;
    for (int k = 0; k < pj; ++k)
      ot[k * TB] = th[k];
    for (int k = 0; k < pj; ++k) {
      double acc = 0.0;
      for (int n = 0; n < max_p; ++n) {
        const double v = (n < pj) ? ot[n * TB] : mj[n];
        acc += Gj[k + n * max_p] * v;
      }
      th[k] = acc;
    }
  }
  for (int k = 0; k < pj; ++k) {
    double acc = 0.0;
    for (int n = 0; n <= k; ++n)
      acc = acc + U[k * max_p + n] * om[n * TB];
    th[k] = acc + th[k];
  }
}

/* k_evolve with omega and the evolved coefficients in registers
 * (max_p <= PMAX): the same operations in the same order as evolve_one
 * (G theta accumulated over n for every k, then chol(W)^T omega), with the
 * old coefficients read once each from the draw's theta row. */
template <int PMAX>
__global__ __launch_bounds__(TB) void k_evolve_reg(
    long N, int m, int max_p, const unsigned *p, const int *off, long P,
    const double *U_W, const double *r_t, const double *c_t,
    const double *beta, const double *G, const double *m_t, uint64_t seed,
    uint32_t epoch, double *lam, double *theta) {
  const int j = blockIdx.y;
  const int pj = (int)p[j];
  const double *U = U_W + (size_t)j * max_p * max_p;
  const double *Gj = G ? G + (size_t)j * max_p * max_p : nullptr;
  const double *mj = m_t + (size_t)j * max_p;
  for (long i = (long)blockIdx.x * TB + threadIdx.x; i < N; i += (long)gridDim.x * TB) {
    double *th = theta + (size_t)i * P + off[j];
    double *lp = lam + (size_t)i * m + j;
    const double eta = sg_beta_eta(r_t[j], beta[j], seed, epoch, (uint32_t)i, (uint32_t)j);
    const double l = *lp * eta;
    *lp = l;
    const double sc = __builtin_sqrt(c_t[j] * l);
    double om[PMAX], nt[PMAX];
#pragma unroll
    for (int q = 0; q < PMAX; q += 2) {
      if (q < pj) {
        double z0, z1;
        sg_normal_pair(sg_rng_block(seed, epoch, SG_PURPOSE_OMEGA, (uint32_t)i,
                                    (uint32_t)j, (uint32_t)(q >> 1)),
                       &z0, &z1);
        om[q] = z0 / sc;
        if (q + 1 < PMAX)
          om[q + 1] = (q + 1 < pj) ? z1 / sc : 0.0;
      }
    }
    if (Gj != nullptr) {
#pragma unroll
      for (int k = 0; k < PMAX; ++k)
        nt[k] = 0.0;
#pragma unroll
      for (int n = 0; n < PMAX; ++n) {
        if (n < max_p) {
          const double v = (n < pj) ? th[n] : mj[n];
#pragma unroll
          for (int k = 0; k < PMAX; ++k)
            if (k < pj)
              nt[k] += Gj[k + n * max_p] * v;
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < PMAX; ++k)
        if (k < pj)
          nt[k] = th[k];
    }
#pragma unroll
    for (int k = 0; k < PMAX; ++k) {
      if (k < pj) {
        double acc = 0.0;
#pragma unroll
        for (int n = 0; n <= k; ++n)
          acc = acc + U[k * max_p + n] * om[n];
        th[k] = acc + nt[k];
      }
    }
  }
}

__global__ __launch_bounds__(TB) void k_evolve(
    long N, int m, int max_p, const unsigned *p, const int *off, long P,
    const double *U_W, const double *r_t, const double *c_t,
    const double *beta, const double *G, const double *m_t, uint64_t seed,
    uint32_t epoch, double *lam, double *theta) {
  extern __shared__ double sbuf[]; /* omega [max_p][TB], old theta [max_p][TB] */
  const int j = blockIdx.y;
  for (long i = (long)blockIdx.x * TB + threadIdx.x; i < N; i += (long)gridDim.x * TB)
    evolve_one(i, j, m, max_p, p, off, P, U_W, r_t, c_t, beta, G, m_t, seed, epoch, lam, theta,
               sbuf + threadIdx.x, sbuf + (size_t)max_p * TB + threadIdx.x);
}

__global__ __launch_bounds__(TB) void k_unpack_theta(long N, int m, int max_p,
                                                     const unsigned *p,
                                                     const int *off, long P,
                                                     const double *theta,
                                                     const double *m_t,
                                                     double *out) {
  const int j = blockIdx.y;
  const int pj = (int)p[j];
  for (long i = (long)blockIdx.x * TB + threadIdx.x; i < N; i += (long)gridDim.x * TB) {
    const double *th = theta + (size_t)i * P + off[j];
    double *o = out + ((size_t)i * m + j) * max_p;
    for (int k = 0; k < max_p; ++k)
      o[k] = (k < pj) ? th[k] : m_t[(size_t)j * max_p + k];
  }
}

} // namespace

void launch_cholesky(const Structure &S, const double *C, double *U,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_cholesky, dim3(S.m), dim3(64),
                     (size_t)S.max_p * S.max_p * sizeof(double), st, S.m, S.max_p, S.p, C, U);
}

void launch_draw(const Structure &S, long N, const double *m_t,
                 const double *U, const double *n_t, const double *s_t,
                 uint64_t seed, uint32_t epoch, double *lam, double *theta,
                 hipStream_t st, const uint32_t *epoch_dev) {
  if (S.max_p <= 32) {
    auto kern = S.max_p <= 8 ? k_draw_reg<8> : S.max_p <= 16 ? k_draw_reg<16>
                : S.max_p <= 24 ? k_draw_reg<24> : k_draw_reg<32>;
    hipLaunchKernelGGL(kern, grid_nm(N, S.m), dim3(TB), 0, st, N, S.m, S.max_p, S.p, S.off,
                       S.P, m_t, U, n_t, s_t, seed, epoch, epoch_dev, lam, theta);
    return;
  }
  const size_t lds = (size_t)S.max_p * TS * sizeof(double);
  hipLaunchKernelGGL(k_draw, grid_nm(N, S.m), dim3(TB), lds, st, N, S.m,
                     S.max_p, S.p, S.off, S.P, m_t, U, n_t, s_t, seed, epoch, epoch_dev,
                     lam, theta);
}

void launch_forecast_rhs(const Structure &S, long N, const double *x,
                         const double *m_t, const double *theta,
                         const double *lam, uint64_t seed, uint32_t epoch,
                         double *rhs, hipStream_t st) {
  hipLaunchKernelGGL(k_forecast_rhs, grid_nm(N, S.m), dim3(TB), 0, st, N, S.m,
                     S.max_p, S.p, S.sp, S.off, S.P, x, m_t, theta, lam, seed,
                     epoch, rhs);
}

void launch_evolve(const Structure &S, long N, const double *U_W,
                   const double *r_t, const double *c_t, const double *beta,
                   const double *G, const double *m_t, uint64_t seed,
                   uint32_t epoch, double *lam, double *theta,
                   hipStream_t st) {
  if (S.max_p <= 32) {
    auto kern = S.max_p <= 8 ? k_evolve_reg<8> : S.max_p <= 16 ? k_evolve_reg<16>
                : S.max_p <= 24 ? k_evolve_reg<24> : k_evolve_reg<32>;
    hipLaunchKernelGGL(kern, grid_nm(N, S.m), dim3(TB), 0, st, N, S.m, S.max_p, S.p, S.off,
                       S.P, U_W, r_t, c_t, beta, G, m_t, seed, epoch, lam, theta);
    return;
  }
  const size_t lds = 2 * (size_t)S.max_p * TB * sizeof(double);
  hipLaunchKernelGGL(k_evolve, grid_nm(N, S.m), dim3(TB), lds, st, N, S.m,
                     S.max_p, S.p, S.off, S.P, U_W, r_t, c_t, beta, G, m_t,
                     seed, epoch, lam, theta);
}

void launch_cholesky_grp(const VbModel *tab, const int *ser_model, int n_series, int max_p,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_cholesky_grp, dim3(n_series), dim3(64),
                     (size_t)max_p * max_p * sizeof(double), st, tab, ser_model);
}

int launch_draw_grp(const VbModel *tab, const int *ser_model, const uint32_t *epochs,
                    int n_series, long max_N, int max_p, hipStream_t st) {
  if (max_p > 32 || n_series > 65535)
    return SGDLM_ERR_ARG;
  auto kern = max_p <= 8 ? k_draw_reg_grp<8> : max_p <= 16 ? k_draw_reg_grp<16>
              : max_p <= 24 ? k_draw_reg_grp<24> : k_draw_reg_grp<32>;
  hipLaunchKernelGGL(kern, grid_nm(max_N, n_series), dim3(TB), 0, st, tab, ser_model, epochs);
  return hipGetLastError() == hipSuccess ? SGDLM_OK : SGDLM_ERR_DEVICE;
}

void launch_unpack_theta(const Structure &S, long N, const double *theta,
                         const double *m_t, double *theta_full,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_unpack_theta, grid_nm(N, S.m), dim3(TB), 0, st, N, S.m,
                     S.max_p, S.p, S.off, S.P, theta, m_t, theta_full);
}

} // namespace sgdlm
