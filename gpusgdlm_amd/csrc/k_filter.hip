/*
 * k_filter.hip -- per-series conjugate normal-gamma filter update and the
 * discount step, one wavefront per series with the max_p x max_p blocks
 * staged in LDS.
 *
 * Replaces the 12-launch chain of SGDLM::compute_posterior
 * (rSGDLM/src/SGDLM/SGDLM.cu:10-113: cut-offs, A = R F, Q = c + F'A,
 * A /= Q, e = y - F'a, n += 1, s = s (n-1+e^2/Q)/n, m += A e,
 * C = (R - Q A A') (n-1+e^2/Q)/n) and compute_one_step_ahead_prior
 * (SGDLM.cu:115-179: r = beta n, R = C / Delta and W = R - C on the p x p
 * block, optional a = G m, R = G R G' with cut-offs) by one fused kernel
 * each.  All arithmetic is FP64 in the operation order of the reference
 * kernels (batchedCutOffMatrix's "row >= p and col >= p" rule included).
 */
#include <hip/hip_runtime.h>

#include "sgdlm_internal.h"

namespace sgdlm {
namespace {

constexpr int WAVE = 64;

__global__ __launch_bounds__(WAVE) void k_posterior(int m, int max_p,
                                                    const unsigned *p,
                                                    double *m_t, double *C_t,
                                                    double *n_t, double *s_t,
                                                    const double *y,
                                                    const double *F) {
  extern __shared__ double sh[];
  const int j = blockIdx.x;
  const int t = threadIdx.x;
  const int mp = max_p, mp2 = max_p * max_p;
  const int P = (int)p[j];
  double *C = sh;          /* mp2 */
  double *Fs = C + mp2;    /* mp  */
  double *A = Fs + mp;     /* mp  */
  double *mt = A + mp;     /* mp  */
  double *sc = mt + mp;    /* 4: Q, e, factor, sqrt(Q) */
  double *Cg = C_t + (size_t)j * mp2;
  double *mg = m_t + (size_t)j * mp;

  for (int q = t; q < mp2; q += WAVE) {
    const int r = q % mp, c = q / mp;
    const double v = Cg[q];
    C[q] = (r >= P && c >= P) ? 0.0 : v;
  }
  for (int r = t; r < mp; r += WAVE) {
    Fs[r] = F[(size_t)j * mp + r];
    mt[r] = (r >= P) ? 0.0 : mg[r];
  }
  __syncthreads();
  for (int r = t; r < mp; r += WAVE) {
    double acc = 0.0;
    for (int c = 0; c < mp; ++c)
      acc += C[r + c * mp] * Fs[c];
    A[r] = acc;
  }
  __syncthreads();
  if (t == 0) {
    double fa = 0.0, fm = 0.0;
    for (int r = 0; r < mp; ++r)
      fa += Fs[r] * A[r];
    const double Q = s_t[j] + fa;
    for (int r = 0; r < mp; ++r)
      fm += Fs[r] * mt[r];
    const double e = y[j] - fm;
    const double n = n_t[j] + 1.0;
    n_t[j] = n;
    s_t[j] = s_t[j] * (n - 1.0 + e * e / Q) / n;
    sc[0] = Q;
    sc[1] = e;
    sc[2] = ((n - 1.0) + e * e / Q) / n;
    sc[3] = __builtin_sqrt(Q);
  }
  __syncthreads();
  const double Q = sc[0], e = sc[1], factor = sc[2], sq = sc[3];
  for (int r = t; r < mp; r += WAVE) {
    const double a = A[r] / Q;
    mg[r] = mt[r] + e * a;
    A[r] = a * sq;
  }
  __syncthreads();
  for (int q = t; q < mp2; q += WAVE) {
    const int r = q % mp, c = q / mp;
    double v = C[q] - A[r] * A[c];
    v *= factor;
    Cg[q] = v;
  }
}

__global__ __launch_bounds__(WAVE) void k_prior(int m, int max_p,
                                                const unsigned *p, double *m_t,
                                                double *C_t, double *n_t,
                                                double *W_t, const double *beta,
                                                const double *delta,
                                                const double *G) {
  extern __shared__ double sh[];
  const int j = blockIdx.x;
  const int t = threadIdx.x;
  const int mp = max_p, mp2 = max_p * max_p;
  const int P = (int)p[j];
  double *Cg = C_t + (size_t)j * mp2;
  double *Wg = W_t + (size_t)j * mp2;
  const double *D = delta + (size_t)j * mp2;
  if (t == 0)
    n_t[j] = n_t[j] * beta[j];
  for (int q = t; q < mp2; q += WAVE) {
    const int r = q % mp, c = q / mp;
    double w = 0.0;
    if (r < P && c < P) {
      const double cv = Cg[q];
      w = cv * (1.0 / D[q] - 1.0);
      Cg[q] = cv / D[q];
    }
    Wg[q] = w;
  }
  if (G == nullptr)
    return;
  double *Cs = sh;       /* mp2: R after discount */
  double *B = Cs + mp2;  /* mp2: G R              */
  double *mb = B + mp2;  /* mp                    */
  const double *Gj = G + (size_t)j * mp2;
  double *mg = m_t + (size_t)j * mp;
  __syncthreads();
  for (int q = t; q < mp2; q += WAVE)
    Cs[q] = Cg[q];
  for (int r = t; r < mp; r += WAVE)
    mb[r] = mg[r];
  __syncthreads();
  for (int r = t; r < mp; r += WAVE) {
    double acc = 0.0;
    for (int c = 0; c < mp; ++c)
      acc += Gj[r + c * mp] * mb[c];
    mg[r] = (r >= P) ? 0.0 : acc;
  }
  for (int q = t; q < mp2; q += WAVE) {
    const int r = q % mp, c = q / mp;
    double acc = 0.0;
    for (int k = 0; k < mp; ++k)
      acc += Gj[r + k * mp] * Cs[k + c * mp];
    B[q] = acc;
  }
  __syncthreads();
  for (int q = t; q < mp2; q += WAVE) {
    const int r = q % mp, c = q / mp;
    double acc = 0.0;
    for (int k = 0; k < mp; ++k)
      acc += B[r + k * mp] * Gj[c + k * mp];
    Cg[q] = (r >= P && c >= P) ? 0.0 : acc;
  }
}

} // namespace

void launch_posterior(const Structure &S, double *m_t, double *C_t,
                      double *n_t, double *s_t, const double *y,
                      const double *F, hipStream_t st) {
  const size_t lds = ((size_t)S.max_p * S.max_p + 3 * S.max_p + 4) * sizeof(double);
  hipLaunchKernelGGL(k_posterior, dim3(S.m), dim3(WAVE), lds, st, S.m, S.max_p,
                     S.p, m_t, C_t, n_t, s_t, y, F);
}

void launch_prior(const Structure &S, double *m_t, double *C_t, double *n_t,
                  double *W_t, const double *beta, const double *delta,
                  const double *G, hipStream_t st) {
  const size_t lds = (2 * (size_t)S.max_p * S.max_p + S.max_p) * sizeof(double);
  (void)allow_dynamic_lds((const void *)k_prior, lds);
  hipLaunchKernelGGL(k_prior, dim3(S.m), dim3(WAVE), lds, st, S.m, S.max_p,
                     S.p, m_t, C_t, n_t, W_t, beta, delta, G);
}

} // namespace sgdlm
