/*
 * k_vb.hip -- importance weights and the VB moment-match.
 *
 * Replaces, from rSGDLM/src/SGDLM/SGDLM.cu:325-503:
 *   sumVector<<<1,1>>> + scale (single-thread normalisation, Q2), by an FP64
 *     log-sum-exp over the log-determinants: w_i = exp(l_i - max) / sum,
 *     non-finite l_i -> weight 0, ESS = 1 / sum w^2 (one workgroup, fixed
 *     reduction tree, deterministic);
 *   meanScalarAndLog / compute_mean_m_t / compute_VB_vector1 /
 *     meanBatchedVecVecT (serial-over-N per-thread loops,
 *     kernel_functions.cuh:928-982,1102-1133), by one chunked pass over
 *     the packed draws: sum w*log(lambda) and, shifted by the naive mean c,
 *     sum w lambda (theta - c)(theta - c)' with sum w lambda (theta - c)
 *     and sum w lambda as an extra row / column, as a p x N . N x p
 *     contraction on FP64 MFMA (v_mfma_f64_16x16x4f64, the 4 draws of a
 *     k-step on the contraction axis); the final kernel forms the means and
 *     V = sum w lambda (theta - mbar)(theta - mbar)' from them;
 *   getrf/getri(max_p) + the two m*N gemmBatched calls for Q and
 *     compute_mean_Q_t (SGDLM.cu:404-465), by a Gauss-Jordan inverse of the
 *     p x p block and Qbar = trace(V^{-1} V), which equals
 *     sum_i w_i lambda_i (theta_i - m)' V^{-1} (theta_i - m) identically;
 *   compute_mean_n_t / compute_mean_s_t / batchedScale (Newton with the
 *     truncated psi/psi1 series, OPTIM_PRECISION 1e-6, OPTIM_MAX_ITER 50,
 *     fallback to n_t; s < 0 -> s_t; C = s V on the block;
 *     kernel_functions.cuh:1006-1085).
 * Partial sums are combined in a fixed chunk order, so results are
 * bit-reproducible run to run.
 */
#include <hip/hip_runtime.h>

#include <math.h>

#include "sgdlm_internal.h"

#define OPTIM_PRECISION 1e-6 /* rSGDLM/src/Makevars.in:1 */
#define OPTIM_MAX_ITER 50    /* rSGDLM/src/Makevars.in:1 */

namespace sgdlm {
namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int WTB = 1024;

__device__ inline bool finite_d(double v) { return v == v && fabs(v) <= 1.79769313486231570815e308; }

__device__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
    v += __shfl_xor(v, o);
  return v;
}
__device__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
    v = fmax(v, __shfl_xor(v, o));
  return v;
}

/* one workgroup: log-sum-exp weights */
__device__ __forceinline__ void weights_body(long N, const double *ld, double *w,
                                             double *stats) {
  __shared__ double red[WTB / 64];
  __shared__ double bc;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double mx = -INFINITY;
  for (long i = tid; i < N; i += WTB) {
    const double v = ld[i];
    if (finite_d(v) && v > mx)
      mx = v;
  }
  mx = wave_max(mx);
  if (lane == 0)
    red[wv] = mx;
  __syncthreads();
  if (tid == 0) {
    double t = red[0];
    for (int q = 1; q < WTB / 64; ++q)
      t = fmax(t, red[q]);
    bc = t;
  }
  __syncthreads();
  mx = bc;
  double s = 0.0;
  for (long i = tid; i < N; i += WTB) {
    const double v = ld[i];
    if (finite_d(v))
      s += exp(v - mx);
  }
  s = wave_sum(s);
  __syncthreads();
  if (lane == 0)
    red[wv] = s;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int q = 0; q < WTB / 64; ++q)
      t += red[q];
    bc = t;
  }
  __syncthreads();
  s = bc;
  double s2 = 0.0, nz = 0.0;
  for (long i = tid; i < N; i += WTB) {
    const double v = ld[i];
    const double wi = finite_d(v) ? exp(v - mx) / s : (s > 0.0 ? 0.0 : NAN);
    w[i] = wi;
    s2 += wi * wi;
    nz += (wi == 0.0) ? 1.0 : 0.0;
  }
  s2 = wave_sum(s2);
  nz = wave_sum(nz);
  __syncthreads();
  if (lane == 0)
    red[wv] = s2;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int q = 0; q < WTB / 64; ++q)
      t += red[q];
    stats[2] = t;
    stats[0] = mx;
    stats[1] = s;
  }
  __syncthreads();
  if (lane == 0)
    red[wv] = nz;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int q = 0; q < WTB / 64; ++q)
      t += red[q];
    stats[3] = t;
  }
}

__global__ __launch_bounds__(WTB) void k_weights(long N, const double *ld,
                                                 double *w, double *stats) {
  weights_body(N, ld, w, stats);
}

/* k_weights of several models: one workgroup per model */
__global__ __launch_bounds__(WTB) void k_weights_grp(const VbModel *__restrict__ tab) {
  const VbModel &md = tab[blockIdx.x];
  weights_body(md.N, md.logdet, md.w, md.stats);
}

/* One pass over the draws of chunk c for series j, on FP64 MFMA
 * (v_mfma_f64_16x16x4f64, p x N . N x p).  With x_k = sqrt(w lambda)
 * (theta_k - c_k) for k < p_j and x_{p_j} = sqrt(w lambda), X = sum x x^T
 * holds sum w lambda (theta - c)(theta - c)^T, sum w lambda (theta - c)
 * (column p_j) and sum w lambda (entry (p_j, p_j)); the shift c is the
 * naive posterior mean the draws are centred on, so forming
 * sum w lambda (theta - mbar)(theta - mbar)^T from X later cancels little.
 * Also sum w log lambda (part1).  NBLK = number of 16-wide coefficient
 * blocks, p_j + 1 <= 16 NBLK.  X goes to part2[(c m + j) (max_p+1)^2]. */
template <int NBLK>
__device__ __forceinline__ void mom_body(long N, int m, int max_p, int chunks,
                                         const unsigned *p, const int *off, long P,
                                         const double *w, const double *lam,
                                         const double *theta, const double *m_t,
                                         double *part1, double *part2, int j, int c,
                                         double (*tiles)[256], double *red) {
  constexpr int NT = NBLK * (NBLK + 1) / 2;
  const long i0 = N * c / chunks, i1 = N * (c + 1) / chunks;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int pj = (int)p[j], mq = max_p + 1;
  double cb[NBLK];
#pragma unroll
  for (int b = 0; b < NBLK; ++b) {
    const int k = 16 * b + l16;
    cb[b] = (k < pj) ? m_t[(size_t)j * max_p + k] : 0.0;
  }
  f64x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t)
    acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
  double wl = 0.0;
  const double *th = theta + off[j];
  /* each wavefront takes 64 draws at a time: lane l computes sqrt(w lambda)
   * and w log lambda of draw base + l once, then 16 MFMA steps of 4 draws
   * each read it back by lane shuffle */
  for (long base = i0 + 64 * wave; base < i1; base += 256) {
    const long id = base + lane;
    double scl = 0.0;
    if (id < i1) {
      const double wi = w[id], li = lam[(size_t)id * m + j];
      scl = sqrt(wi * li);
      wl += wi * log(li);
    }
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
      const long i = base + 4 * q + g;
      const double sc = __shfl(scl, 4 * q + g);
      double x[NBLK];
#pragma unroll
      for (int b = 0; b < NBLK; ++b) {
        const int k = 16 * b + l16;
        x[b] = (i < i1 && k < pj) ? sc * (th[(size_t)i * P + k] - cb[b]) : (k == pj ? sc : 0.0);
      }
      int t = 0;
#pragma unroll
      for (int A = 0; A < NBLK; ++A)
#pragma unroll
        for (int B = A; B < NBLK; ++B) {
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[A], x[B], acc[t], 0, 0, 0);
          ++t;
        }
    }
  }
  wl = wave_sum(wl);
  if (lane == 0)
    red[wave] = wl;
  /* sum the 4 waves' tiles in a fixed order */
  for (int wv = 0; wv < 4; ++wv) {
    if (wave == wv) {
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          /* v_mfma_f64_16x16x4f64 accumulator: register r of lane group g
           * holds row 4r + g, column lane % 16 */
          const int e = (4 * r + g) * 16 + l16;
          tiles[t][e] = (wv == 0) ? acc[t][r] : tiles[t][e] + acc[t][r];
        }
    }
    __syncthreads();
  }
  if (tid == 0)
    part1[(size_t)c * m + j] = ((red[0] + red[1]) + red[2]) + red[3];
  double *out = part2 + ((size_t)c * m + j) * mq * mq;
  int t = 0;
  for (int A = 0; A < NBLK; ++A)
    for (int B = A; B < NBLK; ++B) {
      const int r = 16 * A + tid / 16, col = 16 * B + (tid & 15);
      if (r <= pj && col <= pj) {
        out[r + (size_t)col * mq] = tiles[t][tid];
        out[col + (size_t)r * mq] = tiles[t][tid];
      }
      ++t;
    }
}

/* mom_body with the fewest 16-wide blocks that hold series j's p_j + 1
 * columns (at most NBLK): the blocks beyond them would only accumulate
 * exact zeros and write nothing, so the moments are the same bit for bit
 * and a series with p_j < 16 runs one MFMA per step instead of three (or
 * more).  nb is uniform over the workgroup. */
template <int NBLK>
__device__ __forceinline__ void mom_series(int nb, long N, int m, int max_p, int chunks,
                                           const unsigned *p, const int *off, long P,
                                           const double *w, const double *lam,
                                           const double *theta, const double *m_t,
                                           double *part1, double *part2, int j, int c,
                                           double (*tiles)[256], double *red) {
  if constexpr (NBLK > 1) {
    if (nb < NBLK) {
      mom_series<NBLK - 1>(nb, N, m, max_p, chunks, p, off, P, w, lam, theta, m_t, part1, part2,
                           j, c, tiles, red);
      return;
    }
  }
  mom_body<NBLK>(N, m, max_p, chunks, p, off, P, w, lam, theta, m_t, part1, part2, j, c, tiles,
                 red);
}

template <int NBLK>
__global__ __launch_bounds__(256) void k_mom(long N, int m, int max_p, int chunks,
                                             const unsigned *p, const int *off, long P,
                                             const double *w, const double *lam,
                                             const double *theta, const double *m_t,
                                             double *part1, double *part2) {
  __shared__ double tiles[NBLK * (NBLK + 1) / 2][256];
  __shared__ double red[4];
  const int j = (int)blockIdx.y;
  const int nb = ((int)p[j] + 1 + 15) / 16;
  mom_series<NBLK>(nb, N, m, max_p, chunks, p, off, P, w, lam, theta, m_t, part1, part2, j,
                   (int)blockIdx.x, tiles, red);
}

/* k_mom of several models: blockIdx.y = series of all models, blockIdx.x =
 * chunk (models with fewer chunks leave the high ones idle).  NBLK covers
 * the largest max_p; the extra blocks only add exact zeros, so the moments
 * equal the single-model launch's bit for bit. */
template <int NBLK>
__global__ __launch_bounds__(256) void k_mom_grp(const VbModel *__restrict__ tab,
                                                 const int *__restrict__ ser_model) {
  __shared__ double tiles[NBLK * (NBLK + 1) / 2][256];
  __shared__ double red[4];
  const VbModel &md = tab[ser_model[blockIdx.y]];
  if ((int)blockIdx.x >= md.chunks)
    return;
  const int j = (int)blockIdx.y - md.ser0;
  const int nb = ((int)md.p[j] + 1 + 15) / 16;
  mom_series<NBLK>(nb, md.N, md.m, md.max_p, md.chunks, md.p, md.off, md.P, md.w, md.lam,
                   md.theta, md.m_t, md.part1, md.part2, j, (int)blockIdx.x, tiles, red);
}

__device__ double psi_series(double x) {
  return log(x) - 1.0 / (2.0 * x) - 1.0 / (12.0 * x * x) +
         1.0 / (120.0 * x * x * x * x) - 1.0 / (252.0 * x * x * x * x * x * x) +
         1.0 / (240.0 * x * x * x * x * x * x * x * x) -
         5.0 / (660.0 * x * x * x * x * x * x * x * x * x * x) +
         691.0 / (32760.0 * x * x * x * x * x * x * x * x * x * x * x * x) -
         1.0 / (12.0 * x * x * x * x * x * x * x * x * x * x * x * x * x * x);
}
__device__ double psi1_series(double x) {
  return 1.0 / x + 1.0 / (2.0 * x * x) + 1.0 / (6.0 * x * x * x) -
         1.0 / (30.0 * x * x * x * x * x) +
         1.0 / (42.0 * x * x * x * x * x * x * x) -
         1.0 / (30.0 * x * x * x * x * x * x * x * x * x);
}

/* final per-series moment-match; one wavefront per series */
__device__ __forceinline__ void vb_final_body(
    int m, int max_p, int chunks, const unsigned *p, const double *m_t, const double *n_t,
    const double *s_t, const double *part1, const double *part2,
    double *mean_lambda, double *mean_log_lambda, double *mean_m,
    double *mean_C, double *mean_n, double *mean_s, double *mean_Q,
    int *newton_iter, int j) {
  extern __shared__ double sh[];
  const int t = threadIdx.x;
  const int pj = (int)p[j], mp = max_p, mp2 = max_p * max_p, mq = max_p + 1;
  double *V = sh;                     /* pj x pj (ld pj) */
  double *Aug = V + (size_t)pj * pj;  /* pj x 2pj        */
  double *D = Aug + (size_t)2 * pj * pj; /* pj: sum w lambda (theta - c) */
  __shared__ double sc[4];
  __shared__ int spiv;
  /* the chunks' partial sums, added in chunk order */
  auto xsum = [&](int r, int col) {
    double v = 0.0;
    for (int ch = 0; ch < chunks; ++ch)
      v += part2[((size_t)ch * m + j) * mq * mq + r + (size_t)col * mq];
    return v;
  };
  if (t == 0) {
    double l = 0.0;
    for (int ch = 0; ch < chunks; ++ch)
      l += part1[(size_t)ch * m + j];
    sc[0] = xsum(pj, pj);
    sc[1] = l;
  }
  for (int k = t; k < pj; k += 64)
    D[k] = xsum(k, pj);
  __syncthreads();
  const double lb = sc[0];
  for (int k = t; k < mp; k += 64)
    mean_m[(size_t)j * mp + k] = (k < pj) ? m_t[(size_t)j * mp + k] + D[k] / lb : 0.0;
  for (int e = t; e < pj * pj; e += 64) {
    const int r = e % pj, c = e / pj;
    V[e] = xsum(r, c) - D[r] * D[c] / lb;
  }
  __syncthreads();
  for (int e = t; e < pj * 2 * pj; e += 64) {
    const int r = e / (2 * pj), c = e % (2 * pj);
    Aug[e] = (c < pj) ? V[r + c * pj] : ((c - pj == r) ? 1.0 : 0.0);
  }
  __syncthreads();
  /* Gauss-Jordan with partial pivoting */
  for (int k = 0; k < pj; ++k) {
    if (t == 0) {
      int piv = k;
      double best = fabs(Aug[(size_t)k * 2 * pj + k]);
      for (int r = k + 1; r < pj; ++r) {
        const double v = fabs(Aug[(size_t)r * 2 * pj + k]);
        if (v > best) {
          best = v;
          piv = r;
        }
      }
      spiv = piv;
    }
    __syncthreads();
    const int piv = spiv;
    if (piv != k)
      for (int c = t; c < 2 * pj; c += 64) {
        const double tmp = Aug[(size_t)k * 2 * pj + c];
        Aug[(size_t)k * 2 * pj + c] = Aug[(size_t)piv * 2 * pj + c];
        Aug[(size_t)piv * 2 * pj + c] = tmp;
      }
    __syncthreads();
    const double d = Aug[(size_t)k * 2 * pj + k];
    __syncthreads();
    for (int c = t; c < 2 * pj; c += 64)
      Aug[(size_t)k * 2 * pj + c] /= d;
    __syncthreads();
    for (int r = t; r < pj; r += 64) {
      if (r != k) {
        const double f = Aug[(size_t)r * 2 * pj + k];
        for (int c = 0; c < 2 * pj; ++c)
          Aug[(size_t)r * 2 * pj + c] -= f * Aug[(size_t)k * 2 * pj + c];
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    /* Qbar = trace(V^{-1} V) */
    double Q = 0.0;
    for (int a = 0; a < pj; ++a)
      for (int b = 0; b < pj; ++b)
        Q += Aug[(size_t)a * 2 * pj + pj + b] * V[b + a * pj];
    const double lbar = sc[0], llbar = sc[1];
    const double pd = (double)pj;
    int iter = 0;
    double mn = n_t[j], fun, dfun, step = 0.0;
    do {
      mn -= step;
      fun = log(mn + pd - Q) - psi_series(mn / 2.0) - (pd - Q) / mn -
            log(2.0 * lbar) + llbar;
      dfun = 1.0 / (mn + pd - Q) - 0.5 * psi1_series(mn / 2.0) +
             (pd - Q) / (mn * mn);
      step = fun / dfun;
      if (step > mn)
        step = 0.5 * mn;
    } while (fabs(fun) > OPTIM_PRECISION && iter++ < OPTIM_MAX_ITER);
    const double nbar = (iter >= OPTIM_MAX_ITER) ? n_t[j] : mn;
    double sbar = (nbar + (double)pj - Q) / (nbar * lbar);
    if (sbar < 0.0)
      sbar = s_t[j];
    mean_lambda[j] = lbar;
    mean_log_lambda[j] = llbar;
    mean_n[j] = nbar;
    mean_s[j] = sbar;
    mean_Q[j] = Q;
    newton_iter[j] = iter;
    sc[2] = sbar;
  }
  __syncthreads();
  const double sbar = sc[2];
  for (int e = t; e < mp2; e += 64) {
    const int r = e % mp, c = e / mp;
    mean_C[(size_t)j * mp2 + e] = (r < pj && c < pj) ? sbar * V[r + c * pj] : 0.0;
  }
}

__global__ __launch_bounds__(64) void k_vb_final(
    int m, int max_p, int chunks, const unsigned *p, const double *m_t, const double *n_t,
    const double *s_t, const double *part1, const double *part2,
    double *mean_lambda, double *mean_log_lambda, double *mean_m,
    double *mean_C, double *mean_n, double *mean_s, double *mean_Q,
    int *newton_iter) {
  vb_final_body(m, max_p, chunks, p, m_t, n_t, s_t, part1, part2, mean_lambda, mean_log_lambda,
                mean_m, mean_C, mean_n, mean_s, mean_Q, newton_iter, (int)blockIdx.x);
}

/* k_vb_final over the series of several models */
__global__ __launch_bounds__(64) void k_vb_final_grp(const VbModel *__restrict__ tab,
                                                     const int *__restrict__ ser_model) {
  const VbModel &md = tab[ser_model[blockIdx.x]];
  vb_final_body(md.m, md.max_p, md.chunks, md.p, md.m_t, md.n_t, md.s_t, md.part1, md.part2,
                md.mean_lambda, md.mean_log_lambda, md.mean_m, md.mean_C, md.mean_n, md.mean_s,
                md.mean_Q, md.newton_iter, (int)blockIdx.x - md.ser0);
}

/* After it: the weight statistics of every model side by side, each
 * followed by its LU fallback count (one copy to the host); a model whose
 * count is fixed after the LU (flag_final) gets it stored first. */
__global__ void k_grp_stats(const VbModel *__restrict__ tab, int n, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 5 * n) {
    const VbModel &md = tab[i / 5];
    if (i % 5 < 4) {
      out[i] = md.stats[i % 5];
    } else {
      if (md.flag_final >= 0)
        md.lu_flags[0] = md.flag_final;
      out[i] = (double)md.lu_flags[0];
    }
  }
}

} // namespace

int vb_chunks(long N) {
  long c = N / 2048;
  if (c < 1)
    c = 1;
  if (c > 64)
    c = 64;
  return (int)c;
}

void launch_weights(long N, const double *logdet, double *w, double *stats,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_weights, dim3(1), dim3(WTB), 0, st, N, logdet, w, stats);
}

void launch_moments(const Structure &S, long N, const double *lam,
                    const double *theta, const double *m_t, const double *n_t,
                    const double *s_t, const VbBuffers &B, hipStream_t st) {
  const int chunks = B.chunks;
  const int nblk = (S.max_p + 1 + 15) / 16;
  const dim3 grid(chunks, S.m);
  switch (nblk) {
  case 1:
    hipLaunchKernelGGL(k_mom<1>, grid, dim3(256), 0, st, N, S.m, S.max_p, chunks, S.p, S.off,
                       S.P, B.w, lam, theta, m_t, B.part1, B.part2);
    break;
  case 2:
    hipLaunchKernelGGL(k_mom<2>, grid, dim3(256), 0, st, N, S.m, S.max_p, chunks, S.p, S.off,
                       S.P, B.w, lam, theta, m_t, B.part1, B.part2);
    break;
  case 3:
    hipLaunchKernelGGL(k_mom<3>, grid, dim3(256), 0, st, N, S.m, S.max_p, chunks, S.p, S.off,
                       S.P, B.w, lam, theta, m_t, B.part1, B.part2);
    break;
  case 4:
    hipLaunchKernelGGL(k_mom<4>, grid, dim3(256), 0, st, N, S.m, S.max_p, chunks, S.p, S.off,
                       S.P, B.w, lam, theta, m_t, B.part1, B.part2);
    break;
  default:
    hipLaunchKernelGGL(k_mom<5>, grid, dim3(256), 0, st, N, S.m, S.max_p, chunks, S.p, S.off,
                       S.P, B.w, lam, theta, m_t, B.part1, B.part2);
    break;
  }
  const size_t lds = (3 * (size_t)S.max_p * S.max_p + S.max_p) * sizeof(double);
  (void)allow_dynamic_lds((const void *)k_vb_final, lds);
  hipLaunchKernelGGL(k_vb_final, dim3(S.m), dim3(64), lds, st, S.m, S.max_p,
                     chunks, S.p, m_t, n_t, s_t, B.part1, B.part2, B.mean_lambda,
                     B.mean_log_lambda, B.mean_m, B.mean_C, B.mean_n,
                     B.mean_s, B.mean_Q, B.newton_iter);
}

void launch_weights_grp(const VbModel *tab, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_weights_grp, dim3(n), dim3(WTB), 0, st, tab);
}

int launch_moments_grp(const VbModel *tab, const int *ser_model, int n_series, int max_chunks,
                       int max_p, hipStream_t st) {
  if (n_series > 65535)
    return SGDLM_ERR_ARG;
  const int nblk = (max_p + 1 + 15) / 16;
  const dim3 grid(max_chunks, n_series);
  switch (nblk) {
  case 1:
    hipLaunchKernelGGL(k_mom_grp<1>, grid, dim3(256), 0, st, tab, ser_model);
    break;
  case 2:
    hipLaunchKernelGGL(k_mom_grp<2>, grid, dim3(256), 0, st, tab, ser_model);
    break;
  case 3:
    hipLaunchKernelGGL(k_mom_grp<3>, grid, dim3(256), 0, st, tab, ser_model);
    break;
  case 4:
    hipLaunchKernelGGL(k_mom_grp<4>, grid, dim3(256), 0, st, tab, ser_model);
    break;
  default:
    hipLaunchKernelGGL(k_mom_grp<5>, grid, dim3(256), 0, st, tab, ser_model);
    break;
  }
  const size_t lds = (3 * (size_t)max_p * max_p + max_p) * sizeof(double);
  (void)allow_dynamic_lds((const void *)k_vb_final_grp, lds);
  hipLaunchKernelGGL(k_vb_final_grp, dim3(n_series), dim3(64), lds, st, tab, ser_model);
  return hipGetLastError() == hipSuccess ? SGDLM_OK : SGDLM_ERR_DEVICE;
}

void launch_grp_stats(const VbModel *tab, int n, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_grp_stats, dim3((5 * n + 255) / 256), dim3(256), 0, st, tab, n, out);
}

} // namespace sgdlm
