/*
 * sgdlm_oracle.c -- FP64 CPU restatement of the rSGDLM hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU
 * baseline of bench.py as the checker; never by the product.
 *
 * Each function cites the reference file:line it restates.  The random draws
 * go through the shared deterministic numerics (Philox + samplers) of
 * gpusgdlm_amd/csrc/sgdlm_numerics.h so that the GPU and this oracle consume
 * identical random numbers; those primitives are pinned independently by the
 * tests (Philox known-answer vectors from rocRAND's host engine, elementary
 * functions against libm, sampler distributions against scipy).  Everything
 * else (filter recurrences, LU, weights, moments, Newton moment-match,
 * forecast solve) is written here independently of the GPU code.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include "sgdlm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../gpusgdlm_amd/csrc/sgdlm_numerics.h"

#define OPTIM_PRECISION 1e-6 /* rSGDLM/src/Makevars.in:1 */
#define OPTIM_MAX_ITER 50    /* rSGDLM/src/Makevars.in:1 */

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0)
    omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* ---------------------------------------------------------------------- */
/* a1: SGDLM.cu:10-113 (kernels batchedCutOffVector/Matrix :85-110,       */
/* batchedScale :140-162, computeSt :479-487, batchedComputeMt :381-390,  */
/* batchedScaleCt :392-401, four gemmBatched calls).                       */
/* ---------------------------------------------------------------------- */
void sgo_compute_posterior(int m, int max_p, const unsigned *p, double *m_t,
                           double *C_t, double *n_t, double *s_t,
                           const double *y_t, const double *F_t) {
  int j, r, c;
  const int mp2 = max_p * max_p;
  double *A = (double *)malloc(sizeof(double) * (size_t)max_p);
  for (j = 0; j < m; ++j) {
    const int P = (int)p[j];
    double *mt = m_t + (size_t)j * max_p;
    double *C = C_t + (size_t)j * mp2;
    const double *F = F_t + (size_t)j * max_p;
    double Q, e, fa, fm, factor;
    for (r = P; r < max_p; ++r)
      mt[r] = 0.0;
    for (c = 0; c < max_p; ++c)
      for (r = 0; r < max_p; ++r)
        if (r >= P && c >= P)
          C[r + c * max_p] = 0.0;
    /* A = R F */
    for (r = 0; r < max_p; ++r) {
      double acc = 0.0;
      for (c = 0; c < max_p; ++c)
        acc += C[r + c * max_p] * F[c];
      A[r] = acc;
    }
    /* Q = c + F' A */
    fa = 0.0;
    for (r = 0; r < max_p; ++r)
      fa += F[r] * A[r];
    Q = s_t[j] + fa;
    for (r = 0; r < max_p; ++r)
      A[r] = A[r] / Q;
    /* e = y - F' a */
    fm = 0.0;
    for (r = 0; r < max_p; ++r)
      fm += F[r] * mt[r];
    e = y_t[j] - fm;
    n_t[j] = n_t[j] + 1.0;
    s_t[j] = s_t[j] * (n_t[j] - 1.0 + e * e / Q) / n_t[j];
    for (r = 0; r < max_p; ++r)
      mt[r] = mt[r] + e * A[r];
    for (r = 0; r < max_p; ++r)
      A[r] = A[r] * sqrt(Q);
    for (c = 0; c < max_p; ++c)
      for (r = 0; r < max_p; ++r)
        C[r + c * max_p] = C[r + c * max_p] - A[r] * A[c];
    factor = ((n_t[j] - 1.0) + e * e / Q) / n_t[j];
    for (r = 0; r < mp2; ++r)
      C[r] *= factor;
  }
  free(A);
}

/* ---------------------------------------------------------------------- */
/* a2: SGDLM.cu:115-179 (multiply :437-444, batchedMatrixScaleWithDiff     */
/* :258-288 with SCALE_TRANSFORMATION_INV; G branch :151-178).             */
/* ---------------------------------------------------------------------- */
void sgo_compute_prior(int m, int max_p, const unsigned *p, double *m_t,
                       double *C_t, double *n_t, double *W_t,
                       const double *beta, const double *delta,
                       const double *G) {
  int j, r, c, k;
  const int mp2 = max_p * max_p;
  double *buf = (double *)malloc(sizeof(double) * (size_t)mp2);
  for (j = 0; j < m; ++j) {
    const int P = (int)p[j];
    double *C = C_t + (size_t)j * mp2;
    double *W = W_t + (size_t)j * mp2;
    const double *D = delta + (size_t)j * mp2;
    n_t[j] = n_t[j] * beta[j];
    for (r = 0; r < mp2; ++r)
      W[r] = 0.0;
    for (c = 0; c < P; ++c)
      for (r = 0; r < P; ++r) {
        const int q = r + c * max_p;
        W[q] = C[q] * (1.0 / D[q] - 1.0);
        C[q] = C[q] / D[q];
      }
    if (G != NULL) {
      const double *Gj = G + (size_t)j * mp2;
      double *mt = m_t + (size_t)j * max_p;
      for (r = 0; r < max_p; ++r)
        buf[r] = mt[r];
      for (r = 0; r < max_p; ++r) {
        double acc = 0.0;
        for (c = 0; c < max_p; ++c)
          acc += Gj[r + c * max_p] * buf[c];
        mt[r] = acc;
      }
      for (r = P; r < max_p; ++r)
        mt[r] = 0.0;
      /* buf = G R ; C = buf G' */
      for (c = 0; c < max_p; ++c)
        for (r = 0; r < max_p; ++r) {
          double acc = 0.0;
          for (k = 0; k < max_p; ++k)
            acc += Gj[r + k * max_p] * C[k + c * max_p];
          buf[r + c * max_p] = acc;
        }
      for (c = 0; c < max_p; ++c)
        for (r = 0; r < max_p; ++r) {
          double acc = 0.0;
          for (k = 0; k < max_p; ++k)
            acc += buf[r + k * max_p] * Gj[c + k * max_p];
          C[r + c * max_p] = acc;
        }
      for (c = 0; c < max_p; ++c)
        for (r = 0; r < max_p; ++r)
          if (r >= P && c >= P)
            C[r + c * max_p] = 0.0;
    }
  }
  free(buf);
}

/* ---------------------------------------------------------------------- */
/* a3-a7: SGDLM.cu:623-677, curand_manager.cuh:111-141, cholesky          */
/* kernel_functions.cuh:844-902, makeVB_MVN :71-83.                        */
/* ---------------------------------------------------------------------- */
void sgo_sample(int m, int max_p, const unsigned *p, const double *m_t,
                const double *C_t, const double *n_t, const double *s_t,
                long N, uint64_t seed, uint32_t epoch, double *lambdas,
                double *thetas) {
  const int mp2 = max_p * max_p;
  double *U = (double *)malloc(sizeof(double) * (size_t)m * mp2);
  long i;
  int j;
  for (j = 0; j < m; ++j)
    sg_cholesky_upper((int)p[j], max_p, max_p, C_t + (size_t)j * mp2,
                      U + (size_t)j * mp2);
#pragma omp parallel for schedule(static) private(j)
  for (i = 0; i < N; ++i) {
    for (j = 0; j < m; ++j) {
      const int P = (int)p[j];
      const double *mt = m_t + (size_t)j * max_p;
      double *th = thetas + ((size_t)i * m + j) * max_p;
      int k;
      lambdas[(size_t)i * m + j] = sg_draw_lambda_theta(
          P, max_p, mt, U + (size_t)j * mp2, n_t[j], s_t[j], seed, epoch,
          (uint32_t)i, (uint32_t)j, th, 1);
      for (k = P; k < max_p; ++k)
        th[k] = mt[k];
    }
  }
  free(U);
}

/* ---------------------------------------------------------------------- */
/* a9/a10: dense partial-pivot LU (the role of cublasXgetrfBatched,        */
/* SGDLM.cu:280) and log|det| = sum log|u_kk| (detFromLU                   */
/* kernel_functions.cuh:904-926).  Row-major, blocked right-looking.       */
/* ---------------------------------------------------------------------- */
#define SGO_NB 32
double sgo_lu_logabsdet(int n, double *A) {
  double logdet = 0.0;
  int K, j, i, c, t;
  for (K = 0; K < n; K += SGO_NB) {
    const int jb = (n - K < SGO_NB) ? (n - K) : SGO_NB;
    const int Ke = K + jb;
    /* panel factorisation with partial pivoting (LAPACK getf2 order) */
    for (j = K; j < Ke; ++j) {
      int piv = j;
      double best = fabs(A[(size_t)j * n + j]);
      double d;
      for (i = j + 1; i < n; ++i) {
        const double v = fabs(A[(size_t)i * n + j]);
        if (v > best) {
          best = v;
          piv = i;
        }
      }
      if (!(best > 0.0))
        return -INFINITY;
      if (piv != j) {
        double *ra = A + (size_t)j * n, *rb = A + (size_t)piv * n;
        for (c = 0; c < n; ++c) {
          const double tmp = ra[c];
          ra[c] = rb[c];
          rb[c] = tmp;
        }
      }
      d = A[(size_t)j * n + j];
      logdet += log(fabs(d));
      for (i = j + 1; i < n; ++i) {
        double *ri = A + (size_t)i * n;
        const double *rj = A + (size_t)j * n;
        const double l = ri[j] / d;
        ri[j] = l;
        for (c = j + 1; c < Ke; ++c)
          ri[c] = fma(-l, rj[c], ri[c]);
      }
    }
    if (Ke >= n)
      break;
    /* U12 = L11^{-1} A12 */
    for (j = K + 1; j < Ke; ++j) {
      double *rj = A + (size_t)j * n;
      for (t = K; t < j; ++t) {
        const double l = rj[t];
        const double *rt = A + (size_t)t * n;
        for (c = Ke; c < n; ++c)
          rj[c] = fma(-l, rt[c], rj[c]);
      }
    }
    /* A22 -= L21 U12 */
    for (i = Ke; i < n; ++i) {
      double *ri = A + (size_t)i * n;
      for (t = K; t < Ke; ++t) {
        const double l = ri[t];
        const double *rt = A + (size_t)t * n;
        for (c = Ke; c < n; ++c)
          ri[c] = fma(-l, rt[c], ri[c]);
      }
    }
  }
  return logdet;
}

/* Row-major I - Gamma for one draw (fillGammas kernel_functions.cuh:821-842:
 * Gamma position v = row + col*m). */
static void build_I_minus_Gamma(int m, int max_p, const unsigned *sp,
                                const double *theta_draw, double *A) {
  const unsigned mm = (unsigned)m * (unsigned)m;
  int q;
  memset(A, 0, sizeof(double) * (size_t)m * m);
  for (q = 0; q < m * max_p; ++q) {
    const unsigned v = sp[q];
    if (v < mm) {
      const unsigned row = v % (unsigned)m, col = v / (unsigned)m;
      A[(size_t)row * m + col] = -theta_draw[q];
    }
  }
  for (q = 0; q < m; ++q)
    A[(size_t)q * m + q] = 1.0;
}

void sgo_logdet(int m, int max_p, const unsigned *p, const unsigned *sp,
                long N, const double *thetas, double *logdet, int nthreads) {
  long i;
  (void)p;
  set_threads(nthreads);
#pragma omp parallel
  {
    double *A = (double *)malloc(sizeof(double) * (size_t)m * m);
#pragma omp for schedule(dynamic, 4)
    for (i = 0; i < N; ++i) {
      build_I_minus_Gamma(m, max_p, sp, thetas + (size_t)i * m * max_p, A);
      logdet[i] = sgo_lu_logabsdet(m, A);
    }
    free(A);
  }
}

/* ---------------------------------------------------------------------- */
/* a19: psi, psi1 kernel_functions.cuh:1006-1022, compute_mean_n_t        */
/* :1024-1068 (truncated series, Newton with step clamp and fallback).     */
/* ---------------------------------------------------------------------- */
double sgo_psi(double x) {
  return log(x) - 1.0 / (2.0 * x) - 1.0 / (12.0 * x * x) +
         1.0 / (120.0 * x * x * x * x) - 1.0 / (252.0 * x * x * x * x * x * x) +
         1.0 / (240.0 * x * x * x * x * x * x * x * x) -
         5.0 / (660.0 * x * x * x * x * x * x * x * x * x * x) +
         691.0 / (32760.0 * x * x * x * x * x * x * x * x * x * x * x * x) -
         1.0 / (12.0 * x * x * x * x * x * x * x * x * x * x * x * x * x * x);
}

double sgo_psi1(double x) {
  return 1.0 / x + 1.0 / (2.0 * x * x) + 1.0 / (6.0 * x * x * x) -
         1.0 / (30.0 * x * x * x * x * x) +
         1.0 / (42.0 * x * x * x * x * x * x * x) -
         1.0 / (30.0 * x * x * x * x * x * x * x * x * x);
}

double sgo_mean_n(unsigned p, double mean_Q, double mean_lambda,
                  double mean_log_lambda, double n_t, int *iter_out) {
  int iter = 0;
  double mn = n_t;
  double fun, dfun, step = 0.0;
  const double pd = (double)p;
  do {
    mn -= step;
    fun = log(mn + pd - mean_Q) - sgo_psi(mn / 2.0) - (pd - mean_Q) / mn -
          log(2.0 * mean_lambda) + mean_log_lambda;
    dfun = 1.0 / (mn + pd - mean_Q) - 0.5 * sgo_psi1(mn / 2.0) +
           (pd - mean_Q) / (mn * mn);
    step = fun / dfun;
    if (step > mn)
      step = 0.5 * mn;
  } while (fabs(fun) > OPTIM_PRECISION && iter++ < OPTIM_MAX_ITER);
  if (iter_out)
    *iter_out = iter;
  return (iter >= OPTIM_MAX_ITER) ? n_t : mn;
}

/* In-place inverse of an n x n column-major matrix by partial-pivot
 * Gauss-Jordan (the role of getrf/getriBatched at SGDLM.cu:427-434). */
static int invert_colmajor(int n, int ld, double *A) {
  int *perm = (int *)malloc(sizeof(int) * (size_t)n);
  double *B = (double *)malloc(sizeof(double) * (size_t)n * 2 * n);
  int i, j, k, ok = 1;
  for (i = 0; i < n; ++i)
    for (j = 0; j < 2 * n; ++j)
      B[(size_t)i * 2 * n + j] =
          (j < n) ? A[i + (size_t)j * ld] : (j - n == i ? 1.0 : 0.0);
  for (k = 0; k < n; ++k) {
    int piv = k;
    double best = fabs(B[(size_t)k * 2 * n + k]), d;
    for (i = k + 1; i < n; ++i)
      if (fabs(B[(size_t)i * 2 * n + k]) > best) {
        best = fabs(B[(size_t)i * 2 * n + k]);
        piv = i;
      }
    if (!(best > 0.0)) {
      ok = 0;
      break;
    }
    if (piv != k)
      for (j = 0; j < 2 * n; ++j) {
        double t = B[(size_t)k * 2 * n + j];
        B[(size_t)k * 2 * n + j] = B[(size_t)piv * 2 * n + j];
        B[(size_t)piv * 2 * n + j] = t;
      }
    d = B[(size_t)k * 2 * n + k];
    for (j = 0; j < 2 * n; ++j)
      B[(size_t)k * 2 * n + j] /= d;
    for (i = 0; i < n; ++i)
      if (i != k) {
        const double f = B[(size_t)i * 2 * n + k];
        if (f != 0.0)
          for (j = 0; j < 2 * n; ++j)
            B[(size_t)i * 2 * n + j] -= f * B[(size_t)k * 2 * n + j];
      }
  }
  if (ok)
    for (i = 0; i < n; ++i)
      for (j = 0; j < n; ++j)
        A[i + (size_t)j * ld] = B[(size_t)i * 2 * n + n + j];
  free(perm);
  free(B);
  return ok;
}

/* ---------------------------------------------------------------------- */
/* a3-a20: VB_posterior SGDLM.cu:181-503.                                  */
/* ---------------------------------------------------------------------- */
void sgo_vb(int m, int max_p, const unsigned *p, const unsigned *sp,
            const double *m_t, const double *C_t, const double *n_t,
            const double *s_t, long N, uint64_t seed, uint32_t epoch,
            double *mean_m, double *mean_C, double *mean_n, double *mean_s,
            double *w, sgo_vb_diag *diag, int nthreads) {
  const int mp2 = max_p * max_p;
  double *lam = (double *)malloc(sizeof(double) * (size_t)N * m);
  double *th = (double *)malloc(sizeof(double) * (size_t)N * m * max_p);
  double *ld = (double *)malloc(sizeof(double) * (size_t)N);
  double *Vinv = (double *)malloc(sizeof(double) * (size_t)mp2);
  double S = 0.0, maxld = -INFINITY, ess = 0.0;
  long i;
  int j;

  set_threads(nthreads);
  sgo_sample(m, max_p, p, m_t, C_t, n_t, s_t, N, seed, epoch, lam, th);
  sgo_logdet(m, max_p, p, sp, N, th, ld, nthreads);

  /* IS weights: det = exp(sum log|u_kk|), non-finite -> 0, normalised by a
   * serial sum (detFromLU :904-926, sumVector :25-39, scale :164-184). */
  for (i = 0; i < N; ++i) {
    double d = exp(ld[i]);
    if (isnan(d) || isinf(d))
      d = 0.0;
    w[i] = d;
    if (ld[i] > maxld)
      maxld = ld[i];
  }
  for (i = 0; i < N; ++i)
    S += w[i];
  for (i = 0; i < N; ++i)
    w[i] = w[i] / S;
  for (i = 0; i < N; ++i)
    ess += w[i] * w[i];
  ess = 1.0 / ess;

  for (j = 0; j < m; ++j) {
    const int P = (int)p[j];
    double lbar = 0.0, llbar = 0.0, Qbar = 0.0;
    double *mm = mean_m + (size_t)j * max_p;
    double *CC = mean_C + (size_t)j * mp2;
    int k, x, y, it = 0;
    /* meanScalarAndLog :928-946 */
    for (i = 0; i < N; ++i) {
      const double l = lam[(size_t)i * m + j];
      lbar += w[i] * l;
      llbar += w[i] * log(l);
    }
    /* compute_mean_m_t :948-967 */
    for (k = 0; k < max_p; ++k) {
      double acc = 0.0;
      for (i = 0; i < N; ++i)
        acc += w[i] * lam[(size_t)i * m + j] * th[((size_t)i * m + j) * max_p + k];
      mm[k] = acc / lbar;
    }
    /* compute_VB_vector1 :969-982 (in place) */
    for (i = 0; i < N; ++i) {
      double *t = th + ((size_t)i * m + j) * max_p;
      const double sl = sqrt(lam[(size_t)i * m + j]);
      for (k = 0; k < max_p; ++k)
        t[k] = sl * (t[k] - mm[k]);
    }
    /* meanBatchedVecVecT :1102-1133 */
    for (y = 0; y < max_p; ++y)
      for (x = 0; x < max_p; ++x) {
        if (x < P && y < P) {
          if (x <= y) {
            double acc = 0.0;
            for (i = 0; i < N; ++i) {
              const double *t = th + ((size_t)i * m + j) * max_p;
              acc += w[i] * t[x] * t[y];
            }
            CC[y * max_p + x] = acc;
            CC[x * max_p + y] = acc;
          }
        } else {
          CC[y * max_p + x] = (x == y) ? 1.0 : 0.0;
        }
      }
    /* cut-offs SGDLM.cu:385-399 */
    for (k = P; k < max_p; ++k)
      mm[k] = 0.0;
    for (y = 0; y < max_p; ++y)
      for (x = 0; x < max_p; ++x)
        if (x >= P && y >= P)
          CC[x + y * max_p] = 0.0;
    /* V^{-1} with unit diagonal outside the block, then cut off
     * (SGDLM.cu:404-444) */
    memcpy(Vinv, CC, sizeof(double) * (size_t)mp2);
    for (k = P; k < max_p; ++k)
      Vinv[k * max_p + k] = 1.0;
    invert_colmajor(max_p, max_p, Vinv);
    for (y = 0; y < max_p; ++y)
      for (x = 0; x < max_p; ++x)
        if (x >= P && y >= P)
          Vinv[x + y * max_p] = 0.0;
    /* Q_ij = theta~' V^{-1} theta~ ; Qbar = sum w Q (SGDLM.cu:446-465,
     * compute_mean_Q_t :1087-1100) */
    for (i = 0; i < N; ++i) {
      const double *t = th + ((size_t)i * m + j) * max_p;
      double q = 0.0;
      for (y = 0; y < max_p; ++y) {
        double acc = 0.0;
        for (x = 0; x < max_p; ++x)
          acc += t[x] * Vinv[x + y * max_p];
        q += acc * t[y];
      }
      Qbar += w[i] * q;
    }
    /* compute_mean_n_t, compute_mean_s_t :1024-1085 */
    mean_n[j] = sgo_mean_n((unsigned)P, Qbar, lbar, llbar, n_t[j], &it);
    mean_s[j] = (mean_n[j] + (double)P - Qbar) / (mean_n[j] * lbar);
    if (mean_s[j] < 0.0)
      mean_s[j] = s_t[j];
    /* C = s V on the p x p block (SGDLM.cu:491-500) */
    for (y = 0; y < P; ++y)
      for (x = 0; x < P; ++x)
        CC[x + y * max_p] *= mean_s[j];
    if (diag) {
      if (diag->mean_lambda)
        diag->mean_lambda[j] = lbar;
      if (diag->mean_log_lambda)
        diag->mean_log_lambda[j] = llbar;
      if (diag->mean_Q)
        diag->mean_Q[j] = Qbar;
      if (diag->newton_iter)
        diag->newton_iter[j] = it;
    }
  }
  if (diag) {
    diag->sum_det_weights = S;
    diag->ess = ess;
    diag->max_logdet = maxld;
  }
  free(lam);
  free(th);
  free(ld);
  free(Vinv);
}

/* ---------------------------------------------------------------------- */
/* a22: forecast SGDLM.cu:505-621, batchedComputeNuPlusMu                 */
/* kernel_functions.cuh:1155-1178.  Solve by partial-pivot LU.             */
/* ---------------------------------------------------------------------- */
static int lu_solve(int n, double *A, double *b) {
  int j, i, c;
  for (j = 0; j < n; ++j) {
    int piv = j;
    double best = fabs(A[(size_t)j * n + j]), d;
    for (i = j + 1; i < n; ++i)
      if (fabs(A[(size_t)i * n + j]) > best) {
        best = fabs(A[(size_t)i * n + j]);
        piv = i;
      }
    if (!(best > 0.0))
      return 0;
    if (piv != j) {
      double t;
      for (c = 0; c < n; ++c) {
        t = A[(size_t)j * n + c];
        A[(size_t)j * n + c] = A[(size_t)piv * n + c];
        A[(size_t)piv * n + c] = t;
      }
      t = b[j];
      b[j] = b[piv];
      b[piv] = t;
    }
    d = A[(size_t)j * n + j];
    for (i = j + 1; i < n; ++i) {
      const double l = A[(size_t)i * n + j] / d;
      if (l != 0.0) {
        for (c = j + 1; c < n; ++c)
          A[(size_t)i * n + c] -= l * A[(size_t)j * n + c];
        b[i] -= l * b[j];
      }
    }
  }
  for (j = n - 1; j >= 0; --j) {
    double acc = b[j];
    for (c = j + 1; c < n; ++c)
      acc -= A[(size_t)j * n + c] * b[c];
    b[j] = acc / A[(size_t)j * n + j];
  }
  return 1;
}

void sgo_forecast(int m, int max_p, const unsigned *p, const unsigned *sp,
                  const double *m_t, const double *C_t, const double *n_t,
                  const double *s_t, long N, uint64_t seed, uint32_t e_draw,
                  uint32_t e_nu, const double *x, double *lambdas,
                  double *thetas, int use_existing, double *y, int nthreads) {
  const unsigned mm = (unsigned)m * (unsigned)m;
  long i;
  set_threads(nthreads);
  if (!use_existing)
    sgo_sample(m, max_p, p, m_t, C_t, n_t, s_t, N, seed, e_draw, lambdas,
               thetas);
#pragma omp parallel
  {
    double *A = (double *)malloc(sizeof(double) * (size_t)m * m);
#pragma omp for schedule(dynamic, 4)
    for (i = 0; i < N; ++i) {
      double *yi = y + (size_t)i * m;
      const double *thi = thetas + (size_t)i * m * max_p;
      int j, k;
      for (j = 0; j < m; ++j) {
        double nu = sg_normal_k(seed, e_nu, SG_PURPOSE_NU, (uint32_t)i,
                                (uint32_t)j, 0u);
        nu = nu / sqrt(lambdas[(size_t)i * m + j]);
        for (k = 0; k < max_p; ++k)
          if (sp[(size_t)j * max_p + k] >= mm)
            nu += x[(size_t)j * max_p + k] * thi[(size_t)j * max_p + k];
        yi[j] = nu;
      }
      build_I_minus_Gamma(m, max_p, sp, thi, A);
      if (!lu_solve(m, A, yi))
        for (j = 0; j < m; ++j)
          yi[j] = NAN;
    }
    free(A);
  }
}

/* ---------------------------------------------------------------------- */
/* a23: evolve_lambdas_and_thetas SGDLM.cu:679-762 (sampleBeta             */
/* curand_manager.cuh:146-179, batchedMultiply, scale_omegas with the      */
/* intended indexing, cholesky(W), gemm theta += chol(W)' omega).          */
/* ---------------------------------------------------------------------- */
void sgo_evolve(int m, int max_p, const unsigned *p, const double *W_t,
                const double *r_t, const double *c_t, const double *beta,
                const double *G, long N, uint64_t seed, uint32_t epoch,
                double *lambdas, double *thetas) {
  const int mp2 = max_p * max_p;
  double *U = (double *)malloc(sizeof(double) * (size_t)m * mp2);
  long i;
  int j;
  for (j = 0; j < m; ++j)
    sg_cholesky_upper((int)p[j], max_p, max_p, W_t + (size_t)j * mp2,
                      U + (size_t)j * mp2);
#pragma omp parallel
  {
    double *om = (double *)malloc(sizeof(double) * (size_t)max_p);
    double *buf = (double *)malloc(sizeof(double) * (size_t)max_p);
#pragma omp for schedule(static)
    for (i = 0; i < N; ++i) {
      for (j = 0; j < m; ++j) {
        const int P = (int)p[j];
        double *lam = lambdas + (size_t)i * m + j;
        double *th = thetas + ((size_t)i * m + j) * max_p;
        const double *Uj = U + (size_t)j * mp2;
        double eta = sg_beta_eta(r_t[j], beta[j], seed, epoch, (uint32_t)i,
                                 (uint32_t)j);
        double sc;
        int k, n;
        *lam = *lam * eta;
        sc = sqrt(c_t[j] * (*lam));
        for (k = 0; k < max_p; ++k)
          om[k] = (k < P) ? sg_normal_k(seed, epoch, SG_PURPOSE_OMEGA,
                                        (uint32_t)i, (uint32_t)j, (uint32_t)k) /
                                sc
                          : 0.0;
        if (G != NULL) {
          const double *Gj = G + (size_t)j * mp2;
          for (k = 0; k < max_p; ++k)
            buf[k] = th[k];
          for (k = 0; k < max_p; ++k) {
            double acc = 0.0;
            for (n = 0; n < max_p; ++n)
              acc += Gj[k + n * max_p] * buf[n];
            th[k] = acc;
          }
        }
        for (k = 0; k < max_p; ++k) {
          double acc = 0.0;
          for (n = 0; n <= k; ++n)
            acc = acc + Uj[k * max_p + n] * om[n];
          th[k] = acc + th[k];
        }
      }
    }
    free(om);
    free(buf);
  }
  free(U);
}

/* ----------------------- shared-numerics exports ------------------------ */
void sgo_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                uint32_t k0, uint32_t k1, uint32_t *out4) {
  sg_u32x4 r = sg_philox4x32_10(c0, c1, c2, c3, k0, k1);
  out4[0] = r.v[0];
  out4[1] = r.v[1];
  out4[2] = r.v[2];
  out4[3] = r.v[3];
}
double sgo_log(double x) { return sg_log(x); }
double sgo_exp(double x) { return sg_exp(x); }
void sgo_sincospi(double x, double *s, double *c) { sg_sincospi(x, s, c); }
double sgo_gamma(double shape, double rate, uint64_t seed, uint32_t epoch,
                 uint32_t purpose, uint32_t draw, uint32_t series) {
  return sg_gamma(shape, rate, seed, epoch, purpose, draw, series);
}
double sgo_normal_k(uint64_t seed, uint32_t epoch, uint32_t purpose,
                    uint32_t draw, uint32_t series, uint32_t k) {
  return sg_normal_k(seed, epoch, purpose, draw, series, k);
}
double sgo_beta_eta(double r, double beta, uint64_t seed, uint32_t epoch,
                    uint32_t draw, uint32_t series) {
  return sg_beta_eta(r, beta, seed, epoch, draw, series);
}
void sgo_cholesky_upper(int p, int max_p, const double *C, double *U) {
  sg_cholesky_upper(p, max_p, max_p, C, U);
}
