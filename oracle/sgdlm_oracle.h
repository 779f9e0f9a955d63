/*
 * sgdlm_oracle.h -- FP64 CPU oracle for the SGDLM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gpusgdlm_amd/, the
 * C-ABI library) includes, links or calls this code; only tests/, the smoke
 * check and the CPU-baseline leg of bench.py do, as the checker.
 *
 * Layouts follow the reference's host buffers (R column-major arrays
 * flattened to double*): m_t[j*max_p + k], C_t[j*max_p*max_p + r + c*max_p],
 * lambdas[i*m + j], thetas[(i*m + j)*max_p + k], y[i*m + j].
 */
#ifndef SGDLM_ORACLE_H
#define SGDLM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SGDLM.cu:10-113 -- naive conjugate filter update, in place. */
void sgo_compute_posterior(int m, int max_p, const unsigned *p, double *m_t,
                           double *C_t, double *n_t, double *s_t,
                           const double *y_t, const double *F_t);

/* SGDLM.cu:115-179 -- one-step-ahead prior (discount), in place; G may be
 * NULL (no state evolution matrix). */
void sgo_compute_prior(int m, int max_p, const unsigned *p, double *m_t,
                       double *C_t, double *n_t, double *W_t,
                       const double *beta, const double *delta,
                       const double *G);

/* SGDLM.cu:623-677 -- lambda_ij ~ Gamma(n/2, n s/2), theta_ij ~ N(m, C/(s
 * lambda)) in the reference layout (entries k >= p_j hold m_t[k]). */
void sgo_sample(int m, int max_p, const unsigned *p, const double *m_t,
                const double *C_t, const double *n_t, const double *s_t,
                long N, uint64_t seed, uint32_t epoch, double *lambdas,
                double *thetas);

/* kernel_functions.cuh:821-842,904-926 + getrfBatched: log|det(I - Gamma_i)|
 * per draw by partial-pivot LU in FP64; -inf for a singular or non-finite
 * factorisation.  nthreads <= 0 means OpenMP default. */
void sgo_logdet(int m, int max_p, const unsigned *p, const unsigned *sp,
                long N, const double *thetas, double *logdet, int nthreads);

/* Dense partial-pivot LU log|det| of one row-major n x n matrix (destroyed).
 * Returns -inf if singular. */
double sgo_lu_logabsdet(int n, double *A);

/* Diagnostics of one VB call. */
typedef struct {
  double sum_det_weights;
  double ess;
  double max_logdet;
  double *mean_lambda;     /* m */
  double *mean_log_lambda; /* m */
  double *mean_Q;          /* m */
  int *newton_iter;        /* m */
} sgo_vb_diag;

/* SGDLM.cu:181-503 -- VB posterior by importance sampling with N draws.
 * Outputs mean_m (m*max_p), mean_C (m*max_p^2), mean_n, mean_s (m), w (N).
 * lambdas/thetas may be NULL (then allocated internally). */
void sgo_vb(int m, int max_p, const unsigned *p, const unsigned *sp,
            const double *m_t, const double *C_t, const double *n_t,
            const double *s_t, long N, uint64_t seed, uint32_t epoch,
            double *mean_m, double *mean_C, double *mean_n, double *mean_s,
            double *w, sgo_vb_diag *diag, int nthreads);

/* SGDLM.cu:505-621 -- forecast draws y_i = (I - Gamma_i)^{-1} rhs_i.  If
 * use_existing is 0, lambdas/thetas are drawn from the given prior (epoch
 * e_draw); nu uses epoch e_nu.  y is N x m. */
void sgo_forecast(int m, int max_p, const unsigned *p, const unsigned *sp,
                  const double *m_t, const double *C_t, const double *n_t,
                  const double *s_t, long N, uint64_t seed, uint32_t e_draw,
                  uint32_t e_nu, const double *x, double *lambdas,
                  double *thetas, int use_existing, double *y, int nthreads);

/* SGDLM.cu:679-762 -- multi-step evolution of the forecast samples, with the
 * intended scaling of scale_omegas (omega_k / sqrt(c lambda), zero for
 * k >= p).  G may be NULL. */
void sgo_evolve(int m, int max_p, const unsigned *p, const double *W_t,
                const double *r_t, const double *c_t, const double *beta,
                const double *G, long N, uint64_t seed, uint32_t epoch,
                double *lambdas, double *thetas);

/* kernel_functions.cuh:1006-1068 -- Newton moment-match for n. */
double sgo_mean_n(unsigned p, double mean_Q, double mean_lambda,
                  double mean_log_lambda, double n_t, int *iter_out);
double sgo_psi(double x);
double sgo_psi1(double x);

/* Shared-numerics entry points exported for known-answer tests. */
void sgo_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                uint32_t k0, uint32_t k1, uint32_t *out4);
double sgo_log(double x);
double sgo_exp(double x);
void sgo_sincospi(double x, double *s, double *c);
double sgo_gamma(double shape, double rate, uint64_t seed, uint32_t epoch,
                 uint32_t purpose, uint32_t draw, uint32_t series);
double sgo_normal_k(uint64_t seed, uint32_t epoch, uint32_t purpose,
                    uint32_t draw, uint32_t series, uint32_t k);
double sgo_beta_eta(double r, double beta, uint64_t seed, uint32_t epoch,
                    uint32_t draw, uint32_t series);
void sgo_cholesky_upper(int p, int max_p, const double *C, double *U);

#ifdef __cplusplus
}
#endif

#endif
