/*
 * sgdlm.h -- C-ABI of the MI355X-native SGDLM engine (libsgdlm_hip.so).
 *
 * One entry point per method of the reference's C++ plugin interface
 * SGDLM::HostWrapper<double> (rSGDLM/src/SGDLM/HostWrapper.hpp:13-99), which
 * the Rcpp module gpuSGDLM/SGDLM calls (rSGDLM/src/SGDLM/RcppWrapper.cpp).
 * A C++ implementation of that interface over these functions is in
 * include/SGDLM/HostWrapper.hpp + gpusgdlm_amd/csrc/host_wrapper.cpp.
 *
 * Conventions
 *   - Every function returns an int status (SGDLM_OK on success).  The
 *     reference logs and returns early on errors (logger.hpp:160-165); here
 *     the reason is kept in sgdlm_last_error().
 *   - Host buffers are caller-owned, in the reference's column-major R
 *     layouts flattened to double* / unsigned int*:
 *       p[m], sp[j*max_p + k], beta[m], delta[j*max_p^2 + r + c*max_p],
 *       m_t[j*max_p + k], C_t[j*max_p^2 + r + c*max_p], n_t[m], s_t[m],
 *       y_t[m], F_t[j*max_p + k], G[j*max_p^2 + r + c*max_p],
 *       IS weights[nsim], forecast y[i*m + j], lambdas[i*m + j],
 *       thetas[(i*m + j)*max_p + k].
 *     NULL output pointers are skipped, as in the reference getters.
 *   - Calls are synchronous.  A context is thread-compatible (one thread at
 *     a time), not thread-safe.
 *   - One GPU per context (the reference's no_gpus is clamped to 1).
 */
#ifndef SGDLM_H
#define SGDLM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGDLM_OK 0
#define SGDLM_ERR_ARG 1     /* invalid argument or dimensions              */
#define SGDLM_ERR_STATE 2   /* call not allowed in the current state       */
#define SGDLM_ERR_DEVICE 3  /* HIP runtime error                           */
#define SGDLM_ERR_ALLOC 4   /* device allocation failed                    */
#define SGDLM_ERR_NUMERIC 5 /* e.g. every importance weight is zero        */

/* Arithmetic of the dense det / solve work (LU of I - Gamma). */
#define SGDLM_PREC_FP64 0   /* FP64 LU (MFMA f64 trailing updates)        */
#define SGDLM_PREC_FP32 1   /* FP32 LU on MFMA, FP64 log-det sums          */
#define SGDLM_PREC_BF16X3 2 /* BF16x3 split-precision MFMA trailing update */

/* Kernel kinds of the VB log-determinant LU (sgdlm_get_lu_kernel,
 * sgdlm_set_lu_kernel, sgdlm_get_lu_autotune). */
#define SGDLM_LU_AUTO 0     /* chosen by the plan or the autotuner          */
#define SGDLM_LU_BLOCKED 1  /* dense block through LDS panels (16/32/64)    */
#define SGDLM_LU_REGISTER 2 /* dense block resident in one CU's registers
                               (FP32 / BF16x3, up to 288 rows)              */
#define SGDLM_LU_PIVOTING 3 /* every draw with row interchanges             */
#define SGDLM_LU_WAVE 4     /* small models (m <= 128 FP32, <= 64 FP64): one
                               wavefront per draw, the whole matrix in
                               registers, partial pivoting                  */
#define SGDLM_LU_WAVE_PREFIX 5 /* dense blocks after the sparse prefix of up
                               to 128 rows (FP32) / 64 (FP64): one wavefront
                               per draw, the prefix on its own LDS slots, the
                               Schur complement in registers, partial
                               pivoting                                     */
#define SGDLM_LU_REGISTER_PIVOT 6 /* dense block resident in one CU's
                               registers (FP32, up to 288 rows) with partial
                               pivoting: each 16-wide column panel factored
                               by one wavefront from LDS, its row exchanges
                               applied to the register tiles                */

typedef struct sgdlm_ctx sgdlm_ctx;

/* HostWrapperFactory<double>::create(no_gpus) -> HostWrapperImpl ctor,
 * HostWrapperFactory.cu:5-11, HostWrapperImpl.cu:11-74 (device, stream,
 * generator seeded 1234). */
int sgdlm_create(sgdlm_ctx **ctx, size_t no_gpus);
/* ~HostWrapperImpl, HostWrapperImpl.cu:76-104. */
int sgdlm_destroy(sgdlm_ctx *ctx);
/* Text of the last error of this context (empty if none). */
const char *sgdlm_last_error(const sgdlm_ctx *ctx);

/* HostWrapper.hpp:17-21 getNoSeries / getMaxP /
 * getEvolutionMatrixConfiguration. */
int sgdlm_get_no_series(const sgdlm_ctx *ctx, size_t *m);
int sgdlm_get_max_p(const sgdlm_ctx *ctx, size_t *max_p);
int sgdlm_get_evolution_matrix_configuration(const sgdlm_ctx *ctx,
                                             int *use_state_evolution_matrix);
/* HostWrapper.hpp:23-32 initMemory / clearMemory (HostWrapperImpl.cu:129-300). */
int sgdlm_init_memory(sgdlm_ctx *ctx, size_t m, size_t max_p);
int sgdlm_clear_memory(sgdlm_ctx *ctx);
/* HostWrapper.hpp:34 manageEvoMemory (HostWrapperImpl.cu:230-279). */
int sgdlm_manage_evo_memory(sgdlm_ctx *ctx, int use_state_evolution_matrix);
/* HostWrapper.hpp:36-42 getNSim / getNSimBatch / initSimMemory /
 * clearSimMemory (HostWrapperImpl.cu:302-537).  nsim_batch is kept as the
 * reference's memory hint; results never depend on it. */
int sgdlm_get_nsim(const sgdlm_ctx *ctx, size_t *nsim);
int sgdlm_get_nsim_batch(const sgdlm_ctx *ctx, size_t *nsim_batch);
int sgdlm_init_sim_memory(sgdlm_ctx *ctx, size_t nsim, size_t nsim_batch);
int sgdlm_clear_sim_memory(sgdlm_ctx *ctx);
/* HostWrapper.hpp:44 initForecastPriors (HostWrapperImpl.cu:539-587). */
int sgdlm_init_forecast_priors(sgdlm_ctx *ctx);
/* HostWrapper.hpp:46-48 isPrior() / isPrior(bool). */
int sgdlm_get_is_prior(const sgdlm_ctx *ctx, int *is_prior);
int sgdlm_set_is_prior(sgdlm_ctx *ctx, int is_prior);
/* HostWrapper.hpp:50-58 getParameters / setParameters. */
int sgdlm_get_parameters(const sgdlm_ctx *ctx, double *m_t, double *C_t,
                         double *n_t, double *s_t);
int sgdlm_set_parameters(sgdlm_ctx *ctx, const double *m_t, const double *C_t,
                         const double *n_t, const double *s_t);
/* HostWrapper.hpp:60-64 getDiscountFactors / setDiscountFactors. */
int sgdlm_get_discount_factors(const sgdlm_ctx *ctx, double *beta,
                               double *delta);
int sgdlm_set_discount_factors(sgdlm_ctx *ctx, const double *beta,
                               const double *delta);
/* HostWrapper.hpp:66-68 getEvolutionMatrix / setEvolutionMatrix. */
int sgdlm_get_evolution_matrix(const sgdlm_ctx *ctx, double *G_t);
int sgdlm_set_evolution_matrix(sgdlm_ctx *ctx, const double *G_t);
/* HostWrapper.hpp:70-74 getParentalSets / setParentalSets.  sp values
 * v < m*m are positions row + col*m of Gamma (row = child); anything else
 * (R's NA -> 2^31) marks an external regressor.  Rejected with SGDLM_ERR_ARG
 * (the reference's result would be a write race or undefined): a diagonal
 * position, a position used twice, a simultaneous position at k >= p_j,
 * p_j > max_p. */
int sgdlm_get_parental_sets(const sgdlm_ctx *ctx, unsigned int *p,
                            unsigned int *sp);
int sgdlm_set_parental_sets(sgdlm_ctx *ctx, const unsigned int *p,
                            const unsigned int *sp);
/* HostWrapper.hpp:76-79 computePrior(bool) / computePrior(G, bool)
 * (HostWrapperImpl.cu:799-877; SGDLM.cu:115-179).  G_tp1 may be NULL. */
int sgdlm_compute_prior(sgdlm_ctx *ctx, const double *G_tp1,
                        int evolve_forecast_variables);
/* HostWrapper.hpp:81-83 evolveForecastSamples() / (G)
 * (HostWrapperImpl.cu:879-946; SGDLM.cu:679-762).  G_tp1 may be NULL. */
int sgdlm_evolve_forecast_samples(sgdlm_ctx *ctx, const double *G_tp1);
/* HostWrapper.hpp:85-89 computeForecast (HostWrapperImpl.cu:948-1021;
 * SGDLM.cu:505-621).  y: nsim x m; lambdas, thetas may be NULL. */
int sgdlm_compute_forecast(sgdlm_ctx *ctx, double *y_tp1, const double *x_tp1,
                           double *lambdas, double *thetas,
                           int use_existing_lambdas_and_thetas);
/* HostWrapper.hpp:91-92 computePosterior (HostWrapperImpl.cu:1023-1057;
 * SGDLM.cu:10-113). */
int sgdlm_compute_posterior(sgdlm_ctx *ctx, const double *y_t,
                            const double *F_t);
/* HostWrapper.hpp:94-99 computeVBPosterior (HostWrapperImpl.cu:1059-1228;
 * SGDLM.cu:181-503).  With every output NULL the results stay on the device
 * (no PCIe copy); sgdlm_get_vb_result fetches them later. */
int sgdlm_compute_vb_posterior(sgdlm_ctx *ctx, double *mean_m_t,
                               double *mean_C_t, double *mean_n_t,
                               double *mean_s_t, double *IS_weights,
                               double *sum_det_weights);

/* ---- Extensions (no counterpart in the reference) ---------------------- */

/* LU arithmetic for the VB log-determinants and forecast solves
 * (SGDLM_PREC_*; default SGDLM_PREC_FP64). */
int sgdlm_set_precision(sgdlm_ctx *ctx, int precision);
int sgdlm_get_precision(const sgdlm_ctx *ctx, int *precision);
/* Philox state: key = seed (default 1234, the reference's cuRAND seed at
 * HostWrapperImpl.cu:63-64) and the epoch counter advanced by every call
 * that draws random numbers.  Checkpointable. */
int sgdlm_set_rng_state(sgdlm_ctx *ctx, uint64_t seed, uint32_t epoch);
int sgdlm_get_rng_state(const sgdlm_ctx *ctx, uint64_t *seed,
                        uint32_t *epoch);
/* Results of the last VB call left on the device: same layouts as
 * sgdlm_compute_vb_posterior, plus log_weights (nsim, log of the
 * unnormalised weight, -inf for a zero weight), ess = 1 / sum w^2, and
 * log_sum_det = log(sum_det_weights).  Any pointer may be NULL. */
int sgdlm_get_vb_result(const sgdlm_ctx *ctx, double *mean_m_t,
                        double *mean_C_t, double *mean_n_t, double *mean_s_t,
                        double *IS_weights, double *log_weights, double *ess,
                        double *log_sum_det);
/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): sgdlm_compute_vb_posterior with nsim
 * draws whose every device buffer beyond the model state (draws,
 * log-determinants, LU workspace, weights, moment partials, VB results)
 * is carved from a static 64 MiB array of this library on the context's
 * GPU instead of the simulation pool; it calls no hipMalloc / hipFree and
 * needs no sgdlm_init_sim_memory.  Results go to the caller's host buffers
 * (NULL skips one) as in sgdlm_compute_vb_posterior, and are identical to
 * its results for the same RNG state.  Sizes whose buffers exceed the array
 * fail with SGDLM_ERR_ARG (a draw takes 8 (m + sum_j p_j) + 20 bytes, plus
 * the LU workspace: m = 6 allows nsim up to ~3e5, m = 100 with
 * k ~ U{0..5} ~1.2e4);
 * scratch_bytes (optional) receives the bytes the call needs.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other threads, other contexts) race.  The LU grid may
 * be narrowed to fit, so large problems run slower than through the
 * simulation pool. */
int sgdlm_compute_vb_posterior_noalloc(sgdlm_ctx *ctx, size_t nsim,
                                       double *mean_m_t, double *mean_C_t,
                                       double *mean_n_t, double *mean_s_t,
                                       double *IS_weights,
                                       double *sum_det_weights,
                                       size_t *scratch_bytes);
/* Batched VB for many independent models (one context each, all on this
 * GPU): runs sgdlm_compute_vb_posterior with device-resident results on
 * ctxs[0..count); fetch results with sgdlm_get_vb_result.  Every stage runs
 * as one launch over all models and their draws (the LU as one launch per
 * kernel variant the models' plans use, workgroups split by work) on the
 * first context's stream, from device tables built on the first call for
 * this list and kept by ctxs[0] while the contexts' structures, buffers,
 * seeds and LU plans stay the same.  The models are split into up to four
 * parts of about equal work whose stages run on streams of their own, and
 * the step is captured once as a hipGraph and replayed (run eagerly only
 * while ctxs[0] has stage timing on).  Results equal the
 * one-call-per-model results bit for bit.  A list the group path cannot
 * take (contexts on different devices, max_p > 32, more than 65535 series
 * in all, an LU structure not tuned yet, sgdlm_set_batch_mode(ctxs[0],
 * SGDLM_BATCH_STREAMS)) runs one context per stream instead, enqueued
 * before any wait.  status (optional, count entries) receives each
 * context's status; the return value is the first non-OK status.  Contexts
 * must not be used concurrently from other threads during the call.  Serves
 * the reference's many-instance use (one HostWrapper per model,
 * HostWrapperImpl.cu:1059-1228 once per model) without a host round trip
 * or a launch per model. */
int sgdlm_batch_compute_vb_posterior(sgdlm_ctx *const *ctxs, size_t count,
                                     int *status);
#define SGDLM_BATCH_GROUP 0   /* one launch per stage over all models      */
#define SGDLM_BATCH_STREAMS 1 /* each model's stages on its own stream     */
/* How a batch led by ctx (ctxs[0]) runs.  No counterpart in the reference. */
int sgdlm_set_batch_mode(sgdlm_ctx *ctx, int mode);
/* Batches ctx led as one launch per stage, and one stream per model. */
int sgdlm_get_batch_stats(const sgdlm_ctx *ctx, long *group_calls, long *stream_calls);
/* Per-stage device timing with HIP events recorded on the context's stream
 * around each launch (enable = 1 resets the counters).  stage: "lu" (VB
 * log-det LU), "draw", "weights", "moments", "forecast_lu", "filter". */
int sgdlm_set_timing(sgdlm_ctx *ctx, int enable);
int sgdlm_get_timing(const sgdlm_ctx *ctx, const char *stage, double *ms,
                     long *launches);
/* Number of draws of the last VB or forecast call whose LU needed row
 * interchanges (factored by a partial-pivoting kernel instead of the fast
 * path; every draw when the call's kernel kind pivots every draw). */
int sgdlm_get_lu_fallback_count(const sgdlm_ctx *ctx, long *count);
/* Workgroups / LDS bytes / panel width / workgroup size of the fast VB LU
 * kernel for the current structure and precision (nb = 0: every draw goes
 * through the partial-pivoting kernel). */
int sgdlm_get_lu_config(const sgdlm_ctx *ctx, int *grid, size_t *lds_bytes,
                        int *nb, int *threads);
/* The same for the forecast solves (computeForecast) at the current
 * precision (FP64 factors for SGDLM_PREC_FP64, else FP32 factors with one
 * FP64-residual refinement): nb = 0 means every draw is solved by the
 * partial-pivoting kernel; slots_in_workspace = 1 when the sparse prefix's
 * LDS slots wait in the global workspace during the dense LU (their LDS is
 * shared with the dense panel). */
int sgdlm_get_lu_solve_config(const sgdlm_ctx *ctx, int *grid, size_t *lds_bytes,
                              int *nb, int *threads, int *slots_in_workspace);
/* The forecast solves' kernel kind: SGDLM_LU_REGISTER (FP32 factors of
 * dense blocks of 96..288 rows in registers, written out per chunk of draws
 * and substituted by one wavefront per draw; the default where it fits),
 * SGDLM_LU_BLOCKED (LDS panels, substitutions in the factoring workgroup) or
 * SGDLM_LU_PIVOTING (every draw with row interchanges); pinning one
 * (SGDLM_LU_AUTO: the default).  No counterpart in the reference. */
int sgdlm_get_lu_solve_kernel(const sgdlm_ctx *ctx, int *kind);
int sgdlm_set_lu_solve_kernel(sgdlm_ctx *ctx, int kind);
/* Pins the fast VB LU kernel's workgroup size (256: two workgroups per CU,
 * 512: one) and panel width (16, 32, 64); 0 = automatic.  A configuration
 * that does not fit makes the next VB call fail with SGDLM_ERR_ARG. */
int sgdlm_set_lu_tuning(sgdlm_ctx *ctx, int threads, int nb);
/* Sparse prefix of the VB log-determinant LU: series whose degree in the
 * (symmetrised) parental graph stays <= max_front are eliminated by a
 * minimum-degree schedule before the dense LU of the remaining series
 * (default 32; 0 = dense LU of the whole I - Gamma).  The result is the
 * same determinant; only the rounding differs.  Unless set here, the limit
 * is raised in steps of 4 (up to 64) when that brings the dense block
 * under the size up to which the 64-wide fast LU fits its LDS (288 on
 * gfx950); a value set here is used as given. */
int sgdlm_set_lu_max_front(sgdlm_ctx *ctx, int max_front);
/* The plan for the current parental sets: the front limit it used, series
 * eliminated in the prefix, series in the dense LU, prefix waves and prefix
 * multiply-adds per draw. */
int sgdlm_get_lu_elimination(const sgdlm_ctx *ctx, int *max_front,
                             int *n_prefix, int *n_dense, int *n_waves,
                             long *prefix_fma);
/* The kernel that factors the dense block of the VB log-determinant LU
 * for the current structure and precision (SGDLM_LU_BLOCKED / _REGISTER /
 * _PIVOTING / _WAVE), the dimension it factors (the Schur complement after
 * the sparse prefix rounded up to 16; for _WAVE the whole matrix rounded up
 * to 32, 64 or 128) and the dimension the partial-pivoting kernel factors
 * for a draw that needs row interchanges (the Schur complement when the
 * prefix is re-eliminated, else the whole matrix rounded up to 16; _WAVE
 * pivots every draw itself).  No counterpart in the reference, where cuBLAS picks
 * its getrfBatched kernel (rSGDLM/src/SGDLM/SGDLM.cu:280-281). */
int sgdlm_get_lu_kernel(const sgdlm_ctx *ctx, int *kind, int *dense_dim,
                        int *pivot_dim);
/* Pins the kernel kind of the VB LU (SGDLM_LU_AUTO = plan / autotuner);
 * a kind that does not fit makes the next VB call fail with
 * SGDLM_ERR_ARG.  No counterpart in the reference. */
int sgdlm_set_lu_kernel(sgdlm_ctx *ctx, int kind);
/* Autotuner of the VB LU (default on): on the first VB call for a parental
 * structure and precision, every feasible configuration (kernel kind,
 * workgroup size, panel width, one wavefront per draw for small models,
 * and prefix front limits 24 / default / 48)
 * is timed on up to 4096 of that call's own draws, and the fastest is kept
 * for the structure, in a table shared by all contexts of the process
 * (keyed by a hash of m, max_p, the parental sets, precision and device).
 * It does not run while sgdlm_set_lu_tuning, sgdlm_set_lu_kernel or
 * sgdlm_set_lu_max_front pin a choice, nor in the allocation-free entry
 * point (which uses a configuration already tuned for its structure).
 * Results do not depend on it beyond rounding; processes that must agree
 * bit for bit either pin the configuration or turn tuning off (also for
 * every new context by setting the environment variable
 * SGDLM_LU_AUTOTUNE=0).  No counterpart in the reference. */
int sgdlm_set_lu_autotune(sgdlm_ctx *ctx, int enable);
/* Graph mode (default off): the device part of each VB call (Cholesky,
 * draws, LU log-determinants, weights, moments and moment-match, the copy
 * of the weight statistics) is captured once as a hipGraph and replayed on
 * later calls; the Philox epoch reaches the draw kernels through a device
 * word that a one-thread kernel sets before each replay, and the LU's
 * fallback counter is cleared by a kernel, so no node depends on the call.
 * A call re-captures when anything the graph baked in changed (draw count,
 * buffers, precision, seed, parental structure, LU plan); calls with stage
 * timing on, or whose LU configuration is still to be tuned, run the
 * stages directly.  Results are bit-identical to the direct path.  No
 * counterpart in the reference (SGDLM.cu:232-503 launches each stage). */
int sgdlm_set_graph_mode(sgdlm_ctx *ctx, int enable);
/* How many VB graphs this context captured and how many VB calls it ran
 * by replaying one. */
int sgdlm_get_graph_stats(const sgdlm_ctx *ctx, long *captures, long *replays);
/* The tuned configuration for the current structure (kind = SGDLM_LU_AUTO
 * when none is), the milliseconds the tuning took in this process (0 when
 * taken from the table), the number of configurations timed and whether it
 * came from the table. */
int sgdlm_get_lu_autotune(const sgdlm_ctx *ctx, int *kind, int *threads, int *nb,
                          int *max_front, double *tune_ms, int *candidates,
                          int *from_cache);

#ifdef __cplusplus
}
#endif

#endif
