/*
 * host_wrapper.cpp -- SGDLM::HostWrapper<double> over the C-ABI.
 *
 * The role of HostWrapperImpl<double> (rSGDLM/src/SGDLM/HostWrapperImpl.cu)
 * and HostWrapperFactory<double> (HostWrapperFactory.cu:5-35) for one GPU:
 * every method forwards to one sgdlm_* call.  Failures are logged to stderr
 * and the method returns, like the reference's myAssert / ERROR_LOGGER
 * (rSGDLM/src/logger.hpp:160-165); getters of a failed call return 0.
 */
#include <cstdio>

#include "sgdlm.h"
#include "SGDLM/HostWrapper.hpp"
#include "SGDLM/HostWrapperFactory.hpp"

namespace SGDLM {
namespace {

class HostWrapperImpl final : public HostWrapper<double> {
public:
  explicit HostWrapperImpl(std::size_t no_gpus) {
    if (sgdlm_create(&ctx_, no_gpus) != SGDLM_OK) {
      std::fprintf(stderr, "[SGDLM] ERROR: no HIP device available\n");
      ctx_ = nullptr;
    }
  }
  ~HostWrapperImpl() override { sgdlm_destroy(ctx_); }

  std::size_t getNoSeries() const override {
    std::size_t v = 0;
    check(sgdlm_get_no_series(ctx_, &v), "getNoSeries");
    return v;
  }
  std::size_t getMaxP() const override {
    std::size_t v = 0;
    check(sgdlm_get_max_p(ctx_, &v), "getMaxP");
    return v;
  }
  bool getEvolutionMatrixConfiguration() const override {
    int v = 0;
    check(sgdlm_get_evolution_matrix_configuration(ctx_, &v), "getEvolutionMatrixConfiguration");
    return v != 0;
  }
  void initMemory(std::size_t m, std::size_t max_p) override {
    check(sgdlm_init_memory(ctx_, m, max_p), "initMemory");
  }
  void clearMemory() override { check(sgdlm_clear_memory(ctx_), "clearMemory"); }
  void manageEvoMemory(bool use) override {
    check(sgdlm_manage_evo_memory(ctx_, use ? 1 : 0), "manageEvoMemory");
  }
  std::size_t getNSim() const override {
    std::size_t v = 0;
    check(sgdlm_get_nsim(ctx_, &v), "getNSim");
    return v;
  }
  std::size_t getNSimBatch() const override {
    std::size_t v = 0;
    check(sgdlm_get_nsim_batch(ctx_, &v), "getNSimBatch");
    return v;
  }
  void initSimMemory(std::size_t nsim, std::size_t nsim_batch) override {
    check(sgdlm_init_sim_memory(ctx_, nsim, nsim_batch), "initSimMemory");
  }
  void clearSimMemory() override { check(sgdlm_clear_sim_memory(ctx_), "clearSimMemory"); }
  void initForecastPriors() override {
    check(sgdlm_init_forecast_priors(ctx_), "initForecastPriors");
  }
  bool isPrior() const override {
    int v = 0;
    check(sgdlm_get_is_prior(ctx_, &v), "isPrior");
    return v != 0;
  }
  void isPrior(bool is_prior) override {
    check(sgdlm_set_is_prior(ctx_, is_prior ? 1 : 0), "isPrior");
  }
  void getParameters(double *m, double *C, double *n, double *s) const override {
    check(sgdlm_get_parameters(ctx_, m, C, n, s), "getParameters");
  }
  void setParameters(const double *m, const double *C, const double *n,
                     const double *s) override {
    check(sgdlm_set_parameters(ctx_, m, C, n, s), "setParameters");
  }
  void getDiscountFactors(double *beta, double *delta) const override {
    check(sgdlm_get_discount_factors(ctx_, beta, delta), "getDiscountFactors");
  }
  void setDiscountFactors(const double *beta, const double *delta) override {
    check(sgdlm_set_discount_factors(ctx_, beta, delta), "setDiscountFactors");
  }
  void getEvolutionMatrix(double *G) const override {
    check(sgdlm_get_evolution_matrix(ctx_, G), "getEvolutionMatrix");
  }
  void setEvolutionMatrix(const double *G) override {
    check(sgdlm_set_evolution_matrix(ctx_, G), "setEvolutionMatrix");
  }
  void getParentalSets(unsigned int *p, unsigned int *sp) const override {
    check(sgdlm_get_parental_sets(ctx_, p, sp), "getParentalSets");
  }
  void setParentalSets(const unsigned int *p, const unsigned int *sp) override {
    check(sgdlm_set_parental_sets(ctx_, p, sp), "setParentalSets");
  }
  void computePrior(bool evolve_forecast_variables) override {
    check(sgdlm_compute_prior(ctx_, nullptr, evolve_forecast_variables ? 1 : 0), "computePrior");
  }
  void computePrior(const double *G, bool evolve_forecast_variables) override {
    check(sgdlm_compute_prior(ctx_, G, evolve_forecast_variables ? 1 : 0), "computePrior");
  }
  void evolveForecastSamples() override {
    check(sgdlm_evolve_forecast_samples(ctx_, nullptr), "evolveForecastSamples");
  }
  void evolveForecastSamples(const double *G) override {
    check(sgdlm_evolve_forecast_samples(ctx_, G), "evolveForecastSamples");
  }
  void computeForecast(double *y, const double *x, double *lambdas, double *thetas,
                       bool use_existing) override {
    check(sgdlm_compute_forecast(ctx_, y, x, lambdas, thetas, use_existing ? 1 : 0),
          "computeForecast");
  }
  void computePosterior(const double *y, const double *F) override {
    check(sgdlm_compute_posterior(ctx_, y, F), "computePosterior");
  }
  void computeVBPosterior(double *m, double *C, double *n, double *s, double *w,
                          double *sum_det) override {
    check(sgdlm_compute_vb_posterior(ctx_, m, C, n, s, w, sum_det), "computeVBPosterior");
  }

private:
  bool check(int rc, const char *what) const {
    if (rc == SGDLM_OK)
      return true;
    std::fprintf(stderr, "[SGDLM] ERROR in %s: %s (status %d)\n", what,
                 ctx_ ? sgdlm_last_error(ctx_) : "no context", rc);
    return false;
  }
  sgdlm_ctx *ctx_ = nullptr;
};

} // namespace

template <>
HostWrapper<double> *HostWrapperFactory<double>::create(std::size_t no_gpus) {
  return new HostWrapperImpl(no_gpus);
}

template <>
HostWrapper<double> *HostWrapperFactory<double>::create(std::size_t no_gpus, std::size_t m,
                                                        std::size_t max_p) {
  HostWrapper<double> *w = new HostWrapperImpl(no_gpus);
  w->initMemory(m, max_p);
  return w;
}

template <>
HostWrapper<double> *HostWrapperFactory<double>::create(std::size_t no_gpus, std::size_t m,
                                                        std::size_t max_p,
                                                        bool use_state_evolution_matrix) {
  HostWrapper<double> *w = new HostWrapperImpl(no_gpus);
  w->initMemory(m, max_p);
  w->manageEvoMemory(use_state_evolution_matrix);
  return w;
}

} // namespace SGDLM
