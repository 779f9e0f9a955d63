/*
 * host_wrapper_driver.cpp -- drives SGDLM::HostWrapper<double> the way the
 * reference's Rcpp layer does (rSGDLM/src/SGDLM/RcppWrapper.cpp: constructor,
 * setSimultaneousParents, setDiscountFactors, setPriorParameters,
 * computePosterior, computePrior, computeVBPosterior, computeForecastDebug)
 * and writes every result to a binary file, for comparison with the Python
 * binding of the same engine.
 *
 * Usage: host_wrapper_driver IN OUT
 *   IN  (little-endian): u64 m, max_p, T, nsim;
 *       u32 p[m], sp[max_p*m]; f64 beta[m], delta[max_p^2*m],
 *       m0[max_p*m], C0[max_p^2*m], n0[m], s0[m];
 *       then per t < T: f64 y[m], F[max_p*m]; then f64 F_tp1[max_p*m].
 *   OUT: f64 m_T, C_T, n_T, s_T (posterior at T-1),
 *        VB m, C, n, s, IS weights[nsim], sum_det,
 *        prior m, C, n, s after computePrior,
 *        forecast y[nsim*m], lambdas[nsim*m], thetas[nsim*m*max_p],
 *        u64 getNoSeries, getMaxP, getNSim, getNSimBatch, isPrior.
 */
#include <cstdint>
#include <cstdio>
#include <memory>
#include <vector>

#include "SGDLM/HostWrapperFactory.hpp"

namespace {

template <typename T> bool rd(std::FILE *f, std::vector<T> &v, std::size_t n) {
  v.resize(n);
  return std::fread(v.data(), sizeof(T), n, f) == n;
}
template <typename T> void wr(std::FILE *f, const std::vector<T> &v) {
  std::fwrite(v.data(), sizeof(T), v.size(), f);
}

} // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
    return 2;
  }
  std::FILE *in = std::fopen(argv[1], "rb");
  if (!in) {
    std::perror(argv[1]);
    return 2;
  }
  std::vector<std::uint64_t> hdr;
  if (!rd(in, hdr, 4)) {
    std::fprintf(stderr, "short header\n");
    return 2;
  }
  const std::size_t m = hdr[0], max_p = hdr[1], T = hdr[2], nsim = hdr[3];
  const std::size_t np = m * max_p, npp = m * max_p * max_p;
  std::vector<unsigned int> p, sp;
  std::vector<double> beta, delta, m0, C0, n0, s0, Fn;
  bool ok = rd(in, p, m) && rd(in, sp, np) && rd(in, beta, m) && rd(in, delta, npp) &&
            rd(in, m0, np) && rd(in, C0, npp) && rd(in, n0, m) && rd(in, s0, m);
  std::vector<std::vector<double>> y(T), F(T);
  for (std::size_t t = 0; ok && t < T; ++t)
    ok = rd(in, y[t], m) && rd(in, F[t], np);
  ok = ok && rd(in, Fn, np);
  std::fclose(in);
  if (!ok) {
    std::fprintf(stderr, "short input\n");
    return 2;
  }

  std::unique_ptr<SGDLM::HostWrapper<double>> w(
      SGDLM::HostWrapperFactory<double>::create(1, m, max_p, false));
  w->setParentalSets(p.data(), sp.data());
  w->setDiscountFactors(beta.data(), delta.data());
  w->setParameters(m0.data(), C0.data(), n0.data(), s0.data());
  w->isPrior(true);
  for (std::size_t t = 0; t < T; ++t) {
    w->computePosterior(y[t].data(), F[t].data());
    if (t + 1 < T)
      w->computePrior(false);
  }
  std::vector<double> pm(np), pC(npp), pn(m), ps(m);
  w->getParameters(pm.data(), pC.data(), pn.data(), ps.data());

  w->initSimMemory(nsim, nsim);
  std::vector<double> vm(np), vC(npp), vn(m), vs(m), iw(nsim), sd(1);
  w->computeVBPosterior(vm.data(), vC.data(), vn.data(), vs.data(), iw.data(), sd.data());

  w->computePrior(false);
  std::vector<double> qm(np), qC(npp), qn(m), qs(m);
  w->getParameters(qm.data(), qC.data(), qn.data(), qs.data());

  w->initForecastPriors();
  std::vector<double> fy(nsim * m), fl(nsim * m), ft(nsim * m * max_p);
  w->computeForecast(fy.data(), Fn.data(), fl.data(), ft.data(), false);

  std::vector<std::uint64_t> meta = {w->getNoSeries(), w->getMaxP(), w->getNSim(),
                                     w->getNSimBatch(), w->isPrior() ? 1u : 0u};
  std::FILE *out = std::fopen(argv[2], "wb");
  if (!out) {
    std::perror(argv[2]);
    return 2;
  }
  for (const auto *v : {&pm, &pC, &pn, &ps, &vm, &vC, &vn, &vs, &iw, &sd, &qm, &qC, &qn, &qs,
                        &fy, &fl, &ft})
    wr(out, *v);
  wr(out, meta);
  std::fclose(out);
  std::printf("host_wrapper_driver: m=%zu max_p=%zu T=%zu nsim=%zu done\n", m, max_p, T, nsim);
  return 0;
}
