/*
 * elim_check.cpp -- CPU execution of the elimination plan of
 * gpusgdlm_amd/csrc/lu_symbolic.cpp, in the order the LU kernel runs it
 * (slot initialisation, per wave multipliers then gathered updates, dense
 * Schur-complement assembly from the ELL pair blocks, dense LU without
 * interchanges), in FP64, reading every table through the same offsets the
 * kernel uses.  Lets the CPU tests check the symbolic tables against
 * LAPACK's slogdet.
 */
#include <cmath>
#include <vector>

#include "../../gpusgdlm_amd/csrc/lu_symbolic.h"

extern "C" int elim_check_logdet(int m, int n_ent, const int *row, const int *col,
                                 const int *idx, const double *theta, int max_front,
                                 int max_slots, double *logdet, double *amax, int *info) {
  std::vector<sgdlm::GammaEntry> ents(n_ent);
  for (int e = 0; e < n_ent; ++e)
    ents[e] = {row[e], col[e], idx[e]};
  sgdlm::ElimPlan P;
  sgdlm::build_elim_plan(m, ents, max_front, max_slots, P);
  const int *tb = P.tab.data();
  info[0] = P.n1;
  info[1] = P.n2;
  info[2] = P.M2;
  info[3] = P.ns;
  info[4] = P.n_waves;
  info[5] = P.n_blk;
  info[6] = (int)P.st_pairs.size();
  info[7] = (int)P.tab.size();

  std::vector<double> sl(P.ns);
  for (int s = 0; s < P.ns; ++s) {
    const int v = P.slot_init[s];
    sl[s] = v >= 0 ? -theta[v] : (v == sgdlm::SLOT_ONE ? 1.0 : v == sgdlm::SLOT_MINUS_ONE ? -1.0 : 0.0);
  }
  double mx = 0.0, ld = 0.0;
  int bad = 0;
  const uint32_t lo = 0xffffu;
  for (int w = 0; w < P.n_waves; ++w) {
    for (int it = tb[P.o_mptr + w]; it < tb[P.o_mptr + w + 1]; ++it) {
      const uint32_t x = (uint32_t)tb[P.o_mult + it];
      double d = sl[x >> 16];
      if (!(std::fabs(d) > 0)) {
        bad = 1;
        d = 1.0;
      }
      const double l = sl[x & lo] / d;
      sl[x & lo] = l;
      mx = std::fmax(mx, std::fabs(l));
    }
    for (int t = tb[P.o_tptr + w]; t < tb[P.o_tptr + w + 1]; ++t) {
      const int s = tb[P.o_tslot + t];
      double acc = sl[s];
      for (int e = tb[P.o_cptr + t]; e < tb[P.o_cptr + t + 1]; ++e) {
        const uint32_t c = (uint32_t)tb[P.o_contrib + e];
        acc = std::fma(-sl[c & lo], sl[c >> 16], acc);
      }
      sl[s] = acc;
    }
  }
  for (int q = 0; q < P.n1; ++q) {
    const double p = sl[tb[P.o_piv + q]];
    if (!(std::fabs(p) > 0))
      bad = 1;
    ld += std::log(std::fabs(p));
  }
  const int M2 = P.M2;
  std::vector<double> S((size_t)M2 * M2, 0.0);
  for (int r = 0; r < M2; ++r)
    S[(size_t)r * M2 + r] = 1.0;
  for (size_t e = 0; e < P.sg_off.size(); ++e)
    S[P.sg_off[e]] = -theta[P.sg_idx[e]];
  for (int b = 0; b < P.n_blk; ++b) {
    const int R = tb[P.o_brounds + b], cb = tb[P.o_bcbase + b];
    for (int l = 0; l < 64; ++l) {
      const int o = P.st_off[(size_t)b * 64 + l];
      if (o < 0)
        continue;
      double acc = 0.0;
      for (int r = 0; r < R; ++r) {
        const uint32_t c = P.st_pairs[(size_t)cb + 64 * r + l];
        acc = std::fma(-sl[c & lo], sl[c >> 16], acc);
      }
      S[o] = acc;
    }
  }
  for (int k = 0; k < M2; ++k) {
    double d = S[(size_t)k * M2 + k];
    if (!(std::fabs(d) > 0)) {
      bad = 1;
      d = 1.0;
    }
    ld += std::log(std::fabs(S[(size_t)k * M2 + k]));
    for (int i = k + 1; i < M2; ++i) {
      const double l = S[(size_t)i * M2 + k] / d;
      mx = std::fmax(mx, std::fabs(l));
      for (int j = k + 1; j < M2; ++j)
        S[(size_t)i * M2 + j] = std::fma(-l, S[(size_t)k * M2 + j], S[(size_t)i * M2 + j]);
    }
  }
  *logdet = ld;
  *amax = mx;
  return bad;
}

/* Solve (I - Gamma) y = b through the plan (prefix forward, dense LU of S
 * without interchanges, dense forward/backward, prefix backward), FP64. */
extern "C" int elim_check_solve(int m, int n_ent, const int *row, const int *col, const int *idx,
                                const double *theta, int max_front, const double *b, double *y) {
  std::vector<sgdlm::GammaEntry> ents(n_ent);
  for (int e = 0; e < n_ent; ++e)
    ents[e] = {row[e], col[e], idx[e]};
  sgdlm::ElimPlan P;
  sgdlm::build_elim_plan(m, ents, max_front, 65536, P);
  const int *tb = P.tab.data();
  const uint32_t lo = 0xffffu;
  std::vector<double> sl(P.ns);
  for (int s = 0; s < P.ns; ++s) {
    const int v = P.slot_init[s];
    sl[s] = v >= 0 ? -theta[v] : (v == sgdlm::SLOT_ONE ? 1.0 : v == sgdlm::SLOT_MINUS_ONE ? -1.0 : 0.0);
  }
  for (int w = 0; w < P.n_waves; ++w) {
    for (int it = tb[P.o_mptr + w]; it < tb[P.o_mptr + w + 1]; ++it) {
      const uint32_t x = (uint32_t)tb[P.o_mult + it];
      sl[x & lo] = sl[x & lo] / sl[x >> 16];
    }
    for (int t = tb[P.o_tptr + w]; t < tb[P.o_tptr + w + 1]; ++t) {
      const int s = tb[P.o_tslot + t];
      double acc = sl[s];
      for (int e = tb[P.o_cptr + t]; e < tb[P.o_cptr + t + 1]; ++e) {
        const uint32_t c = (uint32_t)tb[P.o_contrib + e];
        acc = std::fma(-sl[c & lo], sl[c >> 16], acc);
      }
      sl[s] = acc;
    }
  }
  const int M2 = P.M2, n1 = P.n1, n2 = P.n2;
  std::vector<double> S((size_t)M2 * M2, 0.0);
  for (int r = 0; r < M2; ++r)
    S[(size_t)r * M2 + r] = 1.0;
  for (size_t e = 0; e < P.sg_off.size(); ++e)
    S[P.sg_off[e]] = -theta[P.sg_idx[e]];
  for (int bl = 0; bl < P.n_blk; ++bl) {
    const int R = tb[P.o_brounds + bl], cb = tb[P.o_bcbase + bl];
    for (int l = 0; l < 64; ++l) {
      const int o = P.st_off[(size_t)bl * 64 + l];
      if (o < 0)
        continue;
      double acc = 0.0;
      for (int r = 0; r < R; ++r) {
        const uint32_t c = P.st_pairs[(size_t)cb + 64 * r + l];
        acc = std::fma(-sl[c & lo], sl[c >> 16], acc);
      }
      S[o] = acc;
    }
  }
  for (int k = 0; k < M2; ++k)
    for (int i = k + 1; i < M2; ++i) {
      const double l = S[(size_t)i * M2 + k] / S[(size_t)k * M2 + k];
      S[(size_t)i * M2 + k] = l;
      for (int j = k + 1; j < M2; ++j)
        S[(size_t)i * M2 + j] -= l * S[(size_t)k * M2 + j];
    }
  std::vector<double> x(n1 + M2, 0.0);
  for (int w = 0; w < P.n_waves; ++w)
    for (int q = P.wave_q[w]; q < P.wave_q[w + 1]; ++q) {
      double acc = b[P.node_of[q]];
      for (int e = P.fwd_ptr[q]; e < P.fwd_ptr[q + 1]; ++e)
        acc -= sl[P.fwd[e] & lo] * x[P.fwd[e] >> 16];
      x[q] = acc;
    }
  for (int s2 = 0; s2 < n2; ++s2) {
    double acc = b[P.node_of[n1 + s2]];
    for (int e = P.fwd_ptr[n1 + s2]; e < P.fwd_ptr[n1 + s2 + 1]; ++e)
      acc -= sl[P.fwd[e] & lo] * x[P.fwd[e] >> 16];
    x[n1 + s2] = acc;
  }
  double *xd = x.data() + n1;
  for (int i = 0; i < M2; ++i)
    for (int t = 0; t < i; ++t)
      xd[i] -= S[(size_t)i * M2 + t] * xd[t];
  for (int i = M2 - 1; i >= 0; --i) {
    for (int t = i + 1; t < M2; ++t)
      xd[i] -= S[(size_t)i * M2 + t] * xd[t];
    xd[i] /= S[(size_t)i * M2 + i];
  }
  for (int w = P.n_waves - 1; w >= 0; --w)
    for (int q = P.wave_q[w]; q < P.wave_q[w + 1]; ++q) {
      double acc = x[q];
      for (int e = P.bwd_ptr[q]; e < P.bwd_ptr[q + 1]; ++e)
        acc -= sl[P.bwd[e] & lo] * x[P.bwd[e] >> 16];
      x[q] = acc / sl[tb[P.o_piv + q]];
    }
  for (int i = 0; i < n1 + n2; ++i)
    y[P.node_of[i]] = x[i];
  return 0;
}
