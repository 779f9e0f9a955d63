/*
 * lu_symbolic.cpp -- multiple-minimum-degree prefix and gather tables for
 * the log-determinant LU of I - Gamma (host C++).
 *
 * Plain C++17 with no HIP dependency, so the plan can also be built and
 * executed on the CPU by tests.
 */
#include "lu_symbolic.h"

#include <algorithm>
#include <map>
#include <unordered_map>

namespace sgdlm {
namespace {

/* Pivots of one wave may have a degree up to the current minimum + this. */
constexpr int kWaveSlack = 8;

inline int popcount64(uint64_t x) { return __builtin_popcountll(x); }

} // namespace

void build_elim_plan(int m, const std::vector<GammaEntry> &ents, int max_front,
                     int max_slots, ElimPlan &P) {
  P = ElimPlan();
  P.m = m;
  if (max_slots > 65536)
    max_slots = 65536;
  /* room for the constant slots and one base slot per Gamma entry */
  max_slots = std::max(0, max_slots - 3 - (int)ents.size());
  const int W = (m + 63) / 64;
  std::vector<uint64_t> adj((size_t)m * W, 0);
  auto row = [&](int i) { return adj.data() + (size_t)i * W; };
  for (const GammaEntry &e : ents) {
    if (e.row == e.col)
      continue;
    row(e.row)[e.col >> 6] |= 1ull << (e.col & 63);
    row(e.col)[e.row >> 6] |= 1ull << (e.row & 63);
  }
  std::vector<char> alive(m, 1);
  std::vector<int> elim_pos(m, -1);
  std::vector<std::vector<int>> R;     /* neighbours at elimination, sorted */
  std::vector<std::vector<int>> waves; /* indices into P.order            */
  long long slots = 0;

  if (max_front > 0) {
    std::vector<int> deg(m);
    std::vector<uint64_t> blocked(W);
    bool stop = false;
    while (!stop) {
      int dmin = 1 << 30;
      for (int i = 0; i < m; ++i) {
        if (!alive[i])
          continue;
        int d = 0;
        for (int w = 0; w < W; ++w)
          d += popcount64(row(i)[w]);
        deg[i] = d;
        dmin = std::min(dmin, d);
      }
      if (dmin > max_front)
        break;
      const int limit = std::min(max_front, dmin + kWaveSlack);
      std::vector<std::pair<int, int>> cand;
      for (int i = 0; i < m; ++i)
        if (alive[i] && deg[i] <= limit)
          cand.push_back({deg[i], i});
      std::sort(cand.begin(), cand.end());
      std::fill(blocked.begin(), blocked.end(), 0);
      std::vector<int> wave;
      for (const auto &c : cand) {
        const int k = c.second;
        if (blocked[k >> 6] >> (k & 63) & 1)
          continue;
        if (slots + 1 + 2LL * c.first > max_slots) {
          stop = true;
          break;
        }
        slots += 1 + 2LL * c.first;
        wave.push_back(k);
        for (int w = 0; w < W; ++w)
          blocked[w] |= row(k)[w];
        blocked[k >> 6] |= 1ull << (k & 63);
      }
      if (wave.empty())
        break;
      std::vector<int> widx;
      for (int k : wave) {
        std::vector<int> nb;
        for (int w = 0; w < W; ++w)
          for (uint64_t x = row(k)[w]; x; x &= x - 1)
            nb.push_back(64 * w + __builtin_ctzll(x));
        /* clique on the neighbours, k leaves the graph */
        for (int a : nb) {
          uint64_t *ra = row(a);
          for (int w = 0; w < W; ++w)
            ra[w] |= row(k)[w];
          ra[a >> 6] &= ~(1ull << (a & 63));
          ra[k >> 6] &= ~(1ull << (k & 63));
        }
        std::fill(row(k), row(k) + W, 0);
        alive[k] = 0;
        elim_pos[k] = (int)P.order.size();
        widx.push_back((int)P.order.size());
        P.order.push_back(k);
        R.push_back(std::move(nb));
      }
      waves.push_back(std::move(widx));
    }
  }

  P.n1 = (int)P.order.size();
  P.n2 = m - P.n1;
  P.M2 = (P.n2 + 15) / 16 * 16;
  P.n_waves = (int)waves.size();
  P.sep_pos.assign(m, -1);
  {
    int s = 0;
    for (int i = 0; i < m; ++i)
      if (alive[i])
        P.sep_pos[i] = s++;
  }

  /* slots: per prefix pivot k its diagonal, column (i,k) and row (k,j) */
  std::unordered_map<uint64_t, int> slot_of;
  auto key = [](int i, int j) { return ((uint64_t)(uint32_t)i << 32) | (uint32_t)j; };
  std::vector<int> init;
  std::vector<int> piv(P.n1);
  for (int q = 0; q < P.n1; ++q) {
    const int k = P.order[q];
    piv[q] = (int)init.size();
    slot_of[key(k, k)] = (int)init.size();
    init.push_back(SLOT_ONE);
    for (int i : R[q]) {
      slot_of[key(i, k)] = (int)init.size();
      init.push_back(SLOT_ZERO);
    }
    for (int j : R[q]) {
      slot_of[key(k, j)] = (int)init.size();
      init.push_back(SLOT_ZERO);
    }
  }
  const int s_zero = (int)init.size();
  init.push_back(SLOT_ZERO);
  const int s_one = (int)init.size();
  init.push_back(SLOT_ONE);
  const int s_mone = (int)init.size();
  init.push_back(SLOT_MINUS_ONE);
  std::unordered_map<int, int> sg_of; /* S offset -> theta index */
  for (const GammaEntry &e : ents) {
    if (P.sep_pos[e.row] >= 0 && P.sep_pos[e.col] >= 0) {
      const int o = P.sep_pos[e.row] * P.M2 + P.sep_pos[e.col];
      P.sg_off.push_back(o);
      P.sg_idx.push_back(e.idx);
      sg_of[o] = e.idx;
    } else {
      init[slot_of.at(key(e.row, e.col))] = e.idx;
    }
  }

  /* wave programs and the Schur-complement pairs */
  std::vector<int> mptr{0}, tptr{0}, tslot, cptr{0};
  std::vector<uint32_t> mult, contrib;
  std::map<int, std::vector<uint32_t>> st;
  for (const auto &wv : waves) {
    std::map<int, std::vector<uint32_t>> tg;
    for (int q : wv) {
      const int k = P.order[q];
      const uint32_t dsl = (uint32_t)piv[q];
      for (int i : R[q])
        mult.push_back((uint32_t)slot_of.at(key(i, k)) | dsl << 16);
      for (int i : R[q]) {
        const uint32_t ls = (uint32_t)slot_of.at(key(i, k));
        for (int j : R[q]) {
          const uint32_t c = ls | (uint32_t)slot_of.at(key(k, j)) << 16;
          if (P.sep_pos[i] >= 0 && P.sep_pos[j] >= 0) {
            const int o = P.sep_pos[i] * P.M2 + P.sep_pos[j];
            auto &v = st[o];
            if (v.empty()) {
              /* first pair: (-1) * A22_ij */
              int base = s_zero;
              if (i == j) {
                base = s_one;
              } else {
                auto g = sg_of.find(o);
                if (g != sg_of.end()) {
                  base = (int)init.size();
                  init.push_back(g->second);
                }
              }
              v.push_back((uint32_t)s_mone | (uint32_t)base << 16);
            }
            v.push_back(c);
          } else {
            tg[slot_of.at(key(i, j))].push_back(c);
          }
        }
      }
    }
    for (const auto &t : tg) {
      tslot.push_back(t.first);
      contrib.insert(contrib.end(), t.second.begin(), t.second.end());
      cptr.push_back((int)contrib.size());
    }
    mptr.push_back((int)mult.size());
    tptr.push_back((int)tslot.size());
  }
  P.ns = (int)init.size();

  /* Schur targets: sorted by number of pairs (descending), blocks of 64 */
  std::vector<std::pair<int, int>> order_t; /* (-count, offset) */
  for (const auto &t : st)
    order_t.push_back({-(int)t.second.size(), t.first});
  std::sort(order_t.begin(), order_t.end());
  P.n_blk = ((int)order_t.size() + 63) / 64;
  const uint32_t zero_pair = (uint32_t)s_zero | (uint32_t)s_zero << 16;
  std::vector<int> brounds(P.n_blk), bcbase(P.n_blk);
  P.st_off.assign((size_t)P.n_blk * 64, -1);
  for (int b = 0; b < P.n_blk; ++b) {
    const int R0 = -order_t[(size_t)b * 64].first;
    brounds[b] = R0;
    bcbase[b] = (int)P.st_pairs.size();
    P.st_pairs.resize(P.st_pairs.size() + (size_t)R0 * 64, zero_pair);
    for (int l = 0; l < 64; ++l) {
      const size_t t = (size_t)b * 64 + l;
      if (t >= order_t.size())
        break;
      const int o = order_t[t].second;
      P.st_off[t] = o;
      const auto &v = st[o];
      for (size_t r = 0; r < v.size(); ++r)
        P.st_pairs[bcbase[b] + r * 64 + l] = v[r];
    }
  }

  /* LDS table image */
  auto put = [&](int &o, const auto &v) {
    o = (int)P.tab.size();
    for (auto x : v)
      P.tab.push_back((int)x);
  };
  P.slot_init = init;
  P.zero_pair = zero_pair;
  put(P.o_piv, piv);
  put(P.o_mptr, mptr);
  put(P.o_mult, mult);
  put(P.o_tptr, tptr);
  put(P.o_tslot, tslot);
  put(P.o_cptr, cptr);
  put(P.o_contrib, contrib);
  put(P.o_brounds, brounds);
  put(P.o_bcbase, bcbase);
  long long pairs = 0;
  for (const auto &t : st)
    pairs += (long long)t.second.size() - 1;
  P.prefix_fma = (long long)contrib.size() + (long long)mult.size() + pairs;

  /* solve tables */
  std::vector<int> xidx(m, -1); /* series -> x index */
  for (int q = 0; q < P.n1; ++q)
    xidx[P.order[q]] = q;
  P.node_of.assign(P.n1 + P.n2, 0);
  for (int q = 0; q < P.n1; ++q)
    P.node_of[q] = P.order[q];
  for (int i = 0; i < m; ++i)
    if (P.sep_pos[i] >= 0) {
      xidx[i] = P.n1 + P.sep_pos[i];
      P.node_of[P.n1 + P.sep_pos[i]] = i;
    }
  P.wave_q.assign(1, 0);
  for (const auto &wv : waves)
    P.wave_q.push_back(P.wave_q.back() + (int)wv.size());
  std::vector<std::vector<uint32_t>> fl(P.n1 + P.n2);
  P.bwd_ptr.assign(1, 0);
  for (int q = 0; q < P.n1; ++q) {
    const int k = P.order[q];
    for (int i : R[q])
      fl[xidx[i]].push_back((uint32_t)slot_of.at(key(i, k)) | (uint32_t)q << 16);
    for (int j : R[q])
      P.bwd.push_back((uint32_t)slot_of.at(key(k, j)) | (uint32_t)xidx[j] << 16);
    P.bwd_ptr.push_back((int)P.bwd.size());
  }
  P.fwd_ptr.assign(1, 0);
  for (const auto &v : fl) {
    P.fwd.insert(P.fwd.end(), v.begin(), v.end());
    P.fwd_ptr.push_back((int)P.fwd.size());
  }
}

void build_band_plan(const ElimPlan &P, int BR, int stride, int rows, BandPlan &B) {
  B = BandPlan();
  B.BR = BR;
  B.stride = stride;
  B.n_band = rows > 0 ? (rows + BR - 1) / BR : 0;
  const int *brounds = P.tab.data() + P.o_brounds, *bcbase = P.tab.data() + P.o_bcbase;
  /* per band: (pair list, band-local offset) of every Schur target */
  std::vector<std::vector<std::pair<std::vector<uint32_t>, int>>> tg(B.n_band);
  for (int b = 0; b < P.n_blk; ++b)
    for (int l = 0; l < 64; ++l) {
      const int o = P.st_off[(size_t)b * 64 + l];
      if (o < 0)
        continue;
      const int r = o / P.M2, c = o % P.M2;
      std::vector<uint32_t> v;
      for (int q = 0; q < brounds[b]; ++q) {
        const uint32_t x = P.st_pairs[(size_t)bcbase[b] + (size_t)q * 64 + l];
        if (q == 0 || x != P.zero_pair)
          v.push_back(x);
      }
      tg[r / BR].push_back({std::move(v), (r % BR) * stride + c});
    }
  B.blk_ptr.assign(1, 0);
  for (int band = 0; band < B.n_band; ++band) {
    auto &t = tg[band];
    std::stable_sort(t.begin(), t.end(), [](const auto &x, const auto &y) {
      return x.first.size() > y.first.size();
    });
    for (size_t t0 = 0; t0 < t.size(); t0 += 64) {
      const int R0 = (int)t[t0].first.size();
      B.brounds.push_back(R0);
      B.bcbase.push_back((int)B.st_pairs.size());
      B.st_pairs.resize(B.st_pairs.size() + (size_t)R0 * 64, P.zero_pair);
      const size_t b0 = B.st_off.size();
      B.st_off.resize(b0 + 64, -1);
      for (int l = 0; l < 64 && t0 + l < t.size(); ++l) {
        const auto &e = t[t0 + l];
        B.st_off[b0 + l] = e.second;
        for (size_t q = 0; q < e.first.size(); ++q)
          B.st_pairs[(size_t)B.bcbase.back() + q * 64 + l] = e.first[q];
      }
    }
    B.blk_ptr.push_back((int)B.brounds.size());
  }
  B.n_blk = (int)B.brounds.size();
  /* Gamma entries of S by band */
  std::vector<std::vector<std::pair<int, int>>> sg(B.n_band);
  for (size_t e = 0; e < P.sg_off.size(); ++e) {
    const int r = P.sg_off[e] / P.M2, c = P.sg_off[e] % P.M2;
    sg[r / BR].push_back({(r % BR) * stride + c, P.sg_idx[e]});
  }
  B.sg_ptr.assign(1, 0);
  for (int band = 0; band < B.n_band; ++band) {
    for (const auto &e : sg[band]) {
      B.sg_off.push_back(e.first);
      B.sg_idx.push_back(e.second);
    }
    B.sg_ptr.push_back((int)B.sg_off.size());
  }
}

} // namespace sgdlm
