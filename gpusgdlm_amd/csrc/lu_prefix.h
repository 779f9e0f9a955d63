/*
 * lu_prefix.h -- device code shared by the LU kernels: the sparse
 * elimination prefix of I - Gamma on LDS slots (tables of lu_symbolic.h).
 *
 * Replaces, for the series the minimum-degree plan eliminates first, the
 * leading columns of the reference's dense getrfBatched
 * (rSGDLM/src/SGDLM/SGDLM.cu:280-281) on fillGammas' matrices
 * (kernel_functions.cuh:821-842).
 */
#ifndef SGDLM_LU_PREFIX_H
#define SGDLM_LU_PREFIX_H

#include <hip/hip_runtime.h>
#include <math.h>

#include "lu_symbolic.h"

namespace sgdlm {

/* Threshold partial pivoting: the diagonal pivot of the no-interchange LU
 * kernels is accepted while every multiplier satisfies |l| <= 1 / tau with
 * tau = 1/2 (LAPACK's partial pivoting is tau = 1); element growth per step
 * stays below 1 + 1/tau = 3.  Draws that exceed it are refactored with
 * partial pivoting by k_lu_pivot. */
constexpr float kMaxMultiplier = 2.0f;

/* Multiplier bound of the sparse prefix when the partial-pivoting kernel
 * re-eliminates it for a flagged draw and then factors only the dense
 * Schur complement with row interchanges.  The prefix pivots are the unit
 * diagonal of I - Gamma modified by at most a few low-degree eliminations;
 * with |l| <= 64 the log-determinant matches that of partial pivoting on
 * the whole matrix to within the conditioning of the draw.  Beyond it (or
 * at a zero prefix pivot) the whole matrix is factored with interchanges. */
constexpr float kPrefixLooseMultiplier = 64.0f;

/* The prefix part of the plan as the kernels read it. */
struct PrefixTabs {
  int ns, waves, n1, tab_words;
  int o_piv, o_mptr, o_mult, o_tptr, o_tslot, o_cptr, o_contrib;
  const int *slot_init, *tab; /* global copies */
};

/* The barrier between the elimination's phases: the workgroup's, or (WAVE:
 * one wavefront eliminating a draw of its own, tid = its lane, nthreads =
 * 64) the wavefront's LDS ordering. */
template <bool WAVE> __device__ __forceinline__ void prefix_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

/* Loads the prefix tables into LDS tb[] (when load_tab), fills the slots
 * sl[] of one draw (theta of the draw: th) and eliminates the prefix wave by
 * wave: per wave the multipliers l_ik = a_ik / a_kk replace a_ik, then every
 * target slot subtracts its gathered sum of l_ik * a_kj (fixed order, no
 * atomics).  Adds log|a_kk| of the prefix pivots to ls (FP64); a zero pivot
 * or a multiplier beyond lbound (NaN included) sets bad and *eflag; amax
 * tracks max |l|.  Ends with a workgroup barrier (WAVE: the wavefront's LDS
 * ordering, with load_tab false and the tables already in LDS). */
template <typename T, bool WAVE = false>
__device__ void prefix_eliminate(T *sl, int *tb, const double *th, const PrefixTabs &p, int tid,
                                 int nthreads, bool load_tab, int &bad, T &amax, double &ls,
                                 int *eflag, T lbound) {
  constexpr unsigned LO = 0xffffu;
  /* the global reads below are issued GB at a time per thread: one round
   * trip per batch instead of one per element */
  constexpr int GB = 8;
  if (load_tab) {
    for (int w0 = tid; w0 < p.tab_words; w0 += GB * nthreads) {
      int x[GB];
#pragma unroll
      for (int u = 0; u < GB; ++u) {
        const int w = w0 + u * nthreads;
        x[u] = w < p.tab_words ? p.tab[w] : 0;
      }
#pragma unroll
      for (int u = 0; u < GB; ++u) {
        const int w = w0 + u * nthreads;
        if (w < p.tab_words)
          tb[w] = x[u];
      }
    }
  }
  for (int s0 = tid; s0 < p.ns; s0 += GB * nthreads) {
    int code[GB];
#pragma unroll
    for (int u = 0; u < GB; ++u) {
      const int s = s0 + u * nthreads;
      code[u] = s < p.ns ? p.slot_init[s] : SLOT_ZERO;
    }
    T x[GB];
#pragma unroll
    for (int u = 0; u < GB; ++u) {
      const int v = code[u];
      x[u] = v >= 0 ? (T)(-th[v]) : (v == SLOT_ONE ? (T)1 : (v == SLOT_MINUS_ONE ? (T)-1 : (T)0));
    }
#pragma unroll
    for (int u = 0; u < GB; ++u) {
      const int s = s0 + u * nthreads;
      if (s < p.ns)
        sl[s] = x[u];
    }
  }
  prefix_sync<WAVE>();
  const int *mptr = tb + p.o_mptr, *tptr = tb + p.o_tptr, *cptr = tb + p.o_cptr;
  const unsigned *mult = reinterpret_cast<const unsigned *>(tb + p.o_mult);
  const unsigned *contrib = reinterpret_cast<const unsigned *>(tb + p.o_contrib);
  const int *tslot = tb + p.o_tslot;
  for (int w = 0; w < p.waves; ++w) {
    const int m1 = mptr[w + 1];
    for (int it = mptr[w] + tid; it < m1; it += nthreads) {
      const unsigned x = mult[it];
      T d = sl[x >> 16];
      if (!(fabs(d) > (T)0)) {
        bad = 1;
        d = (T)1;
      }
      const T l = sl[x & LO] / d;
      sl[x & LO] = l;
      /* a multiplier beyond the bound, or NaN (which fmax would drop),
       * flags the draw through bad as well as the early-exit flag */
      if (!(fabs(l) <= lbound))
        bad = 1;
      amax = fmax(amax, fabs(l));
      if (bad)
        *eflag = 1;
    }
    prefix_sync<WAVE>();
    const int t1 = tptr[w + 1];
    for (int t = tptr[w] + tid; t < t1; t += nthreads) {
      const int s = tslot[t];
      T acc = sl[s];
      const int e1 = cptr[t + 1];
      for (int e = cptr[t]; e < e1; ++e) {
        const unsigned c = contrib[e];
        acc = fma(-sl[c & LO], sl[c >> 16], acc);
      }
      sl[s] = acc;
    }
    prefix_sync<WAVE>();
  }
  const int *piv = tb + p.o_piv;
  for (int q = tid; q < p.n1; q += nthreads) {
    const T pv = sl[piv[q]];
    if (!(fabs(pv) > (T)0)) {
      bad = 1;
      *eflag = 1;
    }
    ls += log(fabs((double)pv));
  }
  prefix_sync<WAVE>();
}

} // namespace sgdlm

#endif
