/*
 * k_lu_wave.hip -- log|det(I - Gamma_i)| of every importance draw for small
 * models, gfx950: one wavefront per draw, the whole M x M matrix (m rounded
 * up to 32, 64 or 128; identity padding) in registers, one row per lane
 * (two for M = 128), LU with partial pivoting.
 *
 * Replaces fillGammas + cublasXgetrfBatched + detFromLU of the VB loop
 * (rSGDLM/src/SGDLM/SGDLM.cu:243-298, kernel_functions.cuh:821-842,904-926)
 * where the matrices are too small for a workgroup per draw to pay: no
 * workgroup barriers, no LDS, no fallback pass (every draw is pivoted).
 *
 * Per draw, lane i builds row i of I - Gamma from the row's CSR entries
 * (compile-time column selects, so the row stays in registers).  Column k:
 *   - pivot: the largest |a_ik| over the rows not yet used (a wavefront max,
 *     then the lowest lane holding it: a ballot);
 *   - the pivot row is broadcast one entry at a time (readlane with the
 *     pivot lane in a scalar register);
 *   - every other unused row subtracts l_i times it, l_i = a_ik / a_pk.
 * The columns are taken in chunks of 16 whose lower bound is a
 * compile-time constant, so each chunk's updates touch only the columns
 * right of it (half the work of a full-width loop) while the code stays
 * small.  Pivot k stays in lane k % 64; log|det| = sum of log|pivot| in
 * FP64, one log per lane and a fixed-order wavefront sum at the end; a zero
 * pivot gives -inf (weight 0), as in k_lu_pivot.
 */
#include <hip/hip_runtime.h>

#include <math.h>

#include "sgdlm_internal.h"

namespace sgdlm {
namespace {

struct WaveArgs {
  long n_draws;
  int m;
  long P;
  const double *theta;
  const int *g_rowptr, *g_col, *g_idx;
  double *logdet;
};

template <typename T> __device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const T y = __shfl_xor(v, o);
    v = y > v ? y : v;
  }
  return v;
}

__device__ __forceinline__ float bcast(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double bcast(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <typename T, int M, bool GRP>
__global__ __launch_bounds__(256, 1) void k_lu_wave(WaveArgs a0, const WaveArgs *__restrict__ tab,
                                                   GrpMap gm) {
  constexpr int R = (M + 63) / 64; /* rows per lane */
  constexpr int G = 8;             /* pivot-row entries read per batch */
  const int lane = threadIdx.x & 63;
  int bid, nbk;
  const WaveArgs a = grp_args<GRP>(a0, tab, gm, bid, nbk);
  const long wave0 = ((long)bid * blockDim.x + threadIdx.x) >> 6;
  const long nwaves = ((long)nbk * blockDim.x) >> 6;
  for (long d = wave0; d < a.n_draws; d += nwaves) {
    const double *th = a.theta + (size_t)d * a.P;
    T row[R][M];
    bool used[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const int i = lane + 64 * s;
      used[s] = i >= M; /* lanes beyond M take no part */
#pragma unroll
      for (int c = 0; c < M; ++c)
        row[s][c] = (c == i) ? (T)1 : (T)0;
      if (i < a.m) {
        const int e1 = a.g_rowptr[i + 1];
        for (int e = a.g_rowptr[i]; e < e1; ++e) {
          const int col = a.g_col[e];
          const T v = (T)(-th[a.g_idx[e]]);
#pragma unroll
          for (int c = 0; c < M; ++c)
            row[s][c] = (c == col) ? v : row[s][c];
        }
      }
    }

    /* pivot k is kept by lane k % 64 (slot k / 64); logs taken at the end */
    T pv[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      pv[s] = (T)1;
    bool singular = false;
#pragma unroll
    for (int kb = 0; kb < M; kb += 16) {
      for (int k = kb; k < kb + 16; ++k) {
        /* this lane's candidate |a_ik| over its unused rows */
        T best = (T)-1;
        int bslot = 0;
        T ak[R];
#pragma unroll
        for (int s = 0; s < R; ++s) {
          T x = (T)0;
#pragma unroll
          for (int c = kb; c < kb + 16; ++c)
            x = (c == k) ? row[s][c] : x;
          ak[s] = x;
          T v = fabs(x);
          if (used[s] || !(v >= (T)0))
            v = (T)-1;
          if (v > best) {
            best = v;
            bslot = s;
          }
        }
        /* pivot: largest candidate, lowest lane holding it */
        const T mx = wave_max(best);
        const unsigned long long hit = __ballot(best == mx);
        const int p = __builtin_amdgcn_readfirstlane(__builtin_ctzll(hit));
        const int ps = R == 1 ? 0 : __builtin_amdgcn_readfirstlane(__builtin_amdgcn_readlane(bslot, p));
        T piv = bcast(R == 1 ? ak[0] : (ps ? ak[R - 1] : ak[0]), p);
        if (!(fabs(piv) > (T)0)) {
          singular = true;
          piv = (T)1;
        }
#pragma unroll
        for (int s = 0; s < R; ++s)
          if (lane + 64 * s == k)
            pv[s] = piv;
#pragma unroll
        for (int s = 0; s < R; ++s)
          if (s == ps && lane == p)
            used[s] = true;
        T l[R];
#pragma unroll
        for (int s = 0; s < R; ++s)
          l[s] = used[s] ? (T)0 : ak[s] / piv;
        /* the pivot row's entries right of column k, G at a time (the
         * readlanes of a batch cover each other's latency); zero at and
         * left of k so the update leaves those columns unchanged */
#pragma unroll
        for (int c0 = kb; c0 < M; c0 += G) {
          T pc[G];
#pragma unroll
          for (int u = 0; u < G; ++u) {
            const int c = c0 + u;
            pc[u] = bcast(R == 1 ? row[0][c] : (ps ? row[R - 1][c] : row[0][c]), p);
          }
#pragma unroll
          for (int u = 0; u < G; ++u) {
            const int c = c0 + u;
            const T v = (c >= kb + 16 || c > k) ? pc[u] : (T)0;
#pragma unroll
            for (int s = 0; s < R; ++s)
              row[s][c] = fma(-l[s], v, row[s][c]);
          }
        }
      }
    }
    /* log|det| = sum of log|pivot|: one log per lane and slot, then a
     * fixed-order wavefront sum */
    double ls = 0.0;
#pragma unroll
    for (int s = 0; s < R; ++s)
      ls += log(fabs((double)pv[s]));
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
      ls += __shfl_xor(ls, o);
    if (lane == 0)
      a.logdet[d] = singular ? -INFINITY : ls;
  }
}

template <typename T, int M> int launch_wave(const WaveArgs &a, int grid, hipStream_t st) {
  const WaveArgs *tab = nullptr;
  hipLaunchKernelGGL((k_lu_wave<T, M, false>), dim3(grid), dim3(256), 0, st, a, tab, GrpMap());
  return hipGetLastError() == hipSuccess ? SGDLM_OK : SGDLM_ERR_DEVICE;
}

template <typename T, int M> int resident_blocks() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void *)k_lu_wave<T, M, false>, 256, 0) !=
          hipSuccess ||
      n < 1) {
    (void)hipGetLastError();
    n = 1;
  }
  return n;
}

} // namespace

int lu_wave_dim(int M, int prec) {
  if (M <= 32)
    return 32;
  if (M <= 64)
    return 64;
  if (M <= 128 && prec != SGDLM_PREC_FP64)
    return 128;
  return 0;
}

int lu_wave_grid(int wave_M, int prec, int n_cu) {
  int per_cu = 1;
  const bool f64 = prec == SGDLM_PREC_FP64;
  if (wave_M == 32)
    per_cu = f64 ? resident_blocks<double, 32>() : resident_blocks<float, 32>();
  else if (wave_M == 64)
    per_cu = f64 ? resident_blocks<double, 64>() : resident_blocks<float, 64>();
  else
    per_cu = resident_blocks<float, 128>();
  return std::max(1, n_cu) * per_cu;
}

namespace {
WaveArgs wave_args(const Structure &S, long N, const double *theta, double *logdet) {
  WaveArgs a;
  a.n_draws = N;
  a.m = S.m;
  a.P = S.P;
  a.theta = theta;
  a.g_rowptr = S.g_rowptr;
  a.g_col = S.g_col;
  a.g_idx = S.g_idx;
  a.logdet = logdet;
  return a;
}
} // namespace

int lu_grp_wave(const Structure &S, const LuPlan &plan, long N, const double *theta,
                double *logdet, LuGrpKernel *k) {
  const WaveArgs a = wave_args(S, N, theta, logdet);
  const bool f64 = plan.prec == SGDLM_PREC_FP64;
  k->fn = nullptr;
  switch (plan.wave_M) {
  case 32:
    k->fn = f64 ? (const void *)k_lu_wave<double, 32, true> : (const void *)k_lu_wave<float, 32, true>;
    break;
  case 64:
    k->fn = f64 ? (const void *)k_lu_wave<double, 64, true> : (const void *)k_lu_wave<float, 64, true>;
    break;
  case 128:
    if (!f64)
      k->fn = (const void *)k_lu_wave<float, 128, true>;
    break;
  }
  if (!k->fn)
    return SGDLM_ERR_ARG;
  k->threads = 256;
  k->per_cu = std::max(1, plan.wave_grid / std::max(1, plan.grid));
  k->lds = 0;
  k->ws_wg = 0;
  k->units = (N + 3) / 4; /* four wavefronts (draws in flight) per workgroup */
  /* ms per draw on one workgroup of four wavefronts (size model) */
  const double Md = (double)plan.wave_M;
  k->cost = (double)N * 2.5e-5 * Md * (1.0 + Md / 64.0);
  k->args.assign(reinterpret_cast<const unsigned char *>(&a),
                 reinterpret_cast<const unsigned char *>(&a) + sizeof(a));
  k->work_field = -1;
  return SGDLM_OK;
}

int launch_lu_wave(const Structure &S, const LuPlan &plan, long N, const double *theta,
                   double *logdet, hipStream_t st) {
  const WaveArgs a = wave_args(S, N, theta, logdet);
  const long need = (N + 3) / 4; /* four wavefronts per workgroup */
  const int grid = (int)std::max(1L, std::min<long>(plan.wave_grid, need));
  const bool f64 = plan.prec == SGDLM_PREC_FP64;
  switch (plan.wave_M) {
  case 32:
    return f64 ? launch_wave<double, 32>(a, grid, st) : launch_wave<float, 32>(a, grid, st);
  case 64:
    return f64 ? launch_wave<double, 64>(a, grid, st) : launch_wave<float, 64>(a, grid, st);
  case 128:
    if (!f64)
      return launch_wave<float, 128>(a, grid, st);
    break;
  }
  return SGDLM_ERR_ARG;
}

} // namespace sgdlm
