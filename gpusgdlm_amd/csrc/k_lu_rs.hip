/*
 * k_lu_rs.hip -- log|det(I - Gamma_i)| of every importance draw with the
 * dense Schur complement held in registers (FP32 and BF16x3 updates),
 * gfx950.
 *
 * Replaces fillGammas + cublasXgetrfBatched + detFromLU of the VB loop
 * (rSGDLM/src/SGDLM/SGDLM.cu:243-298, kernel_functions.cuh:821-842,904-926)
 * for dense blocks of up to 288 x 288 after the sparse prefix.
 *
 * One 512-thread workgroup per CU, persistent over draws.  Per draw:
 *   1. the sparse prefix (lu_prefix.h) on LDS slots;
 *   2. the Schur complement S (16 MT x 16 MT, identity padding) assembled
 *      through an LDS staging band of BR rows at a time (zero fill, the
 *      entries of I - Gamma, then every prefix-updated entry as the sum over
 *      its slot pairs, ELL blocks of 64) and loaded into registers: the 8
 *      wavefronts form a 2 x 4 grid and wavefront (wr, wc) keeps the 16 x 16
 *      tiles (I, J) with I = wr (mod 2), J = wc (mod 4) as MFMA accumulators
 *      (lane l holds rows 4 (l / 16) + r, r = 0..3, of column l % 16);
 *   3. a right-looking LU over 16-wide block columns k with no row
 *      interchanges, S never leaving the chip:
 *        - the owner of tile (k, k) factors it (one row per lane, pivot rows
 *          broadcast with readlane) and forms L_kk^-1 and U_kk^-1 (column
 *          substitutions, one column per lane);
 *        - U_kJ = L_kk^-1 A_kJ by the owners of row k (the accumulator tile is
 *          the MFMA B operand directly); L_Ik = A_Ik U_kk^-1 for the column
 *          tiles staged in LDS, spread over the other four wavefronts;
 *        - A_IJ -= L_Ik U_kJ on MFMA into the accumulators: FP32
 *          v_mfma_f32_16x16x4f32, or BF16x3 (hi*hi + hi*lo + lo*hi of
 *          v_mfma_f32_16x16x16bf16_1k, FP32 accumulation);
 *      the next block column is updated first and its diagonal tile
 *      factored while the other wavefronts finish the trailing update, so
 *      every step costs two workgroup barriers;
 *   4. log|det| = sum of log|a_kk| over the prefix and dense pivots in FP64,
 *      reduced in a fixed order.
 * Threshold partial pivoting as in k_lu_fast: a draw with a multiplier
 * |l| > 2 (diagonal-tile or L_Ik entries), a zero or NaN pivot is appended
 * to the device list that k_lu_pivot refactors with partial pivoting.
 */
#include <hip/hip_runtime.h>

#include <math.h>

#include <type_traits>
#include <utility>

#include "lu_prefix.h"
#include "sgdlm_internal.h"

namespace sgdlm {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

enum { RS_FP32 = 0, RS_BF16X3 = 1 };

constexpr int TR = 20;         /* LDS row stride of a staged tile (floats) */
constexpr int TSF = 16 * TR;   /* floats per staged tile                   */

struct RsArgs {
  int MT;                      /* tiles per dimension of S                 */
  long n_draws;
  const double *theta;
  long P;
  double *logdet;
  int *flag_count, *flag_list;
  PrefixTabs pre;
  int pre_base, slot_bytes, tab_persist;
  int flag_base, red_base;
  int x_base, l_base, u_base, lu_base, linv_base, uinvt_base, urev_base, piv_base;
  int stage_base, band_rows, n_band;
  const int *blk_ptr, *st_off, *brounds, *bcbase, *sg_ptr, *sg_off, *sg_idx;
  const unsigned *st_pairs;
  unsigned zero_pair;
  /* solves (SOLVE): per draw, at sol + d sol_stride, the MT x MT tiles of
   * the factors (tile (I, J) at (I MT + J) 256 floats, row-major: L_IJ
   * below, U_IJ above the diagonal, the diagonal tile packed with L_kk^-1
   * strictly below and U_kk^-1 on and above its diagonal), then the prefix
   * slots; sol_ok[d] = 1 for a factored draw, 0 for one left to the
   * partial-pivoting kernel */
  float *sol;
  long sol_stride;
  int *sol_ok;
  /* PIV: LDS byte offset of the row-exchange buffer (32 rows x the
   * trailing column tiles) */
  int swap_base;
};

__device__ __forceinline__ f32x4 mfma_f32(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma_bf16(s16x4 a, s16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

/* float -> bf16 bits, round to nearest even (finite inputs). */
__device__ __forceinline__ unsigned short to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float from_bf16(unsigned short h) {
  return __uint_as_float((unsigned)h << 16);
}

/* The two halves of a staged BF16x3 fragment: elements 0-3 hold the bf16
 * high parts, 4-7 the low parts (explicit shuffles: the ext-vector .lo/.hi
 * accessors on this vector type were compiled to the same half). */
__device__ __forceinline__ s16x4 hi_part(u16x8 v) {
  return __builtin_bit_cast(s16x4, __builtin_shufflevector(v, v, 0, 1, 2, 3));
}
__device__ __forceinline__ s16x4 lo_part(u16x8 v) {
  return __builtin_bit_cast(s16x4, __builtin_shufflevector(v, v, 4, 5, 6, 7));
}

__device__ __forceinline__ float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

/* LDS written by some lanes of a wavefront and then read by others of the
 * same wavefront: keep program order (the hardware executes one
 * wavefront's LDS instructions in order). */
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Workgroup barrier that orders LDS only: global stores still in flight
 * (the factors a solve writes out, never read back by the workgroup) are
 * not waited for.  The log-determinant kernels keep __syncthreads. */
template <bool LDS_ONLY> __device__ __forceinline__ void step_sync() {
  if constexpr (LDS_ONLY) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  } else {
    __syncthreads();
  }
}

/* LDS pointers (address space 3): the out-of-line diagonal factorization
 * then uses LDS instructions instead of flat ones. */
typedef __attribute__((address_space(3))) float lds_f;
typedef __attribute__((address_space(3))) f32x4 lds_f4;

__device__ __forceinline__ f32x4 lds_load4(const lds_f *p) {
  return *reinterpret_cast<const lds_f4 *>(p);
}
__device__ __forceinline__ void lds_store4(lds_f *p, f32x4 v) {
  *reinterpret_cast<lds_f4 *>(p) = v;
}

template <typename F, int... I>
__device__ __forceinline__ void static_for_seq(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F> __device__ __forceinline__ void static_for(F &&f) {
  static_for_seq(f, std::make_integer_sequence<int, N>{});
}

/* v from lane CTRL - 0x150 of each row of 16 lanes (row_newbcast), or the
 * row mirrored (CTRL 0x140) */
template <int CTRL> __device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

/* LU without interchanges of the 16 x 16 tile X (LDS, row stride TR) and
 * its inverse factors: one wavefront.  Writes L^-1 row-major to Linv, U^-1
 * transposed to UinvT (UinvT[c][i] = U^-1[i][c]) and |u_ii| to piv[i].  A
 * zero or NaN pivot, or a multiplier beyond kMaxMultiplier sets bad.
 * Every row of 16 lanes eliminates the tile (lane i holds row i; pivot
 * rows by readlane).  The inverses come from the same registers: rows of
 * the lower triangles L (unit) and T = J U J (row i = row 15 - i of U
 * reversed, by a row mirror) are formed right-looking, x_i -= (T_ic /
 * T_cc) x_c with row c broadcast within each row of lanes (DPP), then
 * scaled by 1 / T_ii; T^-1 = J U^-1 J.  Lane rows 0 and 2 take L, 1 and 3
 * take T, and rows 0 / 1 keep the even columns, 2 / 3 the odd ones. */
__device__ void factor_diag_tile(const lds_f *X, lds_f *Linv, lds_f *UinvT, lds_f *piv,
                                 int lane, int &bad) {
  const int i = lane & 15, g = lane >> 4;
  float v[16];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const f32x4 t = lds_load4(X + i * TR + 4 * q);
#pragma unroll
    for (int e = 0; e < 4; ++e)
      v[4 * q + e] = t[e];
  }
  /* pivot row j broadcast from lane j with readlane */
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    float p[16];
#pragma unroll
    for (int c = j; c < 16; ++c)
      p[c] = lane_value(v[c], j);
    const float dj = p[j];
    if (!(fabsf(dj) > 0.f))
      bad = 1;
    const float rd = __builtin_amdgcn_rcpf(dj);
    const bool below = i > j;
    const float l = below ? v[j] * rd : 0.f;
    if (!(fabsf(l) <= kMaxMultiplier))
      bad = 1;
    v[j] = below ? l : v[j];
#pragma unroll
    for (int c = j + 1; c < 16; ++c)
      v[c] = fmaf(-l, p[c], v[c]);
  }
  /* the lane's row index again, opaque to the compiler and available only
   * after the elimination, so that the row masks below are formed where
   * they are used instead of being shared with the elimination's (lane
   * masks live across the function are scalar registers its callers must
   * spill) */
  int io = i;
  asm volatile("" : "+v"(io) : "v"(v[15]));
  const bool tk = (g & 1) != 0;
  float a[16];
  static_for<16>([&](auto tc) {
    constexpr int t = decltype(tc)::value;
    const float m = dpp_f<0x140>(v[15 - t]);
    a[t] = tk ? m : v[t];
  });
  /* the diagonal entry, and the row's entries from the diagonal on set to
   * zero (no lane masks live through the loop below) */
  float dg = a[0];
#pragma unroll
  for (int t = 1; t < 16; ++t)
    dg = (io == t) ? a[t] : dg;
  if (lane < 16) {
    float u = v[0];
#pragma unroll
    for (int c = 1; c < 16; ++c)
      u = (io == c) ? v[c] : u;
    piv[i] = fabsf(u);
  }
#pragma unroll
  for (int t = 0; t < 16; ++t)
    a[t] = (t < io) ? a[t] : 0.f;
  const float rd = tk ? __builtin_amdgcn_rcpf(dg) : 1.f;
  const int h = g >> 1;
  float x[8]; /* columns 2 s + h of the lane's row */
#pragma unroll
  for (int s2 = 0; s2 < 8; ++s2)
    x[s2] = (2 * s2 + h == io) ? 1.f : 0.f;
  static_for<15>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    const float mm = a[c] * dpp_f<0x150 + c>(rd);
    /* row c is zero beyond column c */
#pragma unroll
    for (int s2 = 0; s2 <= c / 2; ++s2)
      x[s2] = fmaf(-mm, dpp_f<0x150 + c>(x[s2]), x[s2]);
  });
  if (!tk) {
#pragma unroll
    for (int s2 = 0; s2 < 8; ++s2)
      Linv[i * TR + 2 * s2 + h] = x[s2];
  } else {
    /* T^-1[i][t] = U^-1[15 - i][15 - t] = UinvT[15 - t][15 - i] */
#pragma unroll
    for (int s2 = 0; s2 < 8; ++s2)
      UinvT[(15 - 2 * s2 - h) * TR + 15 - i] = x[s2] * rd;
  }
}

typedef __attribute__((address_space(3))) int lds_i;

/* f(std::integral_constant<int, i>) for i = 0 .. N - 1, each body inlined
 * separately (constant register indices however long the loop). */
/* Largest (value, lowest index on ties) over the wavefront, by DPP: xor 1
 * and 2 within quads, the half-row and row mirrors, then rows 0 / 2 into
 * rows 1 / 3 and row 1 into rows 2 / 3; lane 63 ends with the result. */
__device__ __forceinline__ int wave_argmax(float bv, int bi) {
  auto step = [&](auto ctrl_c, auto rmask_c) {
    constexpr int CTRL = decltype(ctrl_c)::value, RMASK = decltype(rmask_c)::value;
    const float ov = __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(-3.f), __float_as_int(bv), CTRL, RMASK, 0xF, false));
    const int oi = __builtin_amdgcn_update_dpp(0x7fffffff, bi, CTRL, RMASK, 0xF, false);
    if (ov > bv || (ov == bv && oi < bi)) {
      bv = ov;
      bi = oi;
    }
  };
  step(std::integral_constant<int, 0xB1>{}, std::integral_constant<int, 0xF>{});  /* quad_perm 1,0,3,2 */
  step(std::integral_constant<int, 0x4E>{}, std::integral_constant<int, 0xF>{});  /* quad_perm 2,3,0,1 */
  step(std::integral_constant<int, 0x141>{}, std::integral_constant<int, 0xF>{}); /* row_half_mirror  */
  step(std::integral_constant<int, 0x140>{}, std::integral_constant<int, 0xF>{}); /* row_mirror       */
  step(std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xA>{}); /* row_bcast:15     */
  step(std::integral_constant<int, 0x143>{}, std::integral_constant<int, 0xC>{}); /* row_bcast:31     */
  return __builtin_amdgcn_readlane(bi, 63);
}

/* Partial-pivoting LU of the panel P (tile column k of S from row 16 k:
 * nrows <= 320 rows of 16 columns, row stride TR) by one wavefront, rows
 * r = lane + 64 s read and updated in LDS.  Rows are not exchanged while
 * factoring: the pivot of column j is the largest |a_rj| over the rows not
 * chosen yet (lowest row on ties, NaN below every number), its row read by
 * every lane into scalar registers, and l = a_rj (1 / u_jj) stored in
 * place of a_rj.  Then the rows go to pivot order: pivot j to row j, the
 * rows it displaced from rows below 16 to the rows the pivots left.  The
 * same exchanges, as (destination, source) pairs of panel rows after their
 * count in pairs[0], are applied to the trailing columns by every
 * wavefront.  Also writes |u_jj| to piv[j] and L_kk^-1 (unit lower,
 * row-major) to Linv; a zero or NaN pivot sets bad.  tmp: 32 x 16 floats
 * of scratch. */
__device__ __attribute__((noinline)) void factor_panel_piv(lds_f *P, int nrows, lds_f *tmp,
                                                          lds_i *pairs, lds_f *piv, lds_f *Linv,
                                                          int lane, int &bad) {
  constexpr int RS = 5;
  unsigned live = 0;
#pragma unroll
  for (int s = 0; s < RS; ++s)
    if (lane + 64 * s < nrows)
      live |= 1u << s;
  int pv[16];
  static_for<16>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    constexpr int jq = j / 4, je = j % 4;
    float bv = -2.f;
    int bi = 0x7fffffff;
#pragma unroll
    for (int s = 0; s < RS; ++s) {
      if (64 * s < nrows) {
        float x = fabsf(P[min(lane + 64 * s, nrows - 1) * TR + j]);
        if (x != x)
          x = -1.f;
        if (((live >> s) & 1u) && x > bv) {
          bv = x;
          bi = lane + 64 * s;
        }
      }
    }
    const int p = __builtin_amdgcn_readfirstlane(wave_argmax(bv, bi));
    pv[j] = p;
    /* the pivot row, uniform: scalar registers */
    float u[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const f32x4 t = lds_load4(P + p * TR + 4 * q);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        u[4 * q + e] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(t[e])));
    }
    float dj = u[j];
    if (!(fabsf(dj) > 0.f)) {
      bad = 1;
      dj = 1.f;
    }
    if (lane == (p & 63))
      live &= ~(1u << (p >> 6));
    const float rdj = 1.f / dj;
    /* every lane computes each slot in use (rows past the panel read a
     * clamped row); only the stores of rows still live are predicated */
#pragma unroll
    for (int s = 0; s < RS; ++s) {
      if (64 * s < nrows) {
        const bool lv = (live >> s) & 1u;
        lds_f *row = P + min(lane + 64 * s, nrows - 1) * TR;
        f32x4 t[4] = {};
#pragma unroll
        for (int q = jq; q < 4; ++q)
          t[q] = lds_load4(row + 4 * q);
        const float l = t[jq][je] * rdj;
        t[jq][je] = l;
#pragma unroll
        for (int c = j + 1; c < 16; ++c)
          t[c / 4][c % 4] = fmaf(-l, u[c], t[c / 4][c % 4]);
        if (lv) {
#pragma unroll
          for (int q = jq; q < 4; ++q)
            lds_store4(row + 4 * q, t[q]);
        }
      }
    }
    wave_lds_sync();
  });
  /* the exchanges: pivot j to row j; each displaced row below 16 to a row
   * a pivot left (both in increasing order) */
  if (lane == 0) {
    unsigned pm = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (pv[j] < 16)
        pm |= 1u << pv[j];
    int n = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (pv[j] != j) {
        pairs[1 + 2 * n] = j;
        pairs[2 + 2 * n] = pv[j];
        ++n;
      }
    int r = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (pv[j] >= 16) {
        while ((pm >> r) & 1u)
          ++r;
        pairs[1 + 2 * n] = pv[j];
        pairs[2 + 2 * n] = r;
        ++n;
        ++r;
      }
    pairs[0] = n;
  }
  wave_lds_sync();
  const int n = __builtin_amdgcn_readfirstlane(pairs[0]);
  for (int e = lane; e < 16 * n; e += 64)
    tmp[e] = P[pairs[2 + 2 * (e >> 4)] * TR + (e & 15)];
  wave_lds_sync();
  for (int e = lane; e < 16 * n; e += 64)
    P[pairs[1 + 2 * (e >> 4)] * TR + (e & 15)] = tmp[e];
  wave_lds_sync();
  if (lane < 16) {
    const int i = lane;
    piv[i] = fabsf(P[i * TR + i]);
    /* column i of L_kk^-1 by forward substitution */
    float x[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      float sum = 0.f;
#pragma unroll
      for (int q = 0; q <= r / 4; ++q) {
        const f32x4 t = lds_load4(P + r * TR + 4 * q);
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (4 * q + e < r)
            sum = fmaf(t[e], x[4 * q + e], sum);
      }
      x[r] = (r == i) ? 1.f : (r < i ? 0.f : -sum);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r)
      Linv[r * TR + i] = x[r];
  }
}

template <int WR, int WC, int RT, int CT, int MODE, int MTT, bool GRP, bool SOLVE = false,
          bool PIV = false>
__global__ __launch_bounds__(64 * WR * WC, 2) void k_lu_rs(
    RsArgs a0, const RsArgs *__restrict__ tab, GrpMap gm) {
  constexpr int NW = WR * WC, NT = 64 * NW;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  int bid, nbk;
  const RsArgs a = grp_args<GRP>(a0, tab, gm, bid, nbk);
  constexpr unsigned LO = 0xffffu;
  /* Schur gather: blocks in flight per wavefront, rounds loaded ahead
   * per block (most blocks of the dense blocks' ELL tables have 2 or 3
   * rounds; the rounds beyond a block's own count are skipped by a
   * wavefront-uniform test, so padding costs no LDS reads).  Measured at
   * m = 400, N = 5e4, FP32 LU: (UB, RR) = (1, 4) 25.8 ms, (1, 2..5)
   * 25.8-26.0, (2, 4) 26.7, (4, 3) 27.3, (4, 8) with the padding read
   * 28.1, (8, 3) 33.3 (spills) */
  constexpr int UB = 1;
  constexpr int RR = 4;
  const int tid = threadIdx.x, lane = tid & 63;
  /* wavefront-uniform (scalar) values: tile ownership tests compile to
   * scalar branches */
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int l16 = lane & 15, g = lane >> 4;
  const int wr = wave % WR, wc = wave / WR;
  /* MTT > 0: the dense block is padded to exactly MTT tiles (identity), so
   * the bound checks on tile indices fold at compile time */
  const int MT = MTT > 0 ? MTT : a.MT;

  float *sl = reinterpret_cast<float *>(smem + a.pre_base);
  int *tb = reinterpret_cast<int *>(smem + a.pre_base + a.slot_bytes);
  int *flagA = reinterpret_cast<int *>(smem + a.flag_base);
  int *flagB = flagA + 1;
  double *redd = reinterpret_cast<double *>(smem + a.red_base);
  float *Xb = reinterpret_cast<float *>(smem + a.x_base);
  float *Lb = reinterpret_cast<float *>(smem + a.l_base);
  float *Ub = reinterpret_cast<float *>(smem + a.u_base);
  float *LUb = reinterpret_cast<float *>(smem + a.lu_base);
  float *Linvb = reinterpret_cast<float *>(smem + a.linv_base);
  float *UinvTb = reinterpret_cast<float *>(smem + a.uinvt_base);
  float *pivb = reinterpret_cast<float *>(smem + a.piv_base); /* |u_kk| of the dense pivots */
  float *stage = reinterpret_cast<float *>(smem + a.stage_base);
  /* per-lane bases: a tile of wavefront-local index (ia, jb) sits at a
   * compile-time offset from them (the LDS instructions' immediate), so
   * the unrolled tile loops hold a handful of address registers */
  const int frag = l16 * TR + 4 * g;           /* MFMA operand fragment   */
  float *Lw = Lb + wr * TSF + frag;            /* tile I = wr + WR ia     */
  float *Uw = Ub + wc * TSF + frag;            /* tile J = wc + WC jb     */
  float *Xw = Xb + wr * TSF + 4 * g * TR + l16; /* accumulator layout     */
  constexpr int SST = 16 * WC * CT + 4;        /* staging row stride      */

  for (long d = bid; d < a.n_draws; d += nbk) {
    const double *th = a.theta + (size_t)d * a.P;
    int bad = 0;
    float amax = 0.f;
    double ls = 0.0;
    if (tid == 0) {
      *flagA = 0;
      *flagB = 0;
    }
    prefix_eliminate<float>(sl, tb, th, a.pre, tid, NT, !a.tab_persist || d == (long)bid, bad,
                            amax, ls, flagA, PIV ? kPrefixLooseMultiplier : kMaxMultiplier);
    bool stop = *flagA != 0;
    float *sol_d = nullptr;
    if constexpr (SOLVE) {
      /* the prefix factors for the substitutions */
      sol_d = a.sol + (size_t)d * a.sol_stride;
      if (!stop) {
        float *so = sol_d + (size_t)MT * MT * 256;
        for (int q = tid; q < a.pre.ns; q += NT)
          so[q] = sl[q];
      }
    }

    f32x4 acc[RT][CT];
#pragma unroll
    for (int ia = 0; ia < RT; ++ia)
#pragma unroll
      for (int jb = 0; jb < CT; ++jb)
        acc[ia][jb] = f32x4{0.f, 0.f, 0.f, 0.f};

    /* ---- Schur complement into the accumulators, one LDS band at a time */
    for (int band = 0; band < a.n_band && !stop; ++band) {
      const int r0 = band * a.band_rows;
      const int rows = min(a.band_rows, 16 * MT - r0);
      const int nvec = (rows * SST + 3) >> 2;
      f32x4 *s4 = reinterpret_cast<f32x4 *>(stage);
      for (int q = tid; q < nvec; q += NT)
        s4[q] = f32x4{0.f, 0.f, 0.f, 0.f};
      __syncthreads();
      for (int r = tid; r < rows; r += NT)
        stage[r * SST + r0 + r] = 1.f;
      for (int e = a.sg_ptr[band] + tid; e < a.sg_ptr[band + 1]; e += NT)
        stage[a.sg_off[e]] = (float)(-th[a.sg_idx[e]]);
      __syncthreads();
      const int b1 = a.blk_ptr[band + 1];
      for (int b0 = a.blk_ptr[band] + wave; b0 < b1; b0 += NW * UB) {
        int off[UB], R[UB], cb[UB];
        unsigned pv[UB][RR];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
          const int b = b0 + NW * u;
          const bool ok = b < b1;
          R[u] = ok ? a.brounds[b] : 0;
          cb[u] = ok ? a.bcbase[b] : 0;
          off[u] = ok ? a.st_off[(size_t)b * 64 + lane] : -1;
#pragma unroll
          for (int r = 0; r < RR; ++r)
            pv[u][r] = r < R[u] ? a.st_pairs[(size_t)cb[u] + 64 * r + lane] : a.zero_pair;
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
          float s = 0.f;
#pragma unroll
          for (int r = 0; r < RR; ++r)
            if (r < R[u])
              s = fmaf(-sl[pv[u][r] & LO], sl[pv[u][r] >> 16], s);
          for (int rb = RR; rb < R[u]; rb += RR) {
            unsigned c[RR];
#pragma unroll
            for (int r = 0; r < RR; ++r)
              c[r] = rb + r < R[u] ? a.st_pairs[(size_t)cb[u] + 64 * (rb + r) + lane] : a.zero_pair;
#pragma unroll
            for (int r = 0; r < RR; ++r)
              if (rb + r < R[u])
                s = fmaf(-sl[c[r] & LO], sl[c[r] >> 16], s);
          }
          if (off[u] >= 0)
            stage[off[u]] = s;
        }
      }
      __syncthreads();
#pragma unroll
      for (int ia = 0; ia < RT; ++ia) {
        const int I = wr + WR * ia;
        if (I < MT && 16 * I >= r0 && 16 * I < r0 + rows) {
          const float *srow = stage + (16 * I - r0 + 4 * g) * SST + 16 * wc + l16;
#pragma unroll
          for (int jb = 0; jb < CT; ++jb) {
            const int J = wc + WC * jb;
            if (J < MT) {
#pragma unroll
              for (int r = 0; r < 4; ++r)
                acc[ia][jb][r] = srow[r * SST + 16 * WC * jb];
            }
          }
        }
      }
      __syncthreads();
    }

    /* ---- phase 1 of step k: the column-k tiles to LDS (X), and the
     * diagonal tile's factors from its owner */
    auto phase1 = [&](int k) {
      if (wc == k % WC) {
        const int jk = k / WC;
#pragma unroll
        for (int jb = 0; jb < CT; ++jb) {
          if (jb == jk) {
#pragma unroll
            for (int ia = 0; ia < RT; ++ia) {
              const int I = wr + WR * ia;
              if (I >= k && I < MT) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                  Xw[WR * ia * TSF + r * TR] = acc[ia][jb][r];
              }
            }
          }
        }
        if (!PIV && wr == k % WR) {
          wave_lds_sync();
          factor_diag_tile((const lds_f *)(Xb + k * TSF), (lds_f *)Linvb, (lds_f *)UinvTb,
                           (lds_f *)(pivb + 16 * k), lane, bad);
          if (bad)
            *flagA = 1;
        }
      }
    };

    /* PIV: the panel of column k (staged by its owners, all of them past a
     * workgroup barrier) factored with partial pivoting by the owner of
     * tile (k, k) */
    lds_i *pairs = (lds_i *)LUb;
    auto panel = [&](int k) {
      if (wave == k % WR + WR * (k % WC)) {
        factor_panel_piv((lds_f *)(Xb + k * TSF), 16 * (MT - k), (lds_f *)UinvTb, pairs,
                         (lds_f *)(pivb + 16 * k), (lds_f *)Linvb, lane, bad);
        if (bad)
          *flagA = 1;
      }
    };

    /* PIV: step k's row exchanges on the trailing columns J > k, through
     * LDS: every source row out (WRITE), a workgroup barrier, every
     * destination row in.  Row 16 k + r of tile I sits in the lanes with
     * 4 g <= r % 16 < 4 g + 4, accumulator element r % 4. */
    float *swapb = reinterpret_cast<float *>(smem + a.swap_base);
    auto swap_rows = [&](int k, auto write_c) {
      constexpr bool WRITE = decltype(write_c)::value;
      const int np = __builtin_amdgcn_readfirstlane(pairs[0]);
      const int nt = MT - k - 1;
      for (int q = 0; q < np; ++q) {
        const int row =
            __builtin_amdgcn_readfirstlane(16 * k + pairs[WRITE ? 2 + 2 * q : 1 + 2 * q]);
        const int I = row >> 4;
        if (I % WR != wr)
          continue;
        const int ia = I / WR, lr = row & 15, rr = lr & 3;
        const bool mine = g == (lr >> 2);
        float *buf = swapb + q * nt * 16 + l16;
#pragma unroll
        for (int iac = 0; iac < RT; ++iac) {
          if (iac == ia) {
#pragma unroll
            for (int jb = 0; jb < CT; ++jb) {
              const int J = wc + WC * jb;
              if (J > k && J < MT && mine) {
                float *e = buf + (J - k - 1) * 16;
                f32x4 c = acc[iac][jb];
                if constexpr (WRITE) {
                  *e = rr == 0 ? c[0] : rr == 1 ? c[1] : rr == 2 ? c[2] : c[3];
                } else {
                  const float v = *e;
                  c[0] = rr == 0 ? v : c[0];
                  c[1] = rr == 1 ? v : c[1];
                  c[2] = rr == 2 ? v : c[2];
                  c[3] = rr == 3 ? v : c[3];
                  acc[iac][jb] = c;
                }
              }
            }
          }
        }
      }
    };

    /* A_IJ -= L_Ik U_kJ for the owned tiles of the columns J that sel(J)
     * picks, rows I > k */
    auto update = [&](int k, auto sel) {
#pragma unroll
      for (int jb = 0; jb < CT; ++jb) {
        const int J = wc + WC * jb;
        if (J < MT && sel(J)) {
          if constexpr (MODE == RS_FP32) {
            const f32x4 ub = *reinterpret_cast<const f32x4 *>(Uw + WC * jb * TSF);
#pragma unroll
            for (int ia = 0; ia < RT; ++ia) {
              const int I = wr + WR * ia;
              if (I > k && I < MT) {
                const f32x4 la = *reinterpret_cast<const f32x4 *>(Lw + WR * ia * TSF);
                f32x4 c = acc[ia][jb];
#pragma unroll
                for (int r = 0; r < 4; ++r)
                  c = mfma_f32(la[r], ub[r], c);
                acc[ia][jb] = c;
              }
            }
          } else {
            const u16x8 ub = *reinterpret_cast<const u16x8 *>(Uw + WC * jb * TSF);
            const s16x4 uh = hi_part(ub), ul = lo_part(ub);
#pragma unroll
            for (int ia = 0; ia < RT; ++ia) {
              const int I = wr + WR * ia;
              if (I > k && I < MT) {
                const u16x8 la = *reinterpret_cast<const u16x8 *>(Lw + WR * ia * TSF);
                const s16x4 lh = hi_part(la), ll = lo_part(la);
                f32x4 c = acc[ia][jb];
                c = mfma_bf16(ll, uh, c);
                c = mfma_bf16(lh, ul, c);
                c = mfma_bf16(lh, uh, c);
                acc[ia][jb] = c;
              }
            }
          }
        }
      }
    };

    if (!stop) {
      phase1(0);
      if constexpr (PIV) {
        __syncthreads();
        panel(0);
      }
    }
    for (int k = 0; k < MT && !stop; ++k) {
      step_sync<SOLVE>();
      if (*flagA) {
        stop = true;
        break;
      }
      if constexpr (PIV) {
        swap_rows(k, std::true_type{});
        __syncthreads();
        swap_rows(k, std::false_type{});
      }
      /* ---- phase 3: U_kJ by the owners of row k, L_Ik by the others */
      if (wr == k % WR) {
        const f32x4 li = *reinterpret_cast<const f32x4 *>(Linvb + frag);
        const int ik = k / WR;
        if constexpr (SOLVE) {
          if (wc == k % WC) {
            /* the diagonal tile's inverse factors, packed: L_kk^-1 below,
             * U_kk^-1 on and above the diagonal (element (4g + r, l16)) */
            float *t = sol_d + ((size_t)k * MT + k) * 256;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = 4 * g + r;
              t[row * 16 + l16] = row > l16 ? Linvb[row * TR + l16] : UinvTb[l16 * TR + row];
            }
          }
        }
#pragma unroll
        for (int ia = 0; ia < RT; ++ia) {
          if (ia == ik) {
#pragma unroll
            for (int jb = 0; jb < CT; ++jb) {
              const int J = wc + WC * jb;
              if (J > k && J < MT) {
                f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int r = 0; r < 4; ++r)
                  c = mfma_f32(li[r], acc[ia][jb][r], c);
                /* U_kJ transposed: row l16 holds U[4g .. 4g + 3][l16] */
                if constexpr (MODE == RS_FP32) {
                  *reinterpret_cast<f32x4 *>(Uw + WC * jb * TSF) = c;
                } else {
                  u16x8 h;
#pragma unroll
                  for (int r = 0; r < 4; ++r) {
                    const unsigned short hi = to_bf16(c[r]);
                    h[r] = hi;
                    h[4 + r] = to_bf16(c[r] - from_bf16(hi));
                  }
                  *reinterpret_cast<u16x8 *>(Uw + WC * jb * TSF) = h;
                }
              }
            }
          }
        }
      } else if constexpr (PIV) {
        /* -L_Ik from the factored panel */
        const int rank = ((wr - k % WR + WR) % WR - 1) * WC + wc;
        for (int t = rank; k + 1 + t < MT; t += (WR - 1) * WC) {
          const int I = k + 1 + t;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int o = I * TSF + (4 * g + r) * TR + l16;
            Lb[o] = -Xb[o];
          }
        }
      } else {
        const f32x4 ui = *reinterpret_cast<const f32x4 *>(UinvTb + frag);
        /* the (WR - 1) WC wavefronts outside row k take the column tiles
         * in turn */
        const int rank = ((wr - k % WR + WR) % WR - 1) * WC + wc;
        for (int t = rank; k + 1 + t < MT; t += (WR - 1) * WC) {
          const int I = k + 1 + t;
          const f32x4 xa = *reinterpret_cast<const f32x4 *>(Xb + I * TSF + frag);
          f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int r = 0; r < 4; ++r)
            c = mfma_f32(xa[r], ui[r], c);
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (!(fabsf(c[r]) <= kMaxMultiplier))
              bad = 1;
          /* -L_Ik row-major: row 4g + r, column l16 */
          if constexpr (MODE == RS_FP32) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              Lb[I * TSF + (4 * g + r) * TR + l16] = -c[r];
          } else {
            unsigned short *row = reinterpret_cast<unsigned short *>(Lb + I * TSF);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const float x = -c[r];
              const unsigned short hi = to_bf16(x);
              /* element (4g + r, l16): k-group l16 / 4, position l16 % 4 */
              const int o = (4 * g + r) * (2 * TR) + (l16 >> 2) * 8 + (l16 & 3);
              row[o] = hi;
              row[o + 4] = to_bf16(x - from_bf16(hi));
            }
          }
        }
      }
      if (bad)
        *flagB = 1;
      step_sync<SOLVE>();
      if (*flagB) {
        stop = true;
        break;
      }
      /* ---- phase 4: column k + 1 first, its diagonal tile factored while
       * the remaining columns are updated */
      if (k + 1 < MT) {
        update(k, [&](int J) { return J == k + 1; });
        phase1(k + 1);
        if constexpr (PIV) {
          __syncthreads();
          panel(k + 1);
        }
        update(k, [&](int J) { return J > k + 1; });
      }
      if constexpr (SOLVE) {
        /* step k's U_kJ (LDS, transposed) and L_Ik (LDS, negated) tiles to
         * the draw's factors, row-major, before the next step rewrites them;
         * by the wavefronts other than the one that factored tile k + 1 (the
         * longest phase 4) */
        const int nt = MT - k - 1;
        const int dw = (k + 1) % WR + WR * ((k + 1) % WC);
        const int pw = wave < dw ? wave : wave - 1;
        if (nt > 0 && wave != dw) {
          /* four consecutive columns of one row per lane: one 16-byte store */
          for (int e = pw * 64 + lane; e < 2 * nt * 64; e += NT - 64) {
            const int t = e >> 6, row = (e >> 2) & 15, col = 4 * (e & 3);
            f32x4 v;
            float *dst;
            if (t < nt) {
              const int J = k + 1 + t;
              const float *src = Ub + J * TSF + col * TR + row;
#pragma unroll
              for (int i = 0; i < 4; ++i)
                v[i] = src[i * TR];
              dst = sol_d + (k * MT + J) * 256 + row * 16 + col;
            } else {
              const int I = k + 1 + t - nt;
              v = -*reinterpret_cast<const f32x4 *>(Lb + I * TSF + row * TR + col);
              dst = sol_d + (I * MT + k) * 256 + row * 16 + col;
            }
            *reinterpret_cast<f32x4 *>(dst) = v;
          }
        }
      }
    }

    /* ---- log|det| (FP64, fixed reduction order) and the fallback list */
    step_sync<SOLVE>();
    if (!stop && !SOLVE)
      for (int q = tid; q < 16 * MT; q += NT)
        ls += log((double)pivb[q]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
      ls += __shfl_xor(ls, o);
    if (lane == 0)
      redd[wave] = ls;
    __syncthreads();
    if (tid == 0) {
      double logdet = redd[0];
      for (int w = 1; w < NW; ++w)
        logdet += redd[w];
      if (stop) {
        const int slot = atomicAdd(a.flag_count, 1);
        a.flag_list[slot] = (int)d;
      }
      if (a.logdet)
        a.logdet[d] = stop ? NAN : logdet;
      if constexpr (SOLVE)
        a.sol_ok[d] = stop ? 0 : 1;
    }
    __syncthreads();
  }
}

/* Forecast substitutions from the factors the register kernel wrote
 * (SOLVE): y = (I - Gamma)^{-1} rhs for every draw with sol_ok set, one
 * wavefront per draw, FP64 accumulation.  Prefix forward (wave by wave,
 * the saved slots), the dense part as block rows of 16 (forward: t_K =
 * b_K - sum_{J<K} L_KJ y_J, y_K = L_KK^-1 t_K; backward with U and
 * U_KK^-1), prefix backward; with refine one more pass on the FP64
 * residual rhs - (I - Gamma) y (sparse Gamma in series order), its
 * correction added.  Lane (r, q) = (lane % 16, lane / 16) takes row r of a
 * block and columns 4q .. 4q + 3 of each tile (one 16-byte load); the four
 * column groups are summed by shuffles in a fixed order. */
struct SubstArgs {
  long n_draws;
  int m, MT;
  long P;
  const double *theta, *rhs;
  double *y;
  const float *sol;
  long sol_stride;
  const int *sol_ok;
  int refine;
  int n1, n2, waves;
  const int *wave_q, *node_of, *fwd_ptr, *bwd_ptr, *piv;
  const unsigned *fwd, *bwd;
  const int *g_rowptr, *g_col, *g_idx;
  int wave_bytes; /* LDS per wavefront: x (n1 + 16 MT doubles), 16 doubles */
};

__device__ __forceinline__ void wave_order() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* sum over tiles J0 <= J < J1 of row (r) of tile J, columns 4q .. 4q + 3,
 * times x_J's entries 4q .. 4q + 3 (row: the lane's row of tile J0 = 0
 * of the block row; x: x_0 + 4q); eight (then four) tiles' loads issued
 * together so a wavefront keeps several 1 KB reads in flight; fixed
 * summation order */
__device__ __forceinline__ double tile_row_dot(const float *row, const double *x, int J0,
                                               int J1) {
  double p0 = 0.0, p1 = 0.0;
  int J = J0;
  for (; J + 8 <= J1; J += 8) {
    float4 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      v[u] = *reinterpret_cast<const float4 *>(row + (size_t)(J + u) * 256);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double *xj = x + 16 * (J + u);
      double &p = (u & 1) ? p1 : p0;
      p += (double)v[u].x * xj[0];
      p += (double)v[u].y * xj[1];
      p += (double)v[u].z * xj[2];
      p += (double)v[u].w * xj[3];
    }
  }
  for (; J + 4 <= J1; J += 4) {
    float4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      v[u] = *reinterpret_cast<const float4 *>(row + (size_t)(J + u) * 256);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double *xj = x + 16 * (J + u);
      double &p = (u & 1) ? p1 : p0;
      p += (double)v[u].x * xj[0];
      p += (double)v[u].y * xj[1];
      p += (double)v[u].z * xj[2];
      p += (double)v[u].w * xj[3];
    }
  }
  for (; J < J1; ++J) {
    const float4 v = *reinterpret_cast<const float4 *>(row + (size_t)J * 256);
    const double *xj = x + 16 * J;
    p0 += (double)v.x * xj[0];
    p0 += (double)v.y * xj[1];
    p0 += (double)v.z * xj[2];
    p0 += (double)v.w * xj[3];
  }
  return p0 + p1;
}

/* acc - sum over entries e0 <= e < e1 of lst (slot | x index << 16) of
 * sl[slot] * xs[x index], in the order of e; the list and slot loads of
 * eight entries are issued together (the lists are serial per lane) */
__device__ __forceinline__ double sub_entries(double acc, const unsigned *__restrict__ lst,
                                              int e0, int e1, const float *__restrict__ sl,
                                              const double *xs) {
  constexpr unsigned LO = 0xffffu;
  for (int e = e0; e < e1; e += 8) {
    unsigned c[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      c[u] = e + u < e1 ? lst[e + u] : 0u;
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      v[u] = e + u < e1 ? sl[c[u] & LO] : 0.f;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (e + u < e1)
        acc -= (double)v[u] * xs[c[u] >> 16];
  }
  return acc;
}

__global__ __launch_bounds__(256) void k_lu_subst(SubstArgs a) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = lane & 15, q = lane >> 4;
  double *xs = reinterpret_cast<double *>(smem + (size_t)wave * a.wave_bytes);
  const int n1 = a.n1, n2 = a.n2, MT = a.MT;
  double *xd = xs + n1;
  double *tb = xd + 16 * MT;
  const long nwv = (long)gridDim.x * (blockDim.x >> 6);
  for (long d = (long)blockIdx.x * (blockDim.x >> 6) + wave; d < a.n_draws; d += nwv) {
    if (!a.sol_ok[d])
      continue;
    const float *LU = a.sol + (size_t)d * a.sol_stride;
    const float *sl = LU + (size_t)MT * MT * 256;
    const double *th = a.theta + (size_t)d * a.P;
    const double *rhs = a.rhs + (size_t)d * a.m;
    double *y = a.y + (size_t)d * a.m;
    for (int pass = 0; pass < (a.refine ? 2 : 1); ++pass) {
      auto bval = [&](int node) -> double {
        if (pass == 0)
          return rhs[node];
        double acc = rhs[node] - y[node];
        const int e1 = a.g_rowptr[node + 1];
        for (int e = a.g_rowptr[node]; e < e1; e += 8) {
          int gi[8], gc[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            gi[u] = e + u < e1 ? a.g_idx[e + u] : 0;
            gc[u] = e + u < e1 ? a.g_col[e + u] : 0;
          }
          double tv[8], yv[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            tv[u] = e + u < e1 ? th[gi[u]] : 0.0;
            yv[u] = e + u < e1 ? y[gc[u]] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (e + u < e1)
              acc += tv[u] * yv[u];
        }
        return acc;
      };
      /* forward through the prefix, wave by wave */
      for (int w = 0; w < a.waves; ++w) {
        for (int p = a.wave_q[w] + lane; p < a.wave_q[w + 1]; p += 64) {
          xs[p] = sub_entries(bval(a.node_of[p]), a.fwd, a.fwd_ptr[p], a.fwd_ptr[p + 1], sl, xs);
        }
        wave_order();
      }
      /* right-hand side of the dense part (identity padding: 0) */
      for (int s2 = lane; s2 < 16 * MT; s2 += 64) {
        double acc = 0.0;
        if (s2 < n2)
          acc = sub_entries(bval(a.node_of[n1 + s2]), a.fwd, a.fwd_ptr[n1 + s2],
                            a.fwd_ptr[n1 + s2 + 1], sl, xs);
        xd[s2] = acc;
      }
      wave_order();
      /* dense forward: block rows K = 0 .. MT - 1 */
      for (int K = 0; K < MT; ++K) {
        const float *row = LU + (size_t)K * MT * 256 + r * 16 + 4 * q;
        /* the diagonal tile's load issued with the row's */
        const float4 v = *reinterpret_cast<const float4 *>(row + (size_t)K * 256);
        double part = tile_row_dot(row, xd + 4 * q, 0, K);
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        const double t = xd[16 * K + r] - part;
        if (q == 0)
          tb[r] = t;
        wave_order();
        /* y_K = L_KK^-1 t: unit diagonal, strictly lower part of the tile */
        double s = 0.0;
        if (4 * q < r)
          s += (double)v.x * tb[4 * q];
        if (4 * q + 1 < r)
          s += (double)v.y * tb[4 * q + 1];
        if (4 * q + 2 < r)
          s += (double)v.z * tb[4 * q + 2];
        if (4 * q + 3 < r)
          s += (double)v.w * tb[4 * q + 3];
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (q == 0)
          xd[16 * K + r] = t + s;
        wave_order();
      }
      /* dense backward: block rows K = MT - 1 .. 0 */
      for (int K = MT - 1; K >= 0; --K) {
        const float *row = LU + (size_t)K * MT * 256 + r * 16 + 4 * q;
        const float4 v = *reinterpret_cast<const float4 *>(row + (size_t)K * 256);
        double part = tile_row_dot(row, xd + 4 * q, K + 1, MT);
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        const double t = xd[16 * K + r] - part;
        if (q == 0)
          tb[r] = t;
        wave_order();
        /* x_K = U_KK^-1 t: the tile's part on and above the diagonal */
        double s = 0.0;
        if (4 * q >= r)
          s += (double)v.x * tb[4 * q];
        if (4 * q + 1 >= r)
          s += (double)v.y * tb[4 * q + 1];
        if (4 * q + 2 >= r)
          s += (double)v.z * tb[4 * q + 2];
        if (4 * q + 3 >= r)
          s += (double)v.w * tb[4 * q + 3];
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (q == 0)
          xd[16 * K + r] = s;
        wave_order();
      }
      /* backward through the prefix, waves in reverse order */
      for (int w = a.waves - 1; w >= 0; --w) {
        for (int p = a.wave_q[w] + lane; p < a.wave_q[w + 1]; p += 64) {
          const float pv = sl[a.piv[p]];
          const double acc = sub_entries(xs[p], a.bwd, a.bwd_ptr[p], a.bwd_ptr[p + 1], sl, xs);
          xs[p] = acc / (double)pv;
        }
        wave_order();
      }
      for (int i = lane; i < n1 + n2; i += 64) {
        const int node = a.node_of[i];
        y[node] = (pass == 0) ? xs[i] : y[node] + xs[i];
      }
      /* y of every lane before the residual reads it */
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      __builtin_amdgcn_wave_barrier();
    }
  }
}

} // namespace

size_t rs_panel_bytes(int MT) { return (size_t)(3 * MT + 4) * TSF * sizeof(float) + 64 * MT; }

/* The instantiation that serves MT tiles per dimension: wavefront grid
 * WR x WC (eight wavefronts, two per SIMD, 256 registers each),
 * accumulator tiles RT x CT per wavefront, and the padded tile count when
 * the kernel is compiled for a fixed one (0: MT at run time). */
struct RsShape {
  int WR, WC, RT, CT, MTT;
};
/* nw = 4: four wavefronts (2 x 2 grid, 6 x 6 tiles each) for blocks of up
 * to 12 tiles, two workgroups per CU */
static RsShape rs_shape(int MT, int nw = 8) {
  if (nw == 4 && MT <= 12)
    return {2, 2, 6, 6, 0};
  if (MT <= 12)
    return {2, 4, 6, 3, 0};
  if (MT <= 16)
    return {2, 4, 8, 4, 0};
  return {2, 4, 9, 5, 18};
}

int rs_padded_tiles(int MT, int nw) {
  const RsShape q = rs_shape(MT, nw);
  return q.MTT > 0 ? q.MTT : MT;
}

int rs_stage_stride(int MT, int nw) {
  const RsShape q = rs_shape(MT, nw);
  return 16 * q.WC * q.CT + 4;
}

namespace {

RsArgs rs_args(const Structure &S, const LuPlan &plan, long N, const double *theta,
               double *logdet, int *flags) {
  RsArgs a{};
  a.MT = plan.rs_MT;
  a.n_draws = N;
  a.theta = theta;
  a.P = S.P;
  a.logdet = logdet;
  a.flag_count = flags;
  a.flag_list = flags + 1;
  a.pre = PrefixTabs{S.e_ns, S.e_waves, S.e_n1, S.e_tab_words, S.e_o_piv, S.e_o_mptr, S.e_o_mult,
                     S.e_o_tptr, S.e_o_tslot, S.e_o_cptr, S.e_o_contrib, S.e_slot_init, S.e_tab};
  a.pre_base = (int)S.e_rb_pre_base;
  a.slot_bytes = (int)((((size_t)S.e_ns * 4) + 15) & ~(size_t)15);
  a.tab_persist = S.e_rb_tab_persist;
  a.flag_base = (int)plan.rs_flag_base;
  a.red_base = a.flag_base + 16;
  const int P0 = (int)plan.rs_panel_base;
  a.x_base = P0;
  a.l_base = a.x_base + a.MT * TSF * 4;
  a.u_base = a.l_base + a.MT * TSF * 4;
  a.lu_base = a.u_base + a.MT * TSF * 4;
  a.linv_base = a.lu_base + TSF * 4;
  a.uinvt_base = a.linv_base + TSF * 4;
  a.urev_base = a.uinvt_base + TSF * 4;
  a.piv_base = a.urev_base + TSF * 4;
  a.stage_base = (int)S.e_rb_stage_base;
  a.band_rows = S.e_rb_rows;
  a.n_band = S.e_rb_nband;
  a.blk_ptr = S.e_rb_blk_ptr;
  a.st_off = S.e_rb_st_off;
  a.brounds = S.e_rb_brounds;
  a.bcbase = S.e_rb_bcbase;
  a.sg_ptr = S.e_rb_sg_ptr;
  a.sg_off = S.e_rb_sg_off;
  a.sg_idx = S.e_rb_sg_idx;
  a.st_pairs = S.e_rb_pairs;
  a.zero_pair = S.e_zero_pair;
  a.swap_base = (int)plan.rs_swap_base;
  return a;
}

/* The register kernel for a dense block of MT tiles: single-model and
 * group forms, workgroup size (one workgroup per CU). */
struct RsKern {
  const void *one = nullptr, *grp = nullptr;
  int threads = 0;
};

template <int WR, int WC, int RT, int CT, int MODE, int MTT, bool PIV = false> RsKern rs_k() {
  RsKern k;
  k.one = (const void *)k_lu_rs<WR, WC, RT, CT, MODE, MTT, false, false, PIV>;
  k.grp = (const void *)k_lu_rs<WR, WC, RT, CT, MODE, MTT, true, false, PIV>;
  k.threads = 64 * WR * WC;
  return k;
}

/* The register kernel in solve mode (FP32 factors written out). */
template <int WR, int WC, int RT, int CT, int MTT> const void *rs_solve_k() {
  return (const void *)k_lu_rs<WR, WC, RT, CT, RS_FP32, MTT, false, true>;
}
const void *rs_solve_kern(int MT, int nw) {
  const RsShape q = rs_shape(MT, nw);
  if (MT > kRsMaxTiles || MT < 1)
    return nullptr;
  if (q.WC == 2)
    return rs_solve_k<2, 2, 6, 6, 0>();
  if (q.RT == 6)
    return rs_solve_k<2, 4, 6, 3, 0>();
  if (q.RT == 8)
    return rs_solve_k<2, 4, 8, 4, 0>();
  return rs_solve_k<2, 4, 9, 5, 18>();
}

/* piv: partial pivoting inside the kernel (FP32 only) */
RsKern rs_kern(int MT, bool bf, bool piv, int nw) {
  const RsShape q = rs_shape(MT, nw);
  if (MT > kRsMaxTiles || MT < 1 || (piv && bf))
    return RsKern();
  if (q.WC == 2) {
    if (piv)
      return rs_k<2, 2, 6, 6, RS_FP32, 0, true>();
    return bf ? rs_k<2, 2, 6, 6, RS_BF16X3, 0>() : rs_k<2, 2, 6, 6, RS_FP32, 0>();
  }
  if (piv) {
    if (q.RT == 6)
      return rs_k<2, 4, 6, 3, RS_FP32, 0, true>();
    if (q.RT == 8)
      return rs_k<2, 4, 8, 4, RS_FP32, 0, true>();
    return rs_k<2, 4, 9, 5, RS_FP32, 18, true>();
  }
  if (q.RT == 6)
    return bf ? rs_k<2, 4, 6, 3, RS_BF16X3, 0>() : rs_k<2, 4, 6, 3, RS_FP32, 0>();
  if (q.RT == 8)
    return bf ? rs_k<2, 4, 8, 4, RS_BF16X3, 0>() : rs_k<2, 4, 8, 4, RS_FP32, 0>();
  return bf ? rs_k<2, 4, 9, 5, RS_BF16X3, 18>() : rs_k<2, 4, 9, 5, RS_FP32, 18>();
}

} // namespace

int launch_lu_rs(const Structure &S, const LuPlan &plan, long N, const double *theta,
                 double *logdet, int *flags, hipStream_t st) {
  RsArgs a = rs_args(S, plan, N, theta, logdet, flags);
  const RsKern k = rs_kern(a.MT, plan.prec == SGDLM_PREC_BF16X3, plan.rs_piv != 0, plan.rs_nw);
  if (!k.one)
    return SGDLM_ERR_ARG;
  if (allow_dynamic_lds(k.one, plan.rs_lds) != hipSuccess)
    return SGDLM_ERR_DEVICE;
  const RsArgs *tab = nullptr;
  GrpMap gm;
  void *args[] = {&a, &tab, &gm};
  if (hipLaunchKernel(k.one, dim3(plan.rs_grid), dim3(k.threads), args, plan.rs_lds, st) !=
      hipSuccess)
    return SGDLM_ERR_DEVICE;
  return hipGetLastError() == hipSuccess ? SGDLM_OK : SGDLM_ERR_DEVICE;
}

size_t rs_solve_stride(const Structure &S, int MT) {
  const size_t MTp = (size_t)rs_padded_tiles(MT);
  return (MTp * MTp * 256 + (size_t)S.e_ns + 63) & ~(size_t)63;
}

int lu_subst_shape(const Structure &S, int MT, int n_cu, int *grid, size_t *lds) {
  const int MTp = rs_padded_tiles(MT);
  const size_t wave = (((size_t)S.e_n1 + 16 * (size_t)MTp + 16) * sizeof(double) + 15) &
                      ~(size_t)15;
  *lds = 4 * wave;
  int per_cu = 0;
  if (allow_dynamic_lds((const void *)k_lu_subst, *lds) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_lu_subst, 256,
                                                   *lds) != hipSuccess ||
      per_cu < 1) {
    (void)hipGetLastError();
    return SGDLM_ERR_ARG;
  }
  *grid = std::max(1, n_cu) * per_cu;
  return SGDLM_OK;
}

int launch_lu_rs_factor(const Structure &S, const LuPlan &plan, long n, const double *theta,
                        float *sol, int *sol_ok, int *flags, hipStream_t st) {
  RsArgs a = rs_args(S, plan, n, theta, nullptr, flags);
  a.sol = sol;
  a.sol_stride = (long)plan.rs_sol_stride;
  a.sol_ok = sol_ok;
  const void *k = rs_solve_kern(a.MT, plan.rs_nw);
  if (!k)
    return SGDLM_ERR_ARG;
  if (allow_dynamic_lds(k, plan.rs_lds) != hipSuccess)
    return SGDLM_ERR_DEVICE;
  const RsShape q = rs_shape(a.MT, plan.rs_nw);
  const RsArgs *tab = nullptr;
  GrpMap gm;
  void *args[] = {&a, &tab, &gm};
  const int grid = (int)std::max(1L, std::min<long>(plan.rs_grid, n));
  if (hipLaunchKernel(k, dim3(grid), dim3(64 * q.WR * q.WC), args, plan.rs_lds, st) != hipSuccess)
    return SGDLM_ERR_DEVICE;
  return hipGetLastError() == hipSuccess ? SGDLM_OK : SGDLM_ERR_DEVICE;
}

int launch_lu_subst(const Structure &S, const LuPlan &plan, long n, const double *theta,
                    const double *rhs, double *y, const float *sol, const int *sol_ok,
                    int refine, hipStream_t st) {
  SubstArgs a{};
  a.n_draws = n;
  a.m = S.m;
  a.MT = rs_padded_tiles(plan.rs_MT, plan.rs_nw);
  a.P = S.P;
  a.theta = theta;
  a.rhs = rhs;
  a.y = y;
  a.sol = sol;
  a.sol_stride = (long)plan.rs_sol_stride;
  a.sol_ok = sol_ok;
  a.refine = refine;
  a.n1 = S.e_n1;
  a.n2 = S.e_n2;
  a.waves = S.e_waves;
  a.wave_q = S.e_wave_q;
  a.node_of = S.e_node_of;
  a.fwd_ptr = S.e_fwd_ptr;
  a.bwd_ptr = S.e_bwd_ptr;
  a.piv = S.e_tab + S.e_o_piv;
  a.fwd = S.e_fwd;
  a.bwd = S.e_bwd;
  a.g_rowptr = S.g_rowptr;
  a.g_col = S.g_col;
  a.g_idx = S.g_idx;
  a.wave_bytes = (int)(plan.subst_lds / 4);
  const int grid = (int)std::max(1L, std::min<long>(plan.subst_grid, (n + 3) / 4));
  hipLaunchKernelGGL(k_lu_subst, dim3(grid), dim3(256), plan.subst_lds, st, a);
  return hipGetLastError() == hipSuccess ? SGDLM_OK : SGDLM_ERR_DEVICE;
}

int lu_grp_rs(const Structure &S, const LuPlan &plan, long N, const double *theta,
              double *logdet, int *flags, LuGrpKernel *k) {
  const RsArgs a = rs_args(S, plan, N, theta, logdet, flags);
  const RsKern rk = rs_kern(a.MT, plan.prec == SGDLM_PREC_BF16X3, plan.rs_piv != 0, plan.rs_nw);
  if (!rk.grp)
    return SGDLM_ERR_ARG;
  k->fn = rk.grp;
  k->threads = rk.threads;
  k->per_cu = plan.rs_nw == 4 ? 2 : 1;
  k->lds = plan.rs_lds;
  k->ws_wg = 0;
  k->units = N;
  /* ms per draw on one workgroup (size model; 0.15 ms at 18 tiles) */
  const double r = a.MT / 18.0;
  k->cost = (double)N * ((plan.rs_piv ? 0.19 : 0.15) * r * r + 5e-7 * (double)S.e_fma);
  k->args.assign(reinterpret_cast<const unsigned char *>(&a),
                 reinterpret_cast<const unsigned char *>(&a) + sizeof(a));
  k->work_field = -1;
  return SGDLM_OK;
}

} // namespace sgdlm
