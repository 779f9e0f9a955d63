/*
 * lu_symbolic.h -- symbolic elimination plan for the log-determinant of
 * I - Gamma (host C++, no device code).
 *
 * The dense LU of the reference (cublasXgetrfBatched on m x m matrices,
 * rSGDLM/src/SGDLM/SGDLM.cu:280-281) ignores that I - Gamma has only
 * sum_j k_j off-diagonal entries.  All draws share that pattern, so it is
 * analysed once per parental-set structure:
 *
 *   1. Multiple minimum degree on the symmetrised pattern: independent sets
 *      of low-degree series ("waves") are eliminated together while the
 *      minimum degree stays <= max_front.  The eliminated series form the
 *      sparse prefix (n1 series); the rest is the separator (n2 series).
 *   2. Every entry touched by the prefix elimination gets an LDS slot; per
 *      wave the kernel computes the multipliers l_ik = a_ik / a_kk and then
 *      each target slot gathers its updates sum_k l_ik * a_kj in a fixed
 *      order (deterministic, no atomics).  These tables are small and are
 *      copied into LDS.
 *   3. The Schur complement S = A22 - L21 U12 on the separator is assembled
 *      densely (n2 rounded up to 16, identity padding): every entry of S
 *      that the prefix updates is a sum over (x, y) slot pairs,
 *      S_ij = -sum x * y, whose first pair (-1, A22_ij) carries the entry of
 *      I - Gamma.  The pairs are stored in blocks of 64 entries (one per
 *      lane), round by round (ELL layout), so a wavefront reads them with
 *      coalesced loads; blocks are sorted by their number of rounds.
 *      S is then factored by the dense MFMA LU.
 *
 * log|det(I - Gamma)| = sum over prefix pivots + sum over the dense pivots.
 * The symmetric permutation keeps the unit diagonal on the diagonal; the
 * multiplier bound of the fast path (|l| <= 2) applies to both parts.
 */
#ifndef SGDLM_LU_SYMBOLIC_H
#define SGDLM_LU_SYMBOLIC_H

#include <stdint.h>

#include <vector>

namespace sgdlm {

struct GammaEntry {
  int row, col, idx; /* (I - Gamma)[row][col] = -theta_packed[idx] */
};

/* slot_init codes besides a theta index >= 0 */
enum { SLOT_ZERO = -1, SLOT_ONE = -2, SLOT_MINUS_ONE = -3 };

struct ElimPlan {
  int m = 0;
  int n1 = 0;      /* series eliminated in the sparse prefix              */
  int n2 = 0;      /* separator series (dense Schur complement)           */
  int M2 = 0;      /* n2 rounded up to 16                                 */
  int ns = 0;      /* LDS slots (prefix entries + constant/base slots)    */
  int n_waves = 0;
  int n_blk = 0;   /* blocks of 64 Schur-complement targets               */
  std::vector<int> order;   /* n1 prefix series in elimination order       */
  std::vector<int> sep_pos; /* m: dense index of a separator series, -1    */

  std::vector<int> slot_init; /* ns: theta index or SLOT_* code          */
  uint32_t zero_pair = 0;     /* pair of two zero slots (ELL padding)     */
  /* LDS table image (32-bit words) and the offsets of its parts:
   *   piv     [n1]          slot of each prefix pivot a_kk
   *   mptr    [n_waves + 1] wave ranges of mult
   *   mult    [..]          slot(a_ik) | slot(a_kk) << 16
   *   tptr    [n_waves + 1] wave ranges of the prefix targets
   *   tslot   [..]          prefix target slots
   *   cptr    [targets + 1] ranges of contrib per prefix target
   *   contrib [..]          slot(l_ik) | slot(a_kj) << 16
   *   brounds [n_blk]       rounds of each Schur block
   *   bcbase  [n_blk]       first pair of each Schur block in st_pairs   */
  std::vector<int> tab;
  int o_piv = 0, o_mptr = 0, o_mult = 0, o_tptr = 0, o_tslot = 0, o_cptr = 0,
      o_contrib = 0, o_brounds = 0, o_bcbase = 0;

  std::vector<int> sg_off;  /* Gamma entries inside S: r * M2 + c         */
  std::vector<int> sg_idx;  /*   and their theta index                    */
  std::vector<int> st_off;  /* n_blk * 64 target offsets, -1 = padding    */
  std::vector<uint32_t> st_pairs; /* pair (x | y << 16) of block b, round
                                     r, lane l at bcbase[b] + 64 r + l    */
  long long prefix_fma = 0; /* multiply-adds of the prefix + S gather     */

  /* Triangular solves through the prefix (forecasts), with x indexed
   * [0, n1) = prefix series in elimination order, [n1, n1 + n2) =
   * separator series in dense order:
   *   forward:  z_q = b - sum l * z over fwd entries (waves in order), then
   *             the separator rhs b_s - sum l * z (rows n1.. of fwd_ptr);
   *   backward: x_q = (z_q - sum u * x) / u_qq (waves in reverse order).
   * Entries are slot | index << 16. */
  std::vector<int> wave_q;    /* n_waves + 1: prefix index range per wave  */
  std::vector<int> node_of;   /* n1 + n2: series of each x index           */
  std::vector<int> fwd_ptr;   /* n1 + n2 + 1                               */
  std::vector<uint32_t> fwd;  /* slot(l) | q << 16                         */
  std::vector<int> bwd_ptr;   /* n1 + 1                                    */
  std::vector<uint32_t> bwd;  /* slot(u) | x index << 16                   */
};

/* The Schur-complement tables of a plan regrouped for the
 * register-resident LU, which stages S through LDS one band of BR rows at a
 * time (row stride `stride` floats): the targets and the Gamma entries of
 * each band, with offsets local to the band.  Blocks of 64 targets (ELL,
 * round r of block b at bcbase[b] + 64 r) are sorted by rounds within a
 * band; a band's blocks are blk_ptr[band] .. blk_ptr[band + 1]. */
struct BandPlan {
  int BR = 0, stride = 0, n_band = 0, n_blk = 0;
  std::vector<int> blk_ptr;       /* n_band + 1                             */
  std::vector<int> st_off;        /* n_blk * 64 band-local offsets, -1 pad  */
  std::vector<uint32_t> st_pairs; /* pairs (x | y << 16), zero_pair padding */
  std::vector<int> brounds, bcbase; /* n_blk                                */
  std::vector<int> sg_ptr;        /* n_band + 1                             */
  std::vector<int> sg_off, sg_idx; /* Gamma entries of S, band-local        */
};

/* Regroups P's Schur tables into bands of BR rows (BR a multiple of 16)
 * with LDS row stride `stride` (>= rows); `rows` >= P.M2 is the dense
 * dimension the kernel factors (identity padding beyond P.n2). */
void build_band_plan(const ElimPlan &P, int BR, int stride, int rows, BandPlan &B);

/* Builds the plan.  max_front = largest degree a prefix pivot may have
 * (0 disables the prefix: n1 = 0, S = the whole matrix); max_slots bounds
 * the LDS slots (and must be <= 65536). */
void build_elim_plan(int m, const std::vector<GammaEntry> &ents, int max_front,
                     int max_slots, ElimPlan &plan);

} // namespace sgdlm

#endif
