// Exact-accumulation grid Fock contraction on int8 MFMA (Ozaki-style digit slicing).
//
// F = Phi^T diag(w bw v) Phi = At^T B with At = sqrt|w bw| Phi (per geometry) and
// B = sign(w bw) v At (per build). Both operands are written, per (function, chunk of
// points), as a 35-bit fixed-point integer J in [-2^34, 2^34] relative to a power-of-two
// scale, split into five balanced base-128 digits d1..d5 in [-64, 64]:
//   J = d1 2^28 + d2 2^21 + d3 2^14 + d4 2^7 + d5.
// The products of digit pairs (i, j) with i + j <= 6 (1-based, 15 of the 25) run on
// v_mfma_i32_16x16x64_i8, whose int32 accumulation is exact; each level L = i + j has its own
// accumulator S_L, chained through the MFMAs over the chunk's stages (|S_L| < 2^27 for up to
// 63 stages), so a chunk's block is
//   2^(e_row + e_col) ((((S2 2^-12 + S3 2^-19) + S4 2^-26) + S5 2^-33) + S6 2^-40).
// Chunks are summed in FP64 in a fixed order. The only roundings are the 35-bit representation of
// both operands and the dropped products of level >= 7 (below 2^-35 of the leading one); the
// energy of the returned F stays within 1e-6 Eh of FP64 on molecules whose near-nucleus shells
// make 28-bit digits miss it by up to 7e-5 Eh.
//
// Screening (optional, threshold 2^-S): Gaussian basis functions are local, so in most chunks of
// points most 16 x 16 blocks of F get nothing. With every |At| of a block's functions in a chunk
// below 2^eI (the largest of their chunk exponents), |vs| below 2^eV and at most 2^lg points in
// the chunk, every element of block (I, J)'s chunk sum is below 2^(eI + eJ + eV + lg); blocks
// with eI + eJ + eV + lg <= -S are skipped in that chunk, so each element of F moves by at most
// nchunk 2^-S. The functions are then ordered by how many chunks they reach (the most compact
// first), so the blocks of compact functions are empty far from their centres.
#pragma once

#include <vector>

#include "common.hpp"

namespace napmo {

constexpr int OZ_STAGE = 64;      // points per stage (the MFMA k depth)
constexpr int OZ_DIGITS = 5;      // base-128 digits per operand
constexpr int OZ_MAXB = 5;        // 16 x 16 blocks per wave (five int32 accumulators each)
constexpr int OZ_MAXCOL = 3;      // column blocks per wave (their A digits held in registers)
constexpr int OZ_TASK = 3 * OZ_MAXB + OZ_MAXCOL;  // ints per wave task
constexpr int OZ_WAVES = 8;       // waves per workgroup
constexpr int OZ_MAX_STAGES = 63; // stages per chunk (the level sums stay far inside int32)
constexpr int OZ_MAXROWS = 24;    // block rows of one task (n <= 384)
constexpr int OZ_PTASK = 2 + OZ_MAXROWS + OZ_WAVES * OZ_TASK;  // ints per packed task

// Per-plan layout: chunks of points, 16-function blocks, groups of block rows.
struct OzakiLayout {
  long N = 0, Nq = 0;     // points, padded to nchunk * chunk
  long chunk = 0;          // points per chunk (multiple of OZ_STAGE, <= OZ_MAX_STAGES stages)
  int nchunk = 0;
  int n = 0, ld = 0;       // functions; leading dimension of the digit planes (>= 16 nb)
  int nb = 0;              // 16-function blocks covering n
  int nblk = 0;            // upper blocks nb (nb + 1) / 2
  int ngroups = 0;
  std::vector<int> row0;   // group g: block rows [row0[g], row0[g + 1]), columns [row0[g], nb)
  // [ngroups][OZ_WAVES][OZ_TASK]: OZ_MAXB slots (block row I, column J, index of J in the
  // wave's column list; I = -1: empty), then the wave's column list (-1: none)
  std::vector<int> tasks;
  size_t lds_bytes = 0;    // largest dynamic LDS of a group
  int pack_rows = 0;       // packed tasks: at most this many block rows (their LDS fits)
  size_t pack_lds = 0;     // dynamic LDS of a packed task
  long partial_stride() const { return (long)nblk * 256 + 32; }
};

// Chooses the chunking, the row groups (as many block rows per group as the wave slots take,
// at most wave_blocks blocks and OZ_MAXCOL columns per wave) and the wave tasks. Throws when
// no grouping fits the LDS or the wave slots.
OzakiLayout ozaki_layout(long N, int n, int ld, int wave_blocks, size_t max_lds, long chunk_target = 0);

constexpr int OZ_ZERO_EXP = -2000;  // chunk exponent of a function that is zero in the chunk
constexpr int OZ_NO_CACHE = -1000000;  // packing cache: no tasks formed for the chunk yet

// Per-geometry operand: digits of At = sqrt|wbw| Phi. phi64 [Npad x ld] point-major (Npad >= N),
// wbw [N]; perm [ld] (NULL: identity): operand function f is column perm[f] of phi64; out:
// planes [OZ_DIGITS][ld][Nq] int8 function-major, eA [ld][nchunk] int (OZ_ZERO_EXP where the
// function is zero in the chunk), maxd1 [ld][Nq / OZ_STAGE] uint8 (largest |d1| per stage).
void launch_ozaki_scales(const OzakiLayout &L, const double *phi64, const double *wbw, const int *perm,
                         int *eA, hipStream_t s);
void launch_ozaki_setup(const OzakiLayout &L, const double *phi64, const double *wbw, const int *perm,
                        signed char *planes, int *eA, unsigned char *maxd1, hipStream_t s);
// The screening order: functions by the number of chunks where their chunk exponent reaches
// the largest one less 30 (fewest first, ties by index), from eA [ld][nchunk] of the
// identity order; functions n..ld-1 stay last.
std::vector<int> ozaki_compact_order(const OzakiLayout &L, const std::vector<int> &eA);
// Upper 16 x 16 blocks x chunks that launch_ozaki_fock computes at threshold 2^-screen for the
// potential vs [Nq] (all of them when screen == 0), from eA [ld][nchunk] in operand order.
long ozaki_kept_blocks(const OzakiLayout &L, const std::vector<int> &eA, const std::vector<double> &vs,
                       int screen);
// vs [Nq] = sign(wbw) v for p < N, 0 beyond.
void launch_ozaki_vs(const double *v, const double *wbw, long N, long Nq, double *vs, hipStream_t s);
// Per build: partial [nchunk][partial_stride] FP64 (the 16 x 16 values of every upper block,
// row-major, blocks in upper-triangle row order); d_tasks = L.tasks followed by L.row0;
// screen = S > 0 skips the blocks bounded by 2^-S (0: none).
void launch_ozaki_fock(const OzakiLayout &L, const int *d_tasks, const signed char *planes,
                       const int *eA, const unsigned char *maxd1, const double *vs, int screen,
                       double *partial, hipStream_t s);
// The same contraction with the blocks each chunk keeps packed into as few tasks as the LDS and
// wave slots allow (screen > 0): a packing pass per chunk forms the kept set (the rule above)
// and writes zeros for the skipped blocks' partials, then a grid of one workgroup per compute
// unit takes the tasks in turn. ptasks [nchunk * nb][OZ_PTASK] and ctr [2] are scratch; cache
// [nchunk * nb][OZ_PTASK], ccount [nchunk] and cev [nchunk] (OZ_NO_CACHE at first) keep each
// chunk's tasks across builds, reused while the chunk's |v| exponent has not grown. Equal to
// launch_ozaki_fock at the same threshold, bit for bit.
void launch_ozaki_fock_packed(const OzakiLayout &L, const signed char *planes, const int *eA,
                              const unsigned char *maxd1, const double *vs, int screen, double *partial,
                              int *ptasks, int *ctr, int *cache, int *ccount, int *cev, hipStream_t s);
// F [n x n] (symmetric) = the fixed-order sum over chunks of the partials; element (mu, nu) of
// the operand order is F[perm[mu]][perm[nu]] (perm NULL: identity).
void launch_ozaki_reduce(const OzakiLayout &L, const double *partial, const int *perm, double *F,
                         hipStream_t s);

}  // namespace napmo
