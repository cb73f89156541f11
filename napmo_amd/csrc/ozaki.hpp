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
#pragma once

#include "common.hpp"

namespace napmo {

constexpr int OZ_STAGE = 64;      // points per stage (the MFMA k depth)
constexpr int OZ_DIGITS = 5;      // base-128 digits per operand
constexpr int OZ_MAXB = 5;        // 16 x 16 blocks per wave (five int32 accumulators each)
constexpr int OZ_MAXCOL = 3;      // column blocks per wave (their A digits held in registers)
constexpr int OZ_TASK = 3 * OZ_MAXB + OZ_MAXCOL;  // ints per wave task
constexpr int OZ_WAVES = 8;       // waves per workgroup
constexpr int OZ_MAX_STAGES = 63; // stages per chunk (the level sums stay far inside int32)

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
  long partial_stride() const { return (long)nblk * 256 + 32; }
};

// Chooses the chunking, the row groups (as many block rows per group as the wave slots take,
// at most wave_blocks blocks and OZ_MAXCOL columns per wave) and the wave tasks. Throws when
// no grouping fits the LDS or the wave slots.
OzakiLayout ozaki_layout(long N, int n, int ld, int wave_blocks, size_t max_lds, long chunk_target = 0);

// Per-geometry operand: digits of At = sqrt|wbw| Phi. phi64 [Npad x ld] point-major (Npad >= N),
// wbw [N]; out: planes [OZ_DIGITS][ld][Nq] int8 function-major, eA [ld][nchunk] int, maxd1
// [ld][Nq / OZ_STAGE] uint8 (largest |d1| per stage).
void launch_ozaki_setup(const OzakiLayout &L, const double *phi64, const double *wbw,
                        signed char *planes, int *eA, unsigned char *maxd1, hipStream_t s);
// vs [Nq] = sign(wbw) v for p < N, 0 beyond.
void launch_ozaki_vs(const double *v, const double *wbw, long N, long Nq, double *vs, hipStream_t s);
// Per build: partial [nchunk][partial_stride] FP64 (the 16 x 16 values of every upper block,
// row-major, blocks in upper-triangle row order); d_tasks = L.tasks followed by L.row0.
void launch_ozaki_fock(const OzakiLayout &L, const int *d_tasks, const signed char *planes,
                       const int *eA, const unsigned char *maxd1, const double *vs,
                       double *partial, hipStream_t s);
// F [n x n] (symmetric) = the fixed-order sum over chunks of the partials.
void launch_ozaki_reduce(const OzakiLayout &L, const double *partial, double *F, hipStream_t s);

}  // namespace napmo
