// Exact-accumulation grid Fock contraction on int8 MFMA (see ozaki.hpp for the scheme).
//
// v_mfma_i32_16x16x64_i8 operand layout (lane i of 64), the same for both operands, so the
// order of k inside a lane does not matter: A[m][k] row m = i % 16, k = 16 (i / 16) .. + 15;
// B[k][n] column n = i % 16, same k; D[16 x 16] int32: lane i holds D[4 (i / 16) + r][i % 16].
// Operands come from LDS planes [digit][function][64 points] (64 B per function, its four
// 16-byte segments stored in the order q ^ ((f >> 2) & 3), so the 16 functions of a block
// read one segment each without bank conflicts).
//
// Digit split of J (|J| <= 2^34, integer-valued): the low part lo = J - 2^14 h with
// h = floor((J + 2^13) / 2^14), so lo in [-2^13, 2^13) and |h| <= 2^20; then
//   d1 = floor((h + 2^13) / 2^14), r = h - 2^14 d1, d2 = floor((r + 64) / 128), d3 = r - 128 d2,
//   d4 = floor((lo + 64) / 128), d5 = lo - 128 d4,
// every digit in [-64, 64]. h and lo come from FP64 arithmetic that is exact on these values,
// the rest is int32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "ozaki.hpp"

namespace napmo {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

namespace {

// Smallest e with 2^e >= m (m > 0); 0 for m == 0 (all digits are then zero).
__host__ __device__ inline int pow2_ceil_exp(double m) {
  if (!(m > 0.0)) return 0;
  int ex;
  const double mant = frexp(m, &ex);  // m = mant 2^ex, mant in [0.5, 1)
  return mant == 0.5 ? ex - 1 : ex;
}

// J integer-valued, |J| <= 2^34 -> (h, lo) with J = 2^14 h + lo, lo in [-2^13, 2^13).
__device__ inline void split_hl(double J, int &h, int &lo) {
  const double hd = floor(fma(J, 0x1p-14, 0.5));
  h = (int)hd;
  lo = (int)fma(-hd, 16384.0, J);
}

// (h, lo) -> the five digits d[0..4] (d1..d5).
__device__ inline void split5(int h, int lo, int d[OZ_DIGITS]) {
  d[0] = (h + (1 << 13)) >> 14;
  const int r = h - d[0] * (1 << 14);
  d[1] = (r + 64) >> 7;
  d[2] = r - d[1] * 128;
  d[3] = (lo + 64) >> 7;
  d[4] = lo - d[3] * 128;
}

__device__ inline int sbyte(unsigned w, int b) { return (int)(signed char)(w >> (8 * b)); }
__device__ inline unsigned pack4(int a, int b, int c, int d) {
  return (unsigned)(a & 0xFF) | ((unsigned)(b & 0xFF) << 8) | ((unsigned)(c & 0xFF) << 16) |
         ((unsigned)(d & 0xFF) << 24);
}

__device__ inline size_t seg_off(int nf, int plane, int fl, int q) {
  return ((size_t)(plane * nf + fl) * 4 + (q ^ ((fl >> 2) & 3))) * 16;
}

// ------------------------------------------------------------------ per geometry
// eA[f][c] = exponent of the chunk maximum of |sqrt|wbw| phi| (one workgroup per chunk;
// OZ_ZERO_EXP for a function that is zero in the chunk, whose digits are all zero either way).
__global__ void __launch_bounds__(256) k_oz_scales(const double *__restrict__ phi64, int ld,
                                                   const double *__restrict__ wbw, const int *__restrict__ perm,
                                                   long N, long chunk, int nchunk, int *__restrict__ eA) {
  const int c = blockIdx.x;
  const long p0 = (long)c * chunk, p1 = min(N, p0 + chunk);
  for (int f = threadIdx.x; f < ld; f += blockDim.x) {
    const int col = perm ? perm[f] : f;
    double m = 0.0;
    for (long p = p0; p < p1; ++p) m = fmax(m, fabs(sqrt(fabs(wbw[p])) * phi64[p * ld + col]));
    eA[(size_t)f * nchunk + c] = m > 0.0 ? pow2_ceil_exp(m) : OZ_ZERO_EXP;
  }
}

// Digit planes and per-stage max |d1|: one workgroup per (stage of 64 points, 64 functions).
__global__ void __launch_bounds__(256) k_oz_digits(const double *__restrict__ phi64, int ld,
                                                   const double *__restrict__ wbw,
                                                   const int *__restrict__ perm, long N, long Nq,
                                                   long chunk, int nchunk, const int *__restrict__ eA,
                                                   signed char *__restrict__ planes,
                                                   unsigned char *__restrict__ maxd1) {
  __shared__ double tile[OZ_STAGE][65];
  const long p0 = (long)blockIdx.x * OZ_STAGE;
  const int f0 = blockIdx.y * 64;
  const int tid = threadIdx.x;
  for (int e = tid; e < OZ_STAGE * 64; e += 256) {
    const int pr = e / 64, fc = e % 64;
    const long p = p0 + pr;
    const int col = perm ? perm[f0 + fc] : f0 + fc;
    tile[pr][fc] = p < N ? sqrt(fabs(wbw[p])) * phi64[p * ld + col] : 0.0;
  }
  __syncthreads();
  const int fl = tid >> 2, q = tid & 3, f = f0 + fl;
  const int c = (int)(p0 / chunk);
  const int e = eA[(size_t)f * nchunk + c];
  unsigned w[OZ_DIGITS][4];
  int m1 = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const double J = rint(ldexp(tile[16 * q + k][fl], 34 - e));
    int h, lo, d[OZ_DIGITS];
    split_hl(J, h, lo);
    split5(h, lo, d);
    m1 = max(m1, abs(d[0]));
    const int b = k & 3;
#pragma unroll
    for (int x = 0; x < OZ_DIGITS; ++x) {
      if (b == 0) w[x][k >> 2] = 0u;
      w[x][k >> 2] |= (unsigned)(d[x] & 0xFF) << (8 * b);
    }
  }
#pragma unroll
  for (int k = 0; k < OZ_DIGITS; ++k) {
    u4 v = {w[k][0], w[k][1], w[k][2], w[k][3]};
    *(u4 *)&planes[((size_t)k * ld + f) * Nq + p0 + 16 * q] = v;
  }
  m1 = max(m1, __shfl_xor(m1, 1));
  m1 = max(m1, __shfl_xor(m1, 2));
  if (q == 0) maxd1[(size_t)f * (Nq / OZ_STAGE) + blockIdx.x] = (unsigned char)m1;
}

__global__ void k_oz_vs(const double *__restrict__ v, const double *__restrict__ wbw, long N, long Nq,
                        double *__restrict__ vs) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Nq) return;
  vs[p] = p < N ? (wbw[p] < 0.0 ? -v[p] : v[p]) : 0.0;
}

// --------------------------------------------------------------------- per build
struct OzArgs {
  const signed char *planes;
  const int *eA;
  const unsigned char *maxd1;
  const double *vs;
  const int *tasks;   // [ngroups][OZ_WAVES][OZ_TASK]
  const int *row0;    // [ngroups + 1]
  double *partial;
  long Nq, chunk, pstride;
  int ld, nb, nchunk, ngroups, nstage_total;
  int screen, lgchunk;  // skip blocks bounded by 2^-screen (0: none); 2^lgchunk >= chunk
  // packed mode (ptasks != NULL): workgroups take the packer's tasks [*pntask][OZ_PTASK] in
  // turn (*pnext: the next task to take)
  const int *ptasks;
  const int *pntask;
  int *pnext;
};

__device__ inline v4i mfma8(v4i a, v4i b, v4i c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// One 16 x 16 block: rows r[0..4] (digits of B, the v-weighted side), columns c[0..4]; the 15
// digit products of levels 2..6 (1-based) chained into one int32 accumulator per level
// (acc[L - 2] += T_L), so the stage needs no vector instructions.
__device__ inline void block_products(const v4i r[OZ_DIGITS], const v4i c[OZ_DIGITS], v4i acc[5]) {
  acc[0] = mfma8(r[0], c[0], acc[0]);
  acc[1] = mfma8(r[1], c[0], mfma8(r[0], c[1], acc[1]));
  acc[2] = mfma8(r[2], c[0], mfma8(r[1], c[1], mfma8(r[0], c[2], acc[2])));
  acc[3] = mfma8(r[3], c[0], mfma8(r[2], c[1], mfma8(r[1], c[2], mfma8(r[0], c[3], acc[3]))));
  acc[4] = mfma8(r[4], c[0], mfma8(r[3], c[1], mfma8(r[2], c[2], mfma8(r[1], c[3], mfma8(r[0], c[4], acc[4])))));
}

// Four bytes (the low byte of each) packed little-endian into a word.
__device__ inline unsigned pack_bytes(int b0, int b1, int b2, int b3) {
  const unsigned lo = __builtin_amdgcn_perm((unsigned)b1, (unsigned)b0, 0x0c0c0400u);
  const unsigned hi = __builtin_amdgcn_perm((unsigned)b3, (unsigned)b2, 0x0c0c0400u);
  return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

// B digits of the group's listed row blocks (rowl[0..nrow), local indices) for one stage into
// an LDS buffer: one (function, 8 points) item per thread and pass. JB = rint(JA vs 2^-(eV + eB))
// with JA = 2^14 (2^14 d1 + 2^7 d2 + d3) + 2^7 d4 + d5 read from the digit planes.
__device__ inline void stage_b_digits(const OzArgs &a, const int *rowlist, int nB, const int *rowl, int nrow,
                                      long pn, const double *scB, char *buf, int tid) {
  for (int it = tid; it < nrow * 128; it += 64 * OZ_WAVES) {
    const int fl = 16 * rowl[it >> 7] + ((it >> 3) & 15), s8 = it & 7;
    const int f = 16 * rowlist[fl >> 4] + (fl & 15);
    const long p = pn + 8 * s8;
    uint2 dg[OZ_DIGITS];
#pragma unroll
    for (int k = 0; k < OZ_DIGITS; ++k) dg[k] = *(const uint2 *)&a.planes[((size_t)k * a.ld + f) * a.Nq + p];
    const double sc = scB[fl];
    unsigned out[OZ_DIGITS][2];
#pragma unroll
    for (int wd = 0; wd < 2; ++wd) {
      const double2 va = *(const double2 *)&a.vs[p + 4 * wd], vb = *(const double2 *)&a.vs[p + 4 * wd + 2];
      const double vsv[4] = {va.x, va.y, vb.x, vb.y};
      const unsigned w0 = wd ? dg[0].y : dg[0].x, w1 = wd ? dg[1].y : dg[1].x, w2 = wd ? dg[2].y : dg[2].x,
                     w3 = wd ? dg[3].y : dg[3].x, w4 = wd ? dg[4].y : dg[4].x;
      int d[4][OZ_DIGITS];
#pragma unroll
      for (int by = 0; by < 4; ++by) {
        const int hi = sbyte(w0, by) * (1 << 14) + sbyte(w1, by) * 128 + sbyte(w2, by);
        const int lo = sbyte(w3, by) * 128 + sbyte(w4, by);
        const double JA = fma((double)hi, 16384.0, (double)lo);
        const double JB = rint(JA * (vsv[by] * sc));
        int h, l;
        split_hl(JB, h, l);
        split5(h, l, d[by]);
      }
#pragma unroll
      for (int k = 0; k < OZ_DIGITS; ++k) out[k][wd] = pack_bytes(d[0][k], d[1][k], d[2][k], d[3][k]);
    }
#pragma unroll
    for (int k = 0; k < OZ_DIGITS; ++k)
      *(uint2 *)&buf[seg_off(nB, k, fl, s8 >> 1) + 8 * (s8 & 1)] = make_uint2(out[k][0], out[k][1]);
  }
}

// One (chunk c, set of block rows) task of the contraction: the block rows rows[0..nrows) (NULL:
// the contiguous rows r0, r0 + 1, ...) against the columns of the waves' tasks wtask (OZ_TASK
// ints per wave). The B digits of the rows are formed per stage into a double-buffered LDS
// region; each wave reads the A digits of its (at most OZ_MAXCOL) column blocks straight from
// the planes into registers and owns at most OZ_MAXB blocks. The two halves of the workgroup
// (one wave of each per SIMD) run the stage's MFMAs and the next stage's digit formation in
// opposite order, so matrix and vector work of a SIMD overlap; one barrier per stage. With
// screen (static groups only) the blocks bounded by 2^-screen are skipped here; packed tasks
// hold kept blocks only.
__device__ inline void ozaki_task(const OzArgs &a, char *lds, int c, int nrows, const int *rows, int r0,
                                  const int *wtask, bool screen) {
  const int nB = 16 * nrows;
  const size_t bufsz = (size_t)OZ_DIGITS * nB * 64;
  char *ldsB = lds;                                         // [2][OZ_DIGITS][nB][64]
  double *scB = (double *)(lds + 2 * bufsz);                // [nB]
  int *eRow = (int *)(scB + nB);                           // [nB]
  double *maxv = (double *)(eRow + nB + (nB & 1));          // [OZ_MAX_STAGES]
  int *s_eV_ptr = (int *)(maxv + OZ_MAX_STAGES);             // [1] (all LDS is dynamic)
  int *s_nrow = s_eV_ptr + 1;                                // [1]
  int *eBlk = s_nrow + 1;                                    // [nb]: block exponents, this chunk
  int *rowl = eBlk + a.nb;                                   // [nrows]: local row blocks with work
  int *rowlist = rowl + OZ_MAXROWS;                          // [nrows]: local -> block row
  int *rowslot = rowlist + OZ_MAXROWS;                       // [nb]: block row -> local
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ns = (int)(a.chunk / OZ_STAGE);
  const long pc = (long)c * a.chunk;
  const int stage0 = (int)(pc / OZ_STAGE);
  const int rfirst = rows ? rows[0] : r0;
  for (int s = tid; s < nrows; s += 64 * OZ_WAVES) {
    const int I = rows ? rows[s] : r0 + s;
    rowlist[s] = I;
    rowslot[I] = s;
  }

  // ---- prologue: stage maxima of |vs|, the chunk exponent eV, and per B function the
  // scale of its digits: bound = 2^-6 max_t (maxv_t 2^-eV) (maxd1 + 1) >= max |vs At| 2^-(eA+eV)
  for (int t = wave; t < ns; t += OZ_WAVES) {
    double m = fabs(a.vs[pc + (long)t * OZ_STAGE + lane]);
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if (lane == 0) maxv[t] = m;
  }
  __syncthreads();
  if (tid == 0) {
    double m = 0.0;
    for (int t = 0; t < ns; ++t) m = fmax(m, maxv[t]);
    *s_eV_ptr = pow2_ceil_exp(m);
  }
  __syncthreads();
  const int eV = *s_eV_ptr;
  // screening: the largest chunk exponent of each column block's functions, then the row
  // blocks with at least one block left (every row block when screening is off)
  auto skipped = [&](int I, int J) { return eBlk[I] + eBlk[J] + eV + a.lgchunk <= -a.screen; };
  if (screen) {
    for (int b = rfirst + wave; b < a.nb; b += OZ_WAVES) {
      int e = lane < 16 ? a.eA[(size_t)(16 * b + lane) * a.nchunk + c] : OZ_ZERO_EXP;
      for (int o = 8; o > 0; o >>= 1) e = max(e, __shfl_xor(e, o));
      if (lane == 0) eBlk[b] = e;
    }
    __syncthreads();
    if (tid == 0) {
      int nr = 0;
      for (int s = 0; s < nrows; ++s) {
        const int I = rows ? rows[s] : r0 + s;
        bool need = false;
        for (int J = I; J < a.nb && !need; ++J) need = !skipped(I, J);
        if (need) rowl[nr++] = s;
      }
      *s_nrow = nr;
    }
  } else if (tid == 0) {
    for (int s = 0; s < nrows; ++s) rowl[s] = s;
    *s_nrow = nrows;
  }
  __syncthreads();
  const int nrow = *s_nrow;
  for (int fl = tid; fl < nB; fl += 64 * OZ_WAVES) {
    const int f = 16 * rowlist[fl >> 4] + (fl & 15);
    const unsigned char *md = a.maxd1 + (size_t)f * a.nstage_total + stage0;
    double bound = 0.0;
    for (int t = 0; t < ns; ++t) bound = fmax(bound, ldexp(maxv[t], -eV) * (double)(md[t] + 1));
    bound = ldexp(bound, -6);
    const int eB = pow2_ceil_exp(bound);
    scB[fl] = ldexp(1.0, -eV - eB);
    eRow[fl] = a.eA[(size_t)f * a.nchunk + c] + eV + eB;
  }
  __syncthreads();

  v4i acc[OZ_MAXB][5];
#pragma unroll
  for (int k = 0; k < OZ_MAXB; ++k)
#pragma unroll
    for (int L = 0; L < 5; ++L) acc[k][L] = (v4i){0, 0, 0, 0};
  // this wave's task, read once into scalar registers
  const int *task = wtask + (size_t)wave * OZ_TASK;
  int tI[OZ_MAXB], tJ[OZ_MAXB], tC[OZ_MAXB], cJ[OZ_MAXCOL];
#pragma unroll
  for (int k = 0; k < OZ_MAXB; ++k) {
    tI[k] = __builtin_amdgcn_readfirstlane(task[3 * k]);
    tJ[k] = __builtin_amdgcn_readfirstlane(task[3 * k + 1]);
    tC[k] = __builtin_amdgcn_readfirstlane(task[3 * k + 2]);
  }
#pragma unroll
  for (int cc = 0; cc < OZ_MAXCOL; ++cc) cJ[cc] = __builtin_amdgcn_readfirstlane(task[3 * OZ_MAXB + cc]);
  // slots whose block is skipped keep zero accumulators (their partial is written as 0), and
  // columns that no remaining slot uses are not loaded
  bool live[OZ_MAXB];
#pragma unroll
  for (int k = 0; k < OZ_MAXB; ++k)
    live[k] = tI[k] >= 0 && !(screen && __builtin_amdgcn_readfirstlane((int)skipped(tI[k], tJ[k])));
#pragma unroll
  for (int cc = 0; cc < OZ_MAXCOL; ++cc) {
    bool used = false;
#pragma unroll
    for (int k = 0; k < OZ_MAXB; ++k) used |= live[k] && tC[k] == cc;
    if (!used) cJ[cc] = -1;
  }
  const int r16 = lane & 15, q = lane >> 4;
  // waves 0-3 and 4-7 (one of each per SIMD) run the MFMAs of a stage and the digit formation
  // of the next in opposite order; both load their column digits before the formation, whose
  // vector work then covers the load latency
  const bool mfma_first = ((wave >> 2) & 1) == 0;
  v4i cf[OZ_MAXCOL][OZ_DIGITS];
  auto load_columns = [&](long p0) {
#pragma unroll
    for (int cc = 0; cc < OZ_MAXCOL; ++cc) {
      const int J = cJ[cc];
      if (J < 0) continue;
      const signed char *src = a.planes + (size_t)(16 * J + r16) * a.Nq + p0 + 16 * q;
#pragma unroll
      for (int d = 0; d < OZ_DIGITS; ++d) cf[cc][d] = *(const v4i *)(src + (size_t)d * a.ld * a.Nq);
    }
  };
  auto mfma_phase = [&](const char *bufB) {
#pragma unroll
    for (int k = 0; k < OZ_MAXB; ++k) {
      const int I = tI[k], ci = tC[k];
      if (!live[k]) continue;
      v4i ra[OZ_DIGITS];
      const int fl = 16 * rowslot[I] + r16;
#pragma unroll
      for (int d = 0; d < OZ_DIGITS; ++d) ra[d] = *(const v4i *)&bufB[seg_off(nB, d, fl, q)];
      if (ci == 0)
        block_products(ra, cf[0], acc[k]);
      else if (ci == 1)
        block_products(ra, cf[1], acc[k]);
      else
        block_products(ra, cf[2], acc[k]);
    }
  };
  if (nrow > 0) {  // (the same for the whole workgroup)
    stage_b_digits(a, rowlist, nB, rowl, nrow, pc, scB, ldsB, tid);
    __syncthreads();
    if (mfma_first) load_columns(pc);
    for (int t = 0; t < ns; ++t) {
      const long p0 = pc + (long)t * OZ_STAGE;
      const char *bufB = ldsB + (t & 1) * bufsz;
      char *bufN = ldsB + ((t + 1) & 1) * bufsz;
      const bool more = t + 1 < ns;
      if (mfma_first) {
        mfma_phase(bufB);
        if (more) {
          load_columns(p0 + OZ_STAGE);
          stage_b_digits(a, rowlist, nB, rowl, nrow, p0 + OZ_STAGE, scB, bufN, tid);
        }
      } else {
        load_columns(p0);
        if (more) stage_b_digits(a, rowlist, nB, rowl, nrow, p0 + OZ_STAGE, scB, bufN, tid);
        mfma_phase(bufB);
      }
      __syncthreads();
    }
  }
  // ---- epilogue: scaled FP64 partial of every block of this wave
  double *part = a.partial + (size_t)c * a.pstride;
#pragma unroll
  for (int k = 0; k < OZ_MAXB; ++k) {
    const int I = tI[k], J = tJ[k];
    if (I < 0) continue;
    const int blk = I * a.nb - I * (I - 1) / 2 + (J - I);
    const int col = 16 * J + r16;
    const int ec = a.eA[(size_t)col * a.nchunk + c];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = 4 * q + r;
      const double val = (((acc[k][0][r] * 0x1p-12 + acc[k][1][r] * 0x1p-19) + acc[k][2][r] * 0x1p-26) +
                          acc[k][3][r] * 0x1p-33) + acc[k][4][r] * 0x1p-40;
      part[(size_t)blk * 256 + m * 16 + r16] = ldexp(val, eRow[16 * rowslot[I] + m] + ec);
    }
  }
}

// Workgroup = (row group g, chunk c): block rows [r0, r1) against every column block >= r0.
__global__ void __launch_bounds__(64 * OZ_WAVES) k_ozaki_fock(OzArgs a) {
  extern __shared__ __align__(16) char lds[];
  // chunk c on XCD c % 8 for every group (workgroups are dealt to the 8 XCDs in turn)
  const int id = blockIdx.x;
  const int c8 = id % 8, rest = id / 8;
  const int g = rest % a.ngroups, c = (rest / a.ngroups) * 8 + c8;
  if (c >= a.nchunk) return;
  const int r0 = a.row0[g], r1 = a.row0[g + 1];
  ozaki_task(a, lds, c, r1 - r0, nullptr, r0, a.tasks + (size_t)g * OZ_WAVES * OZ_TASK, a.screen > 0);
}

// Packed tasks: every workgroup takes the next task until none is left (the task count was
// written by the packing pass before this launch).
__global__ void __launch_bounds__(64 * OZ_WAVES) k_ozaki_fock_packed(OzArgs a) {
  extern __shared__ __align__(16) char lds[];
  __shared__ int s_task;
  const int total = *a.pntask;
  for (;;) {
    if (threadIdx.x == 0) s_task = atomicAdd(a.pnext, 1);
    __syncthreads();
    const int t = s_task;
    __syncthreads();  // s_task read by all; the previous task's LDS use is over
    if (t >= total) break;
    const int *rec = a.ptasks + (size_t)t * OZ_PTASK;
    ozaki_task(a, lds, rec[0], rec[1], rec + 2, 0, rec + 2 + OZ_MAXROWS, a.screen > 0);
  }
}

__global__ void k_oz_pack_reset(int *ctr) {
  if (threadIdx.x == 0) ctr[0] = 0, ctr[1] = 0;
}

struct OzPackArgs {
  const int *eA;
  const double *vs;
  double *partial;
  int *ptasks, *ctr;
  int *cache, *ccount, *cev;  // per chunk: its last tasks [nb][OZ_PTASK], their count, their eV
  long chunk, pstride;
  int nchunk, nb, screen, lgchunk, rmax;
};

// Packing pass, one wave per chunk: the chunk exponent of |vs| and the block exponents (as the
// contraction forms them), the kept blocks (all of them when screen == 0), zeros for the
// skipped blocks' partials, then (the whole wave, uniformly) the kept block rows in order, grouped greedily into
// tasks of at most rmax rows whose kept blocks, column by column, fit the waves' slots as
// pack_waves fills them; each task is appended to ptasks at an index taken from ctr[0].
__global__ void __launch_bounds__(64) k_oz_pack(OzPackArgs p) {
  __shared__ int eb[OZ_MAXROWS];
  __shared__ int s_eV;
  const int c = blockIdx.x, lane = threadIdx.x, nb = p.nb;
  const long pc = (long)c * p.chunk;
  double m = 0.0;
  // the chunk's stages 8 at a time (independent loads in flight; chunk = 64 * stages)
  const int ns = (int)(p.chunk / OZ_STAGE);
  int t0 = 0;
  for (; t0 + 8 <= ns; t0 += 8) {
    double x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = p.vs[pc + (long)(t0 + k) * OZ_STAGE + lane];
#pragma unroll
    for (int k = 0; k < 8; ++k) m = fmax(m, fabs(x[k]));
  }
  for (; t0 < ns; ++t0) m = fmax(m, fabs(p.vs[pc + (long)t0 * OZ_STAGE + lane]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if (lane == 0) s_eV = pow2_ceil_exp(m);
  if (lane < nb) {
    int e = OZ_ZERO_EXP;
    for (int i = 0; i < 16; ++i) e = max(e, p.eA[(size_t)(16 * lane + i) * p.nchunk + c]);
    eb[lane] = e;
  }
  __syncthreads();
  const int eV = s_eV;
  // live (lane I): bit J = block (I, J) kept; colm (lane J): bit I = block (I, J) kept. The
  // packing below runs on the whole wave with uniform control flow and reads them per lane
  // with readlane, so its sequential steps are register operations.
  unsigned live_r = 0;
  if (lane < nb)
    for (int J = lane; J < nb; ++J)
      if (p.screen <= 0 || eb[lane] + eb[J] + eV + p.lgchunk > -p.screen) live_r |= 1u << J;
  unsigned colm_r = 0;
  for (int I = 0; I < nb; ++I) {
    const unsigned li = (unsigned)__builtin_amdgcn_readlane((int)live_r, I);
    if (I <= lane) colm_r |= ((li >> lane) & 1u) << I;
  }
  auto liveI = [&](int I) { return (unsigned)__builtin_amdgcn_readlane((int)live_r, I); };
  auto colmJ = [&](int J) { return (unsigned)__builtin_amdgcn_readlane((int)colm_r, J); };
  double *part = p.partial + (size_t)c * p.pstride;
  for (int I = 0; I < nb; ++I) {
    const unsigned li = liveI(I);
    for (int J = I; J < nb; ++J)
      if (!((li >> J) & 1)) {
        const int blk = I * nb - I * (I - 1) / 2 + (J - I);
        for (int e = lane; e < 256; e += 64) part[(size_t)blk * 256 + e] = 0.0;
      }
  }
  // the kept blocks of the rows rm, column by column, rows in order, into the waves' slots
  // as pack_waves fills them (out and out2: two copies of the OZ_WAVES * OZ_TASK ints, written
  // by lane 0, or NULL to test the fit)
  auto fits = [&](unsigned rm, int *out, int *out2) {
    if (out && lane == 0)
      for (int e = 0; e < OZ_WAVES * OZ_TASK; ++e) out[e] = -1, out2[e] = -1;
    int total = 0;
    for (int J = 0; J < nb; ++J) total += __popc(colmJ(J) & rm);
    int it = 0, J = 0;
    unsigned mj = colmJ(0) & rm;
    for (int w = 0; w < OZ_WAVES && it < total; ++w) {
      const int left = OZ_WAVES - w;
      const int target = min(OZ_MAXB, (total - it + left - 1) / left);
      int ncol = 0, k = 0, last = -1;
      while (it < total && k < target) {
        while (mj == 0) mj = colmJ(++J) & rm;  // it < total: a kept block remains
        if (J != last) {
          if (ncol == OZ_MAXCOL) break;
          if (out && lane == 0) out[w * OZ_TASK + 3 * OZ_MAXB + ncol] = J, out2[w * OZ_TASK + 3 * OZ_MAXB + ncol] = J;
          last = J, ++ncol;
        }
        const int I = __ffs(mj) - 1;
        mj &= mj - 1;
        if (out && lane == 0) {
          out[w * OZ_TASK + 3 * k] = I, out[w * OZ_TASK + 3 * k + 1] = J, out[w * OZ_TASK + 3 * k + 2] = ncol - 1;
          out2[w * OZ_TASK + 3 * k] = I, out2[w * OZ_TASK + 3 * k + 1] = J, out2[w * OZ_TASK + 3 * k + 2] = ncol - 1;
        }
        ++k, ++it;
      }
    }
    return it == total;
  };
  // a chunk whose |v| exponent has not grown since its tasks were formed reuses them: their
  // blocks include every block kept now, and the contraction's rule skips the others
  {
    const int cev = p.cev[c];
    if (cev != OZ_NO_CACHE && eV <= cev) {
      const int ng = p.ccount[c];
      if (ng == 0) return;
      int tf = 0;
      if (lane == 0) tf = atomicAdd(&p.ctr[0], ng);
      tf = __builtin_amdgcn_readfirstlane(tf);
      const int *src = p.cache + (size_t)c * nb * OZ_PTASK;
      int *dst = p.ptasks + (size_t)tf * OZ_PTASK;
      for (int e = lane; e < ng * OZ_PTASK; e += 64) dst[e] = src[e];
      return;
    }
  }
  auto emit = [&](unsigned rm, int *rec, int *rec2) {  // the task into both records
    if (lane == 0) {
      rec[0] = rec2[0] = c;
      int nr = 0;
      for (unsigned b = rm; b; b &= b - 1, ++nr) rec[2 + nr] = rec2[2 + nr] = __ffs(b) - 1;
      rec[1] = rec2[1] = nr;
      for (int s = nr; s < OZ_MAXROWS; ++s) rec[2 + s] = rec2[2 + s] = -1;
    }
    fits(rm, rec + 2 + OZ_MAXROWS, rec2 + 2 + OZ_MAXROWS);
  };
  // up to 3 blocks per wave always fit (a wave then never needs a fourth column), so the
  // wave filling is simulated only past OZ_WAVES * OZ_MAXCOL kept blocks
  // the row groups first, then one atomic for the chunk's task slots
  unsigned grp[OZ_MAXROWS];
  int ngrp = 0;
  unsigned cur = 0;
  int ncur = 0, tcur = 0;
  for (int I = 0; I < nb; ++I) {
    const unsigned li = liveI(I);
    if (!li) continue;
    const unsigned cand = cur | (1u << I);
    const int tc = tcur + __popc(li);
    if (ncur > 0 && (ncur + 1 > p.rmax || (tc > OZ_WAVES * OZ_MAXCOL && !fits(cand, nullptr, nullptr)))) {
      grp[ngrp++] = cur;
      cur = 1u << I, ncur = 1, tcur = __popc(li);
    } else {
      cur = cand, ++ncur, tcur = tc;
    }
  }
  if (ncur > 0) grp[ngrp++] = cur;
  // the tasks into the build's list and into this chunk's cache slots
  if (lane == 0) p.ccount[c] = ngrp, p.cev[c] = eV;
  if (ngrp == 0) return;
  int tfirst = 0;
  if (lane == 0) tfirst = atomicAdd(&p.ctr[0], ngrp);
  tfirst = __builtin_amdgcn_readfirstlane(tfirst);
  for (int k = 0; k < ngrp; ++k)
    emit(grp[k], p.ptasks + (size_t)(tfirst + k) * OZ_PTASK, p.cache + ((size_t)c * nb + k) * OZ_PTASK);
}

// F[mu][nu] = F[nu][mu] = sum over chunks (in order) of the block element, mu <= nu < n.
__global__ void k_ozaki_reduce(const double *__restrict__ partial, long pstride, int nchunk, int nb,
                               int n, const int *__restrict__ perm, double *__restrict__ F) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const int mu = (int)(e / n), nu = (int)(e % n);
  if (mu > nu) return;
  const int I = mu >> 4, J = nu >> 4;
  const long off = ((long)(I * nb - I * (I - 1) / 2 + (J - I))) * 256 + (mu & 15) * 16 + (nu & 15);
  // loads issued 8 at a time, added in chunk order
  double s = 0.0;
  int c = 0;
  for (; c + 8 <= nchunk; c += 8) {
    double x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = partial[(long)(c + j) * pstride + off];
#pragma unroll
    for (int j = 0; j < 8; ++j) s += x[j];
  }
  for (; c < nchunk; ++c) s += partial[(long)c * pstride + off];
  const int x = perm ? perm[mu] : mu, y = perm ? perm[nu] : nu;
  F[(long)x * n + y] = s;
  F[(long)y * n + x] = s;
}

}  // namespace

// ------------------------------------------------------------------------ host
// Blocks (I, J) of rows [r0, r1) x columns [r0, nb), I <= J, column-major, cut into OZ_WAVES
// contiguous runs of at most wave_blocks blocks and OZ_MAXCOL distinct columns, balanced.
static bool pack_waves(int r0, int r1, int nb, int wave_blocks, std::vector<int> &tab) {
  std::vector<std::pair<int, int>> blocks;
  for (int J = r0; J < nb; ++J)
    for (int I = r0; I < r1 && I <= J; ++I) blocks.push_back({I, J});
  const int total = (int)blocks.size();
  tab.assign((size_t)OZ_WAVES * OZ_TASK, -1);
  size_t it = 0;
  for (int w = 0; w < OZ_WAVES && it < blocks.size(); ++w) {
    const int left = OZ_WAVES - w;
    const int target = std::min(wave_blocks, (int)((total - (int)it + left - 1) / left));
    int *t = &tab[(size_t)w * OZ_TASK];
    int *cols = t + 3 * OZ_MAXB;
    int ncol = 0, k = 0;
    while (it < blocks.size() && k < target) {
      const int I = blocks[it].first, J = blocks[it].second;
      int ci = -1;
      for (int x = 0; x < ncol; ++x)
        if (cols[x] == J) ci = x;
      if (ci < 0) {
        if (ncol == OZ_MAXCOL) break;
        cols[ncol] = J;
        ci = ncol++;
      }
      t[3 * k] = I, t[3 * k + 1] = J, t[3 * k + 2] = ci;
      ++k, ++it;
    }
  }
  return it == blocks.size();
}

static size_t group_lds(int r0, int r1, int nb) {
  const size_t nB = 16 * (size_t)(r1 - r0);
  return 2 * OZ_DIGITS * nB * 64 + nB * 12 + 8 + OZ_MAX_STAGES * 8 + 16 + 8 * (size_t)nb + 8 * OZ_MAXROWS;
}

OzakiLayout ozaki_layout(long N, int n, int ld, int wave_blocks, size_t max_lds, long chunk_target) {
  NAPMO_REQUIRE(N > 0 && n > 0, "ozaki layout: empty grid or basis");
  NAPMO_REQUIRE(wave_blocks >= 1 && wave_blocks <= OZ_MAXB, "ozaki layout: blocks per wave out of range");
  OzakiLayout L;
  L.N = N, L.n = n, L.ld = ld;
  L.nb = (int)div_up(n, 16);
  NAPMO_REQUIRE(16 * L.nb <= ld && ld % 64 == 0, "ozaki layout: leading dimension too small");
  L.nblk = L.nb * (L.nb + 1) / 2;
  long chunk = chunk_target > 0 ? chunk_target : div_up(N, 256);
  chunk = std::min<long>(std::max<long>(round_up(chunk, OZ_STAGE), OZ_STAGE), (long)OZ_MAX_STAGES * OZ_STAGE);
  L.chunk = chunk;
  L.nchunk = (int)div_up(N, chunk);
  L.Nq = (long)L.nchunk * chunk;
  // row groups: from the first free block row, the most rows whose blocks the waves take
  const int nb = L.nb;
  std::vector<int> row0{0};
  std::vector<int> tasks, tab;
  size_t lds = 0;
  int r0 = 0;
  while (r0 < nb) {
    int best = -1;
    for (int r1 = r0 + 1; r1 <= nb; ++r1) {
      if (group_lds(r0, r1, nb) > max_lds || !pack_waves(r0, r1, nb, wave_blocks, tab)) break;
      best = r1;
    }
    if (best < 0)
      throw Error("int8 Fock contraction: basis of " + std::to_string(n) +
                  " functions does not fit the wave slots (use the FP64 mode)");
    pack_waves(r0, best, nb, wave_blocks, tab);
    tasks.insert(tasks.end(), tab.begin(), tab.end());
    lds = std::max(lds, group_lds(r0, best, nb));
    row0.push_back(best);
    r0 = best;
  }
  L.ngroups = (int)row0.size() - 1;
  L.row0 = row0;
  L.tasks = tasks;
  L.lds_bytes = lds;
  // packed tasks: as many block rows as the LDS takes (16 bytes left for the task index)
  for (int R = std::min(nb, OZ_MAXROWS); R >= 1; --R)
    if (group_lds(0, R, nb) + 16 <= max_lds) {
      L.pack_rows = R;
      L.pack_lds = group_lds(0, R, nb);
      break;
    }
  return L;
}

void launch_ozaki_scales(const OzakiLayout &L, const double *phi64, const double *wbw, const int *perm,
                         int *eA, hipStream_t s) {
  k_oz_scales<<<L.nchunk, 256, 0, s>>>(phi64, L.ld, wbw, perm, L.N, L.chunk, L.nchunk, eA);
  NAPMO_HIP(hipGetLastError());
}

void launch_ozaki_setup(const OzakiLayout &L, const double *phi64, const double *wbw, const int *perm,
                        signed char *planes, int *eA, unsigned char *maxd1, hipStream_t s) {
  launch_ozaki_scales(L, phi64, wbw, perm, eA, s);
  k_oz_digits<<<dim3((unsigned)(L.Nq / OZ_STAGE), L.ld / 64), 256, 0, s>>>(
      phi64, L.ld, wbw, perm, L.N, L.Nq, L.chunk, L.nchunk, eA, planes, maxd1);
  NAPMO_HIP(hipGetLastError());
}

long ozaki_kept_blocks(const OzakiLayout &L, const std::vector<int> &eA, const std::vector<double> &vs,
                       int screen) {
  NAPMO_REQUIRE(eA.size() == (size_t)L.ld * L.nchunk && vs.size() == (size_t)L.Nq,
                "int8 screening count: table sizes");
  if (screen <= 0) return (long)L.nchunk * L.nblk;
  int lg = 0;
  while ((1L << lg) < L.chunk) ++lg;
  long kept = 0;
  std::vector<int> eb(L.nb);
  for (int c = 0; c < L.nchunk; ++c) {
    double m = 0.0;
    for (long p = (long)c * L.chunk; p < (long)(c + 1) * L.chunk; ++p) m = std::max(m, std::fabs(vs[p]));
    const int eV = pow2_ceil_exp(m);
    for (int b = 0; b < L.nb; ++b) {
      eb[b] = OZ_ZERO_EXP;
      for (int i = 0; i < 16; ++i) eb[b] = std::max(eb[b], eA[(size_t)(16 * b + i) * L.nchunk + c]);
    }
    for (int I = 0; I < L.nb; ++I)
      for (int J = I; J < L.nb; ++J) kept += eb[I] + eb[J] + eV + lg > -screen;
  }
  return kept;
}

std::vector<int> ozaki_compact_order(const OzakiLayout &L, const std::vector<int> &eA) {
  NAPMO_REQUIRE(eA.size() == (size_t)L.ld * L.nchunk, "int8 screening order: exponent table size");
  int top = OZ_ZERO_EXP;
  for (int f = 0; f < L.n; ++f)
    for (int c = 0; c < L.nchunk; ++c) top = std::max(top, eA[(size_t)f * L.nchunk + c]);
  std::vector<int> reach(L.n, 0), order(L.ld);
  for (int f = 0; f < L.n; ++f)
    for (int c = 0; c < L.nchunk; ++c) reach[f] += eA[(size_t)f * L.nchunk + c] >= top - 30;
  for (int f = 0; f < L.ld; ++f) order[f] = f;
  std::stable_sort(order.begin(), order.begin() + L.n, [&](int x, int y) { return reach[x] < reach[y]; });
  return order;
}

void launch_ozaki_vs(const double *v, const double *wbw, long N, long Nq, double *vs, hipStream_t s) {
  k_oz_vs<<<div_up(Nq, 256), 256, 0, s>>>(v, wbw, N, Nq, vs);
  NAPMO_HIP(hipGetLastError());
}

void launch_ozaki_fock(const OzakiLayout &L, const int *d_tasks, const signed char *planes,
                       const int *eA, const unsigned char *maxd1, const double *vs, int screen,
                       double *partial, hipStream_t s) {
  static bool attr = false;
  if (!attr) {
    NAPMO_HIP(hipFuncSetAttribute((const void *)k_ozaki_fock, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)max_lds_per_block()));
    attr = true;
  }
  NAPMO_REQUIRE(L.lds_bytes <= max_lds_per_block(), "int8 Fock contraction: LDS layout too large");
  OzArgs a;
  a.planes = planes, a.eA = eA, a.maxd1 = maxd1, a.vs = vs;
  a.tasks = d_tasks;
  a.row0 = d_tasks + (size_t)L.ngroups * OZ_WAVES * OZ_TASK;
  a.partial = partial;
  a.Nq = L.Nq, a.chunk = L.chunk, a.pstride = L.partial_stride();
  a.ld = L.ld, a.nb = L.nb, a.nchunk = L.nchunk, a.ngroups = L.ngroups;
  a.nstage_total = (int)(L.Nq / OZ_STAGE);
  NAPMO_REQUIRE(screen >= 0 && screen <= 1000, "int8 screening threshold out of range");
  a.screen = screen;
  a.lgchunk = 0;
  while ((1L << a.lgchunk) < L.chunk) ++a.lgchunk;
  a.ptasks = nullptr, a.pntask = nullptr, a.pnext = nullptr;
  const int nc8 = (int)round_up(L.nchunk, 8);
  k_ozaki_fock<<<nc8 * L.ngroups, 64 * OZ_WAVES, L.lds_bytes, s>>>(a);
  NAPMO_HIP(hipGetLastError());
}

void launch_ozaki_fock_packed(const OzakiLayout &L, const signed char *planes, const int *eA,
                              const unsigned char *maxd1, const double *vs, int screen, double *partial,
                              int *ptasks, int *ctr, int *cache, int *ccount, int *cev, hipStream_t s) {
  static bool attr = false;
  if (!attr) {
    // the kernel's own 4-byte task slot is static LDS: leave 16 bytes for it
    NAPMO_HIP(hipFuncSetAttribute((const void *)k_ozaki_fock_packed, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)max_lds_per_block() - 16));
    attr = true;
  }
  NAPMO_REQUIRE(L.pack_rows >= 1 && L.nb <= OZ_MAXROWS && L.pack_lds + 16 <= max_lds_per_block(),
                "int8 Fock contraction: packed tasks do not fit");
  NAPMO_REQUIRE(screen >= 0 && screen <= 1000, "int8 screening threshold out of range");
  int lg = 0;
  while ((1L << lg) < L.chunk) ++lg;
  k_oz_pack_reset<<<1, 64, 0, s>>>(ctr);
  NAPMO_HIP(hipGetLastError());
  OzPackArgs pa;
  pa.eA = eA, pa.vs = vs, pa.partial = partial, pa.ptasks = ptasks, pa.ctr = ctr;
  pa.cache = cache, pa.ccount = ccount, pa.cev = cev;
  pa.chunk = L.chunk, pa.pstride = L.partial_stride();
  pa.nchunk = L.nchunk, pa.nb = L.nb, pa.screen = screen, pa.lgchunk = lg, pa.rmax = L.pack_rows;
  k_oz_pack<<<L.nchunk, 64, 0, s>>>(pa);
  NAPMO_HIP(hipGetLastError());
  OzArgs a;
  a.planes = planes, a.eA = eA, a.maxd1 = maxd1, a.vs = vs;
  a.tasks = nullptr, a.row0 = nullptr;
  a.partial = partial;
  a.Nq = L.Nq, a.chunk = L.chunk, a.pstride = L.partial_stride();
  a.ld = L.ld, a.nb = L.nb, a.nchunk = L.nchunk, a.ngroups = L.ngroups;
  a.nstage_total = (int)(L.Nq / OZ_STAGE);
  a.screen = screen, a.lgchunk = lg;
  a.ptasks = ptasks, a.pntask = ctr, a.pnext = ctr + 1;
  int dev = 0, cus = 0;
  NAPMO_HIP(hipGetDevice(&dev));
  NAPMO_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  k_ozaki_fock_packed<<<cus, 64 * OZ_WAVES, L.pack_lds, s>>>(a);
  NAPMO_HIP(hipGetLastError());
}

void launch_ozaki_reduce(const OzakiLayout &L, const double *partial, const int *perm, double *F,
                         hipStream_t s) {
  const long nn = (long)L.n * L.n;
  k_ozaki_reduce<<<div_up(nn, 256), 256, 0, s>>>(partial, L.partial_stride(), L.nchunk, L.nb, L.n, perm, F);
  NAPMO_HIP(hipGetLastError());
}

}  // namespace napmo
