// Mixed-precision and orbital-form contractions for gfx950:
//   * Phi split once per geometry into function-major bf16 hi / lo and float32 copies;
//   * F = Phi^T diag(s) Phi with the 3-term bf16 split on v_mfma_f32_16x16x32_bf16,
//     FP32 accumulation over blocks of 256 points flushed into FP64 registers;
//   * rho_p = eta sum_i (Phi C)_pi^2 with FP64 MFMA (orbital form of the density);
//   * P = eta C C^T; an MFMA lane-layout self test.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace napmo {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ inline unsigned short bf16_bits(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf16_float(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

// ---------------------------------------------------------- per-geometry splits
// phi64 [Npad x ld] point-major -> hi, lo (bf16 bits) and f32, function-major [ld x Npad].
__global__ void k_split_phi(const double *__restrict__ phi64, int ld, long Npad,
                            unsigned short *__restrict__ hi, unsigned short *__restrict__ lo,
                            float *__restrict__ f32) {
  __shared__ double tile[32][33];
  const long p0 = (long)blockIdx.x * 32;
  const int m0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) tile[r][tx] = phi64[(p0 + r) * ld + m0 + tx];
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const double v = tile[tx][r];  // point p0 + tx, function m0 + r
    const float x = (float)v;
    const unsigned short h = bf16_bits(x);
    const unsigned short l = bf16_bits(x - bf16_float(h));
    const size_t o = (size_t)(m0 + r) * Npad + p0 + tx;
    hi[o] = h;
    lo[o] = l;
    f32[o] = x;
  }
}

void launch_split_phi(const double *phi64, int ld, long Npad, unsigned short *hi, unsigned short *lo,
                      float *f32, hipStream_t s) {
  NAPMO_REQUIRE(ld % 32 == 0 && Npad % 32 == 0, "split_phi: padding mismatch");
  k_split_phi<<<dim3(Npad / 32, ld / 32), 256, 0, s>>>(phi64, ld, Npad, hi, lo, f32);
  NAPMO_HIP(hipGetLastError());
}

// ------------------------------------------------------------- BF16x3 Fock build
// v_mfma_f32_16x16x32_bf16 operand layout (lane i of 64):
//   A[16 x 32]: lane i holds A[i % 16][8 (i / 16) .. 8 (i / 16) + 7]
//   B[32 x 16]: lane i holds B[8 (i / 16) .. + 7][i % 16]
//   D[16 x 16]: lane i holds D[4 (i / 16) + r][i % 16], r = 0..3
// A[m][k] = Phi[p0+k][mu0+m], B[k][n] = s[p0+k] Phi[p0+k][nu0+n]; both staged in LDS
// k-contiguous ([row][k]) from the function-major copies. Products hi*hi + hi*lo + lo*hi.
constexpr int KB = 32;         // points per LDS stage (MFMA k-depth)
constexpr int KPAD = KB + 8;   // LDS row stride in bf16
constexpr int FLUSH = 256;     // points accumulated in FP32 before the FP64 flush

__global__ void __launch_bounds__(256) k_fock_bf16x3(const unsigned short *__restrict__ Ahi,
                                                     const unsigned short *__restrict__ Alo,
                                                     const float *__restrict__ B32,
                                                     const double *__restrict__ s, int nt,
                                                     int symmetric, long Npad, long chunk,
                                                     double *__restrict__ partial) {
  __shared__ unsigned short sAh[TILE][KPAD], sAl[TILE][KPAD], sBh[TILE][KPAD], sBl[TILE][KPAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int ti, tj;
  if (!symmetric) {
    ti = blockIdx.x / nt, tj = blockIdx.x % nt;
  } else {
    int i = 0, rem = blockIdx.x;
    while (rem >= nt - i) rem -= nt - i, ++i;
    ti = i, tj = i + rem;
  }
  const int mu0 = ti * TILE, nu0 = tj * TILE;
  const long pbeg = (long)blockIdx.y * chunk;
  const long pend = min(Npad, pbeg + chunk);
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  f4 acc[2][2];
  double acc64[2][2][4];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      acc[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
      for (int r = 0; r < 4; ++r) acc64[a][b][r] = 0.0;
    }
  // staging: thread t copies 8 consecutive k of row t / 4 (k segment (t % 4) * 8)
  const int srow = tid >> 2, sk = (tid & 3) * 8;
  int since_flush = 0;
  for (long k0 = pbeg; k0 < pend; k0 += KB) {
    {
      const size_t ga = (size_t)(mu0 + srow) * Npad + k0 + sk;
      *(u16x8 *)&sAh[srow][sk] = *(const u16x8 *)&Ahi[ga];
      *(u16x8 *)&sAl[srow][sk] = *(const u16x8 *)&Alo[ga];
      const size_t gb = (size_t)(nu0 + srow) * Npad + k0 + sk;
      u16x8 bh, bl;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float x = B32[gb + e] * (float)s[k0 + sk + e];
        const unsigned short h = bf16_bits(x);
        bh[e] = h;
        bl[e] = bf16_bits(x - bf16_float(h));
      }
      *(u16x8 *)&sBh[srow][sk] = bh;
      *(u16x8 *)&sBl[srow][sk] = bl;
    }
    __syncthreads();
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      ah[a] = __builtin_bit_cast(bf16x8, *(const u16x8 *)&sAh[wr + a * 16 + r16][kq]);
      al[a] = __builtin_bit_cast(bf16x8, *(const u16x8 *)&sAl[wr + a * 16 + r16][kq]);
      bh[a] = __builtin_bit_cast(bf16x8, *(const u16x8 *)&sBh[wc + a * 16 + r16][kq]);
      bl[a] = __builtin_bit_cast(bf16x8, *(const u16x8 *)&sBl[wc + a * 16 + r16][kq]);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
      }
    __syncthreads();
    since_flush += KB;
    if (since_flush >= FLUSH) {
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          for (int r = 0; r < 4; ++r) acc64[a][b][r] += (double)acc[a][b][r];
          acc[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
        }
      since_flush = 0;
    }
  }
  double *out = partial + (long)blockIdx.y * ((long)gridDim.x * TILE * TILE + 32) +
                (long)blockIdx.x * TILE * TILE;
  const int col16 = lane & 15, rg = lane >> 4;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int r = 0; r < 4; ++r)
        out[(wr + a * 16 + rg * 4 + r) * TILE + wc + b * 16 + col16] =
            acc64[a][b][r] + (double)acc[a][b][r];
}

void launch_fock_bf16x3(const FockShape &sh, const unsigned short *Ahi, const unsigned short *Alo,
                        const float *B32, const double *s, double *partial, hipStream_t st) {
  NAPMO_REQUIRE(sh.chunk % KB == 0 && sh.Npad % KB == 0, "fock_bf16x3: chunking mismatch");
  dim3 grid(sh.ntiles, sh.nchunk);
  k_fock_bf16x3<<<grid, 256, 0, st>>>(Ahi, Alo, B32, s, sh.nt, sh.symmetric ? 1 : 0, sh.Npad,
                                      sh.chunk, partial);
  NAPMO_HIP(hipGetLastError());
}

// ------------------------------------------- BF16x3 Fock build, one pass over Phi
// One workgroup per (chunk of points, group of up to 16 upper 64 x 64 tiles) computes
// those tiles of F for the chunk, so each point's Phi row is read from HBM once per
// tile group (once per build for ld <= 320). Per stage of 32 points the chunk's hi / lo
// rows of all ld functions go to LDS ([function][k]); the B-operand split of s Phi is
// formed once per element while staging, and the next stage's segments are prefetched
// into registers while the MFMAs of the current one run. Wave w accumulates tiles
// 2w, 2w + 1 of its group (4 x 4 blocks of 16 x 16 each) in FP32 over the chunk; the
// FP32 chunk partials are summed in FP64, in fixed order, by k_fock_reduce.
constexpr int TPW = 2;          // tiles per wave
constexpr int PASS_WAVES = 8;   // waves per workgroup (256 VGPRs each)
constexpr int PASS_UNITS = 3;   // 8-point staging segments per thread (ld * 4 / 512 <= 3)

__global__ void __launch_bounds__(512) k_fock_bf16x3_pass(const unsigned short *__restrict__ Ahi,
                                                          const unsigned short *__restrict__ Alo,
                                                          const double *__restrict__ s, int ld,
                                                          int nt, int ntiles, long Npad,
                                                          long chunk,
                                                          float *__restrict__ partial) {
  extern __shared__ unsigned short lds[];
  unsigned short *sAh = lds, *sAl = lds + ld * KPAD, *sBh = lds + 2 * ld * KPAD,
                 *sBl = lds + 3 * ld * KPAD;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tile0 = blockIdx.y * PASS_WAVES * TPW;
  int tr[TPW], tc[TPW];
  bool has[TPW];
#pragma unroll
  for (int k = 0; k < TPW; ++k) {
    const int t = tile0 + wave * TPW + k;
    has[k] = t < ntiles;
    int i = 0, rem = has[k] ? t : 0;
    while (rem >= nt - i) rem -= nt - i, ++i;
    tr[k] = i * TILE, tc[k] = (i + rem) * TILE;
  }
  f4 acc[TPW][4][4];
#pragma unroll
  for (int k = 0; k < TPW; ++k)
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[k][a][b] = (f4){0.f, 0.f, 0.f, 0.f};
  const long pbeg = (long)blockIdx.x * chunk;
  const long pend = min(Npad, pbeg + chunk);
  const int units = ld * (KB / 8);
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  __shared__ float sS[KB];
  // staging units of this thread; past the last unit a thread repeats the last one (same
  // values to the same LDS slots), so the staging code has no branches
  constexpr int NTHREADS = 64 * PASS_WAVES;
  int unit[PASS_UNITS];
#pragma unroll
  for (int q = 0; q < PASS_UNITS; ++q) unit[q] = min(tid + q * NTHREADS, units - 1);
  u16x8 ph[PASS_UNITS], pl[PASS_UNITS];
  double pref_s = 0.0;
  auto prefetch = [&](long k0) {
    if (tid < KB) pref_s = s[k0 + tid];
#pragma unroll
    for (int q = 0; q < PASS_UNITS; ++q) {
      const size_t g = (size_t)(unit[q] >> 2) * Npad + k0 + ((unit[q] & 3) << 3);
      ph[q] = *(const u16x8 *)&Ahi[g];
      pl[q] = *(const u16x8 *)&Alo[g];
    }
  };
  if (pbeg < pend) prefetch(pbeg);
  for (long k0 = pbeg; k0 < pend; k0 += KB) {
    if (tid < KB) sS[tid] = (float)pref_s;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PASS_UNITS; ++q) {
      const int f = unit[q] >> 2, seg = (unit[q] & 3) << 3;
      *(u16x8 *)&sAh[f * KPAD + seg] = ph[q];
      *(u16x8 *)&sAl[f * KPAD + seg] = pl[q];
      u16x8 bh, bl;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float x = (bf16_float(ph[q][e]) + bf16_float(pl[q][e])) * sS[seg + e];
        const unsigned short hb = bf16_bits(x);
        bh[e] = hb;
        bl[e] = bf16_bits(x - bf16_float(hb));
      }
      *(u16x8 *)&sBh[f * KPAD + seg] = bh;
      *(u16x8 *)&sBl[f * KPAD + seg] = bl;
    }
    __syncthreads();
    if (k0 + KB < pend) prefetch(k0 + KB);
#pragma unroll
    for (int k = 0; k < TPW; ++k) {
      if (has[k]) {
        bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int ra = (tr[k] + a * 16 + r16) * KPAD + kq;
          const int rb = (tc[k] + a * 16 + r16) * KPAD + kq;
          ah[a] = __builtin_bit_cast(bf16x8, *(const u16x8 *)&sAh[ra]);
          al[a] = __builtin_bit_cast(bf16x8, *(const u16x8 *)&sAl[ra]);
          bh[a] = __builtin_bit_cast(bf16x8, *(const u16x8 *)&sBh[rb]);
          bl[a] = __builtin_bit_cast(bf16x8, *(const u16x8 *)&sBl[rb]);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            acc[k][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[k][a][b], 0, 0, 0);
            acc[k][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[k][a][b], 0, 0, 0);
            acc[k][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[k][a][b], 0, 0, 0);
          }
      }
    }
    __syncthreads();
  }
  const int col16 = lane & 15, rg = lane >> 4;
#pragma unroll
  for (int k = 0; k < TPW; ++k) {
    if (!has[k]) continue;
    float *out = partial + (long)blockIdx.x * ((long)ntiles * TILE * TILE + 32) +
                 (long)(tile0 + wave * TPW + k) * TILE * TILE;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          out[(a * 16 + rg * 4 + r) * TILE + b * 16 + col16] = acc[k][a][b][r];
  }
}

FockShape fock_pass_shape(int ld, long Npad, long wg_target) {
  FockShape sh;
  sh.ld = ld;
  sh.nt = ld / TILE;
  sh.ntiles = sh.nt * (sh.nt + 1) / 2;
  sh.symmetric = true;
  sh.Npad = Npad;
  const int groups = (int)div_up(sh.ntiles, PASS_WAVES * TPW);
  // about 256 workgroups (or wg_target) over all tile groups, chunks a multiple of 32 points
  long chunk = round_up(div_up(Npad * groups, wg_target > 0 ? wg_target : 256), KB);
  sh.chunk = std::max<long>(chunk, KB);
  sh.nchunk = (int)div_up(Npad, sh.chunk);
  return sh;
}

bool fock_pass_supported(int ld) {
  const size_t lds = (size_t)4 * ld * KPAD * sizeof(unsigned short) + KB * sizeof(float);
  return lds <= max_lds_per_block() && ld * (KB / 8) <= PASS_UNITS * 64 * PASS_WAVES;
}

void launch_fock_bf16x3_pass(const FockShape &sh, const unsigned short *Ahi,
                             const unsigned short *Alo, const double *s, float *partial,
                             hipStream_t st) {
  NAPMO_REQUIRE(fock_pass_supported(sh.ld), "fock pass: basis too large for one workgroup");
  const size_t lds = (size_t)4 * sh.ld * KPAD * sizeof(unsigned short);
  const int groups = (int)div_up(sh.ntiles, PASS_WAVES * TPW);
  k_fock_bf16x3_pass<<<dim3(sh.nchunk, groups), dim3(64 * PASS_WAVES), lds, st>>>(
      Ahi, Alo, s, sh.ld, sh.nt, sh.ntiles, sh.Npad, sh.chunk, partial);
  NAPMO_HIP(hipGetLastError());
}

// ------------------------------------- BF16x3 Fock build, symmetric one-pass form
// F = Phi^T diag(s) Phi = G^T diag(sigma) G with G = diag(sqrt|s|) Phi and sigma = sign(s),
// so one operand matrix serves both MFMA sides: per stage of 32 points only G's bf16 hi / lo
// go to LDS (half the staging of the two-operand pass), which leaves room for two LDS
// buffers: the next stage is converted and stored while the MFMAs of the current one run,
// with one barrier per stage. sigma enters as a sign flip of the B operand's K entries
// (exact in bf16). Products per 16 x 16 block: Gl_a^T Gh_b + Gh_a^T Gl_b + Gh_a^T Gh_b (and
// Gl_a^T Gl_a on the diagonal blocks, where that dropped term would be a square), FP32
// accumulation per chunk, the same partial layout and reduction as the pass above.
// Measured on 472,000 points: the MFMA's FP32 accumulation lets sums of one-signed terms
// drift low as a chunk grows (the P-weighted trace of F moves by -6e-4 Eh at 1,844 points
// per chunk, -1e-4 at 461), so this form is meant for short chunks (1024+ workgroups),
// where it is also the more accurate one (max|dF| 1.3e-5 against 8.5e-5).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(512) k_fock_syrk_pass(const float *__restrict__ phi32,
                                                        const double *__restrict__ s, int ld,
                                                        int nt, int ntiles, long Npad, long chunk,
                                                        float *__restrict__ partial) {
  extern __shared__ unsigned short lds[];  // [buffer][hi, lo][ld][KPAD]
  const size_t arr = (size_t)ld * KPAD;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tile0 = blockIdx.y * PASS_WAVES * TPW;
  int tr[TPW], tc[TPW];
  bool has[TPW];
#pragma unroll
  for (int k = 0; k < TPW; ++k) {
    const int t = tile0 + wave * TPW + k;
    has[k] = t < ntiles;
    int i = 0, rem = has[k] ? t : 0;
    while (rem >= nt - i) rem -= nt - i, ++i;
    tr[k] = i * TILE, tc[k] = (i + rem) * TILE;
  }
  f4 acc[TPW][4][4];
#pragma unroll
  for (int k = 0; k < TPW; ++k)
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[k][a][b] = (f4){0.f, 0.f, 0.f, 0.f};
  const long pbeg = (long)blockIdx.x * chunk;
  const long pend = min(Npad, pbeg + chunk);
  const int units = ld * (KB / 8);
  const int seg = (tid & 3) * 8;  // every unit of this thread covers points seg .. seg + 7
                                  // (a repeated last unit has seg = 24 = (units - 1) & 3)
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  // staging units of this thread; past the last unit a thread repeats the last one (same
  // values to the same LDS slots), so the staging code has no branches
  constexpr int NTHREADS = 64 * PASS_WAVES;
  int unit[PASS_UNITS];
#pragma unroll
  for (int q = 0; q < PASS_UNITS; ++q) unit[q] = min(tid + q * NTHREADS, units - 1);
  f4 ph[PASS_UNITS][2];
  double sv = 0.0;  // s of stage point (lane & 31)
  auto load = [&](long k0) {
    sv = s[k0 + (lane & 31)];
#pragma unroll
    for (int q = 0; q < PASS_UNITS; ++q) {
      const float *src = phi32 + (size_t)(unit[q] >> 2) * Npad + k0 + ((unit[q] & 3) << 3);
      ph[q][0] = *(const f4 *)src;
      ph[q][1] = *(const f4 *)(src + 4);
    }
  };
  // G = sqrt|s| Phi split into bf16 hi / lo for the loaded stage; returns the stage's sign
  // mask (bit k: s of point k is negative)
  auto convert = [&](int buf) -> unsigned {
    const float as = sqrtf((float)fabs(sv));
    const unsigned mask = (unsigned)__ballot(sv < 0.0);
    float r[8], rl[8];  // scales of this thread's segment and of the last unit's (24 .. 31)
#pragma unroll
    for (int e = 0; e < 8; ++e) r[e] = __shfl(as, seg + e), rl[e] = __shfl(as, 24 + e);
    unsigned short *gh = lds + (size_t)(2 * buf) * arr, *gl = gh + arr;
#pragma unroll
    for (int q = 0; q < PASS_UNITS; ++q) {
      const int u = unit[q];
      const int useg = (u & 3) << 3;
      u16x8 hb, lb;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float x = ph[q][e >> 2][e & 3] * (useg == seg ? r[e] : rl[e]);
        const unsigned short h = bf16_bits(x);
        hb[e] = h;
        lb[e] = bf16_bits(x - bf16_float(h));
      }
      const int off = (u >> 2) * KPAD + useg;
      *(u16x8 *)&gh[off] = hb;
      *(u16x8 *)&gl[off] = lb;
    }
    return mask;
  };
  auto mfma_stage = [&](int buf, unsigned mask) {
    // sign flips of this lane's B entries (points kq .. kq + 7), two bf16 per dword
    const unsigned bits = (mask >> kq) & 0xFFu;
    u32x4 flip;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      flip[j] = (((bits >> (2 * j)) & 1u) << 15) | (((bits >> (2 * j + 1)) & 1u) << 31);
    const unsigned short *gh = lds + (size_t)(2 * buf) * arr, *gl = gh + arr;
#pragma unroll
    for (int k = 0; k < TPW; ++k) {
      if (has[k]) {
        bf16x8 ah[4], al[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int ra = (tr[k] + a * 16 + r16) * KPAD + kq;
          ah[a] = __builtin_bit_cast(bf16x8, *(const u16x8 *)&gh[ra]);
          al[a] = __builtin_bit_cast(bf16x8, *(const u16x8 *)&gl[ra]);
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int rb = (tc[k] + b * 16 + r16) * KPAD + kq;
          const bf16x8 bh = __builtin_bit_cast(bf16x8, *(const u32x4 *)&gh[rb] ^ flip);
          const bf16x8 bl = __builtin_bit_cast(bf16x8, *(const u32x4 *)&gl[rb] ^ flip);
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            acc[k][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh, acc[k][a][b], 0, 0, 0);
            acc[k][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl, acc[k][a][b], 0, 0, 0);
            if (a == b && tr[k] == tc[k])  // the square term of the diagonal blocks
              acc[k][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bl, acc[k][a][b], 0, 0, 0);
            acc[k][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh, acc[k][a][b], 0, 0, 0);
          }
        }
      }
    }
  };
  unsigned mask = 0;
  if (pbeg < pend) {
    load(pbeg);
    mask = convert(0);
  }
  __syncthreads();
  int buf = 0;
  for (long k0 = pbeg; k0 < pend; k0 += KB) {
    const bool next = k0 + KB < pend;
    if (next) load(k0 + KB);
    mfma_stage(buf, mask);
    unsigned mask_next = 0;
    if (next) mask_next = convert(buf ^ 1);
    __syncthreads();
    mask = mask_next;
    buf ^= 1;
  }
  const int col16 = lane & 15, rg = lane >> 4;
#pragma unroll
  for (int k = 0; k < TPW; ++k) {
    if (!has[k]) continue;
    float *out = partial + (long)blockIdx.x * ((long)ntiles * TILE * TILE + 32) +
                 (long)(tile0 + wave * TPW + k) * TILE * TILE;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          out[(a * 16 + rg * 4 + r) * TILE + b * 16 + col16] = acc[k][a][b][r];
  }
}

bool fock_syrk_supported(int ld) {
  const size_t lds = (size_t)8 * ld * KPAD * sizeof(unsigned short) / 2;
  return lds <= max_lds_per_block() && ld * (KB / 8) <= PASS_UNITS * 64 * PASS_WAVES;
}

void launch_fock_syrk_pass(const FockShape &sh, const float *phi32, const double *s,
                           float *partial, hipStream_t st) {
  NAPMO_REQUIRE(fock_syrk_supported(sh.ld), "fock syrk pass: basis too large for one workgroup");
  NAPMO_REQUIRE(sh.chunk % KB == 0 && sh.Npad % KB == 0, "fock syrk pass: chunking mismatch");
  const size_t lds = (size_t)4 * sh.ld * KPAD * sizeof(unsigned short);
  const int groups = (int)div_up(sh.ntiles, PASS_WAVES * TPW);
  k_fock_syrk_pass<<<dim3(sh.nchunk, groups), dim3(64 * PASS_WAVES), lds, st>>>(
      phi32, s, sh.ld, sh.nt, sh.ntiles, sh.Npad, sh.chunk, partial);
  NAPMO_HIP(hipGetLastError());
}

// P[i][j] = eta sum_k C[i][k] C[j][k] (C [n x nocc] row-major), P [n x n] row-major.
// One workgroup per 16 x 16 tile of P: the tile's 16 + 16 rows of C go through LDS in chunks
// of GRAM_K columns; each thread sums its element over k in order.
constexpr int GRAM_K = 64;

__global__ void __launch_bounds__(256) k_gram(const double *__restrict__ C, int n, int nocc,
                                              double eta, double *__restrict__ P) {
  __shared__ double ci[16][GRAM_K + 1], cj[16][GRAM_K + 1];
  const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
  const int i0 = blockIdx.y * 16, j0 = blockIdx.x * 16;
  const int i = i0 + ti, j = j0 + tj;
  double s = 0.0;
  for (int k0 = 0; k0 < nocc; k0 += GRAM_K) {
    const int kc = min(GRAM_K, nocc - k0);
    for (int e = threadIdx.x; e < 16 * GRAM_K; e += 256) {
      const int r = e / GRAM_K, k = e % GRAM_K;
      ci[r][k] = (i0 + r < n && k < kc) ? C[(long)(i0 + r) * nocc + k0 + k] : 0.0;
      cj[r][k] = (j0 + r < n && k < kc) ? C[(long)(j0 + r) * nocc + k0 + k] : 0.0;
    }
    __syncthreads();
    for (int k = 0; k < kc; ++k) s += ci[ti][k] * cj[tj][k];
    __syncthreads();
  }
  if (i < n && j < n) P[(long)i * n + j] = eta * s;
}

void launch_gram(const double *C, int n, int nocc, double eta, double *P, hipStream_t s) {
  if (n <= 0) return;
  k_gram<<<dim3((unsigned)div_up(n, 16), (unsigned)div_up(n, 16)), 256, 0, s>>>(C, n, nocc, eta, P);
  NAPMO_HIP(hipGetLastError());
}

// C [n x nocc] row-major -> zero-padded [ld x ldc]
__global__ void k_pad2d(const double *__restrict__ src, int rows, int cols, int ldd,
                        double *__restrict__ dst, int drows) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)drows * ldd) return;
  const int i = (int)(e / ldd), j = (int)(e % ldd);
  dst[e] = (i < rows && j < cols) ? src[(long)i * cols + j] : 0.0;
}

__global__ void k_pad_group(const PadMember *__restrict__ mem, int nmem) {
  const long g = blockIdx.x;
  const PadMember &m = mem[group_member(mem, nmem, g)];
  const long e = (g - m.block0) * blockDim.x + threadIdx.x;
  if (e >= (long)m.drows * m.dcols) return;
  const int i = (int)(e / m.dcols), j = (int)(e % m.dcols);
  m.dst[e] = (i < m.rows && j < m.cols) ? m.src[(long)i * m.cols + j] : 0.0;
}

void launch_pad_group(const PadMember *d_mem, int nmem, long nblocks, hipStream_t st) {
  if (nmem <= 0 || nblocks <= 0) return;
  k_pad_group<<<dim3((unsigned)nblocks), 256, 0, st>>>(d_mem, nmem);
  NAPMO_HIP(hipGetLastError());
}

void launch_pad2d(const double *src, int rows, int cols, double *dst, int drows, int dcols,
                  hipStream_t s) {
  k_pad2d<<<div_up((long)drows * dcols, 256), 256, 0, s>>>(src, rows, cols, dcols, dst, drows);
  NAPMO_HIP(hipGetLastError());
}

// ------------------------------------------------------------ calibration copy
// Plain streaming copy of n doubles (16 B per lane per iteration): the reference kernel
// whose FETCH_SIZE / WRITE_SIZE readings scale the HBM-traffic counters.
__global__ void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, long n2) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

void launch_copy(const double *src, double *dst, long n, hipStream_t s) {
  NAPMO_REQUIRE(n % 2 == 0, "copy: even length required");
  k_copy<<<4096, 256, 0, s>>>((const double2 *)src, (double2 *)dst, n / 2);
  NAPMO_HIP(hipGetLastError());
}

// -------------------------------------------------------------- layout self test
// One wave computes D = A B for A[m][k] = m + k / 64, B[k][n] = 1 + (k - n) / 128
// through the MFMA builtin and writes D with the output layout the kernels assume.
__global__ void k_selftest_mfma(int which, double *out) {
  const int lane = threadIdx.x;
  if (which == 0) {  // f64 16x16x4
    const double a = (lane % 16) + (lane / 16) / 64.0;
    const double b = 1.0 + ((lane / 16) - (lane % 16)) / 128.0;
    d4 acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    for (int r = 0; r < 4; ++r) out[((lane / 16) + 4 * r) * 16 + lane % 16] = acc[r];
  } else {  // bf16 16x16x32 (values exact in bf16)
    bf16x8 a, b;
    for (int e = 0; e < 8; ++e) {
      const int k = 8 * (lane / 16) + e;
      a[e] = (__bf16)(float)((lane % 16) + k / 64.0);
      b[e] = (__bf16)(float)(1.0 + (k - (lane % 16)) / 128.0);
    }
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    for (int r = 0; r < 4; ++r) out[(4 * (lane / 16) + r) * 16 + lane % 16] = acc[r];
  }
}

void launch_selftest_mfma(int which, double *out, hipStream_t s) {
  k_selftest_mfma<<<1, 64, 0, s>>>(which, out);
  NAPMO_HIP(hipGetLastError());
}

}  // namespace napmo
