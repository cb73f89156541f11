// Grid density on FP64 MFMA: X = Phi B over a tile of points, then one of two epilogues.
//   P form       (B = P [ld x ld]):     rho_partial[ct][p] = sum_{c in tile ct} X_pc Phi_pc
//   orbital form (B = C [ld x ldc]):    rho_partial[ct][p] = sum_{c in tile ct} X_pc^2
//   product      (B [ld x ldb]):        X [Npad x ldb] itself (the gradient's Phi P)
// (napmo/wavefunction/psi_hybrid.py:144-154 and napmo/wavefunction/psi_numeric.py:361-369).
//
// Workgroup: 128 points x 16 NB columns, 4 waves; wave w owns points 32 w .. 32 w + 31
// (two 16-row blocks) and all NB column blocks. K runs in stages of DKS (16 or 32): the next stage's
// Phi rows and B rows are loaded into registers while the MFMAs of the current stage run,
// then stored to the other LDS buffer (one barrier per stage). LDS rows are padded so each
// 16 x 4 fragment read of a wave touches every bank once per half.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace napmo {

typedef double d4 __attribute__((ext_vector_type(4)));

// Smallest e with 2^e >= m (m > 0).
__device__ inline int dens_pow2_ceil_exp(double m) {
  int ex;
  const double mant = frexp(m, &ex);
  return mant == 0.5 ? ex - 1 : ex;
}
typedef double d2 __attribute__((ext_vector_type(2)));

namespace {

constexpr int DPT = 128;        // points per workgroup
constexpr int ASTRIDE = DPT + 16;
constexpr int MODE_P = 0, MODE_ORB = 1, MODE_STORE_X = 2;

constexpr int bstride(int nb) { return (16 * nb) % 32 == 16 ? 16 * nb : 16 * nb + 16; }

// Workgroup g runs on XCD g % 8 (round-robin dispatch); the column tiles of one point
// tile are given ids that are congruent mod 8, so they share that XCD's L2 and the point
// tile's Phi rows come from HBM once: g = 8 i + x, column tile i % nct, point tile
// (i / nct) * 8 + x.
__device__ inline bool xcd_tile(int g, int nct, int npt, int &pt, int &ct) {
  const int x = g & 7, i = g >> 3;
  ct = i % nct;
  pt = (i / nct) * 8 + x;
  return pt < npt;
}

// The kernel body for workgroup g (blockIdx.x, or the local index of a grouped launch).
template <int NB, int MODE, int DKS, int PF>
__device__ __forceinline__ void density_mfma_body(const double *__restrict__ phi, int ld,
                                                  long Npad, const double *__restrict__ B,
                                                  int ldb, int kend, double scale, long nvalid,
                                                  double *__restrict__ partial, int g) {
  constexpr int BS = bstride(NB);
  __shared__ double As[2][DKS][ASTRIDE];  // As[.][k][m] = Phi[p0 + m][k0 + k]
  __shared__ double Bs[2][DKS][BS];       // Bs[.][k][c] = B[k0 + k][c0 + c]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int pt, ct;
  if (!xcd_tile(g, ldb / (16 * NB), (int)(Npad / DPT), pt, ct)) return;
  const long p0 = (long)pt * DPT;
  const int c0 = ct * 16 * NB;
  // this thread's share of a stage: Phi row (tid / 2), 8 columns from (tid % 2) * 8;
  // B elements tid + q * 256 (q < NB) of the 16 x 16 NB stage
  constexpr int HALF = DKS / 2, NA = HALF / 2, NBQ = NB * DKS / 16;
  const double *arow = phi + (p0 + (tid >> 1)) * ld + (tid & 1) * HALF;
  auto load = [&](d2 (&ra)[NA], double (&rb)[NBQ], int k0) {
#pragma unroll
    for (int i = 0; i < NA; ++i) ra[i] = *(const d2 *)(arow + k0 + 2 * i);
#pragma unroll
    for (int q = 0; q < NBQ; ++q) {
      const int e = tid + q * 256, k = e / (16 * NB), c = e % (16 * NB);
      rb[q] = B[(long)(k0 + k) * ldb + c0 + c];
    }
  };
  auto store = [&](const d2 (&ra)[NA], const double (&rb)[NBQ], int buf) {
    const int m = tid >> 1, kb = (tid & 1) * HALF;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      As[buf][kb + 2 * i][m] = ra[i].x;
      As[buf][kb + 2 * i + 1][m] = ra[i].y;
    }
#pragma unroll
    for (int q = 0; q < NBQ; ++q) {
      const int e = tid + q * 256, k = e / (16 * NB), c = e % (16 * NB);
      Bs[buf][k][c] = rb[q];
    }
  };
  d4 acc[2][NB];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = (d4){0, 0, 0, 0};
  const int kr = lane >> 4, mi = lane & 15;
  auto compute = [&](int buf) {
#pragma unroll
    for (int ks = 0; ks < DKS; ks += 4) {
      double a[2], b[NB];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[buf][ks + kr][wave * 32 + i * 16 + mi];
#pragma unroll
      for (int j = 0; j < NB; ++j) b[j] = Bs[buf][ks + kr][j * 16 + mi];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  };
  const int nstage = kend / DKS;  // K stages holding real functions (the rest are zero padding)
  d2 ra0[NA], ra1[NA];
  double rb0[NBQ], rb1[NBQ];
  if (PF == 1) {
    // one stage in flight: load st + 1 while computing st
    load(ra0, rb0, 0);
    store(ra0, rb0, 0);
    __syncthreads();
    for (int st = 0; st < nstage; ++st) {
      const int buf = st & 1;
      if (st + 1 < nstage) load(ra0, rb0, (st + 1) * DKS);
      compute(buf);
      if (st + 1 < nstage) store(ra0, rb0, buf ^ 1);
      __syncthreads();
    }
  } else {
    // two stages in flight: registers hold st + 1 (to store) and st + 2 (arriving) while
    // stage st is computed from LDS
    load(ra0, rb0, 0);
    store(ra0, rb0, 0);
    if (nstage > 1) load(ra1, rb1, DKS);
    __syncthreads();
    for (int st = 0; st < nstage; st += 2) {
      if (st + 2 < nstage) load(ra0, rb0, (st + 2) * DKS);
      compute(0);
      if (st + 1 < nstage) store(ra1, rb1, 1);
      __syncthreads();
      if (st + 1 >= nstage) break;
      if (st + 3 < nstage) load(ra1, rb1, (st + 3) * DKS);
      compute(1);
      if (st + 2 < nstage) store(ra0, rb0, 0);
      __syncthreads();
    }
  }
  // f64 16x16x4 output: lane i, element r -> row (i / 16) + 4 r, column i % 16
  if (MODE == MODE_STORE_X) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long p = p0 + wave * 32 + i * 16 + kr + 4 * r;
#pragma unroll
        for (int j = 0; j < NB; ++j) partial[p * ldb + c0 + j * 16 + mi] = acc[i][j][r];
      }
    return;
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const long p = p0 + wave * 32 + i * 16 + kr + 4 * r;
      double sum = 0.0;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const double x = acc[i][j][r];
        sum += MODE == MODE_ORB ? x * x : x * phi[p * ld + c0 + j * 16 + mi];
      }
      sum += __shfl_xor(sum, 8);
      sum += __shfl_xor(sum, 4);
      sum += __shfl_xor(sum, 2);
      sum += __shfl_xor(sum, 1);
      if (mi == 0 && p < nvalid) partial[(long)ct * Npad + p] = scale * sum;
    }
}

template <int NB, int MODE, int DKS, int PF>
__global__ void __launch_bounds__(256) k_density_mfma(const double *__restrict__ phi, int ld,
                                                      long Npad, const double *__restrict__ B,
                                                      int ldb, int kend, double scale,
                                                      long nvalid,
                                                      double *__restrict__ partial) {
  density_mfma_body<NB, MODE, DKS, PF>(phi, ld, Npad, B, ldb, kend, scale, nvalid, partial,
                                       blockIdx.x);
}

// Many plans in one launch: workgroup blockIdx.x belongs to the member whose block range
// holds it (block0 multiples of 8, so the XCD-aware tile order of each member is kept).
template <int NB, int MODE, int DKS, int PF>
__global__ void __launch_bounds__(256) k_density_mfma_group(const DensityMember *__restrict__ mem,
                                                            int nmem) {
  const long g = blockIdx.x;
  const DensityMember &d = mem[group_member(mem, nmem, g)];
  density_mfma_body<NB, MODE, DKS, PF>(d.phi, d.ld, d.Npad, d.B, d.ldb, d.kend, d.scale, d.nvalid,
                                       d.partial, (int)(g - d.block0));
}

// Orbital form with C resident in LDS: the whole zero-padded C [kend x 16 NB] is copied
// once per workgroup into fragment order, Cf[g][j][lane][s] = C[16 g + 4 (lane / 16) + s]
// [16 j + lane % 16], and every wave then streams its own Phi rows from HBM straight into
// MFMA operands with no LDS staging and no barrier: for K group g (16 functions) lane
// (kr, mi) loads Phi[row mi][16 g + 4 kr .. + 3] as one 32-byte vector, and step s of the
// group multiplies element s of it by Cf[g][j][lane][s] (the k order inside a group is
// permuted identically on both operands). A wave owns 16 RB rows at a time and walks the
// row tiles with a stride of all waves in the grid; two groups are in flight per wave.
//
// SCR (screening, sc.S > 0): phi is the plan's Phi with its columns in the order sc.perm
// (row k of C is C[sc.perm[k]]), sc.eT [Npad / 16][ld / 16] bounds every |Phi| of a 16-row,
// 16-function block by 2^eT, and each workgroup bounds sum over a group's functions of
// max_c |C| by 2^eC; the products of a 16-row tile and group with eT + eC <= -S (each of
// their terms of X below 2^-S) are neither loaded nor multiplied.
template <int NB, int RB, int WAVES, int PF, bool SCR>
__global__ void __launch_bounds__(64 * WAVES) k_density_orb_stream(
    const double *__restrict__ phi, int ld, long Npad, const double *__restrict__ C, int nrow,
    int ncol, int kend, double scale, long nvalid, long ntiles, double *__restrict__ out,
    DensityScreen sc) {
  extern __shared__ d4 Cf[];  // [G][NB][64] vectors of 4 steps (then, SCR: eC [G] int)
  const int G = kend / 16;
  // C [nrow x ncol] row-major as given; entries past it are the zero padding
  for (int e = threadIdx.x; e < G * NB * 256; e += 64 * WAVES) {
    const int st = e & 3, ln = (e >> 2) & 63, gj = e >> 8;
    const int j = gj % NB, g = gj / NB;
    const int k = 16 * g + 4 * (ln >> 4) + st, c = 16 * j + (ln & 15);
    const int kk = SCR && k < nrow ? sc.perm[k] : k;
    ((double *)Cf)[e] = (k < nrow && c < ncol) ? C[(long)kk * ncol + c] : 0.0;
  }
  int *eC = (int *)(Cf + G * NB * 64);
  if constexpr (SCR) {
    // eC[g]: exponent bounding sum over the group's 16 functions of max_c |C| (16 lanes each)
    for (int e = threadIdx.x; e < G * 16; e += 64 * WAVES) {
      const int k = e;  // 16 g + t
      double m = 0.0;
      if (k < nrow) {
        const double *row = C + (long)sc.perm[k] * ncol;
        for (int c = 0; c < ncol; ++c) m = fmax(m, fabs(row[c]));
      }
      m += __shfl_xor(m, 8);
      m += __shfl_xor(m, 4);
      m += __shfl_xor(m, 2);
      m += __shfl_xor(m, 1);
      if ((e & 15) == 0) eC[e >> 4] = m > 0.0 ? dens_pow2_ceil_exp(m) : -4000;
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, kr = lane >> 4, mi = lane & 15;
  const long nw = (long)gridDim.x * WAVES;
  const long ntile16 = Npad / 16;
  const int ldg = ld / 16;
  // SCR: the block exponents of a wave's next tile are loaded one tile ahead
  auto tile_exp = [&](long t, int rb) -> int {
    const long t16 = min(t * RB + rb, ntile16 - 1);
    return lane < G ? (int)sc.eT[t16 * ldg + lane] : 0;
  };
  int eTn[RB];
  if constexpr (SCR) {
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) eTn[rb] = tile_exp((long)blockIdx.x * WAVES + wave, rb);
  }
  for (long t = (long)blockIdx.x * WAVES + wave; t < ntiles; t += nw) {
    const long r0 = t * 16 * RB;
    const double *ap[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      const long row = min(r0 + rb * 16 + mi, Npad - 1);  // rows past Npad: read, never written
      ap[rb] = phi + row * ld + 4 * kr;
    }
    // SCR: bit g of live[rb] = the (16-row tile, group g) products are kept
    unsigned long long live[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      live[rb] = ~0ull;
      if constexpr (SCR) {
        const bool keep = lane < G && eTn[rb] + eC[lane] > -sc.S;
        live[rb] = __ballot(keep);
        eTn[rb] = tile_exp(t + nw, rb);
      }
    }
    d4 acc[RB][NB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[rb][j] = (d4){0, 0, 0, 0};
    auto load = [&](d4 (&a)[RB], int g) {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
        if (!SCR || ((live[rb] >> g) & 1)) a[rb] = *(const d4 *)(ap[rb] + 16 * g);
    };
    auto compute = [&](const d4 (&a)[RB], int g) {
      d4 b[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) b[j] = Cf[(g * NB + j) * 64 + lane];
      if constexpr (!SCR) {
#pragma unroll
        for (int st = 0; st < 4; ++st)
#pragma unroll
          for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int j = 0; j < NB; ++j)
              acc[rb][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb][st], b[j][st], acc[rb][j], 0, 0, 0);
      } else {
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
          if ((live[rb] >> g) & 1) {
#pragma unroll
            for (int st = 0; st < 4; ++st)
#pragma unroll
              for (int j = 0; j < NB; ++j)
                acc[rb][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb][st], b[j][st], acc[rb][j], 0, 0, 0);
          }
      }
    };
    d4 a[PF][RB];
    if constexpr (!SCR) {
      // a ring of PF groups in flight: slot i holds group g0 + i
#pragma unroll
      for (int i = 0; i < PF; ++i)
        if (i < G) load(a[i], i);
      for (int g0 = 0; g0 < G; g0 += PF) {
#pragma unroll
        for (int i = 0; i < PF; ++i) {
          const int g = g0 + i;
          if (g < G) {
            compute(a[i], g);
            if (g + PF < G) load(a[i], g + PF);
          }
        }
      }
    } else {
      // the ring walks only the groups some 16-row tile of this wave keeps, in increasing
      // order (each accumulator adds its kept groups in the same order as the full ring)
      unsigned long long rest = 0;
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) rest |= live[rb];
      if (G < 64) rest &= (1ull << G) - 1;
      int gl[PF];
#pragma unroll
      for (int i = 0; i < PF; ++i) {
        gl[i] = rest ? (int)__builtin_ctzll(rest) : -1;
        rest &= rest - 1;
        if (gl[i] >= 0) load(a[i], gl[i]);
      }
      while (gl[0] >= 0) {
#pragma unroll
        for (int i = 0; i < PF; ++i) {
          if (gl[i] < 0) break;
          compute(a[i], gl[i]);
          gl[i] = rest ? (int)__builtin_ctzll(rest) : -1;
          rest &= rest - 1;
          if (gl[i] >= 0) load(a[i], gl[i]);
        }
      }
    }
    // f64 16x16x4 output: lane i, element r -> row (i / 16) + 4 r, column i % 16
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long p = r0 + rb * 16 + kr + 4 * r;
        double sum = 0.0;
#pragma unroll
        for (int j = 0; j < NB; ++j) sum += acc[rb][j][r] * acc[rb][j][r];
        sum += __shfl_xor(sum, 8);
        sum += __shfl_xor(sum, 4);
        sum += __shfl_xor(sum, 2);
        sum += __shfl_xor(sum, 1);
        if (mi == 0 && p < nvalid) out[p] = scale * sum;
      }
  }
}

template <int NB, int RB, int WAVES, int PF, bool SCR>
void launch_orb_stream_scr(const double *phi, int ld, long Npad, const double *C, int nrow, int ncol,
                           int kend, double scale, long nvalid, double *out, const DensityScreen &sc,
                           hipStream_t s) {
  constexpr int STREAM_RB = RB, STREAM_WAVES = WAVES;
  const size_t bytes = (size_t)kend * 16 * NB * sizeof(double) + (SCR ? 4 * (size_t)round_up(kend / 16, 4) : 0);
  const long ntiles = div_up(Npad, 16L * STREAM_RB);
  auto kern = k_density_orb_stream<NB, RB, WAVES, PF, SCR>;
  static thread_local int configured = -1;  // per instantiation: dynamic LDS above 64 KB
  int dev = 0;
  NAPMO_HIP(hipGetDevice(&dev));
  if (configured != dev) {
    NAPMO_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)max_lds_per_block()));
    configured = dev;
  }
  int occ = 0, cus = 0;
  NAPMO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 64 * STREAM_WAVES, bytes));
  NAPMO_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  NAPMO_REQUIRE(occ >= 1, "density_orb_stream: no occupancy");
  const long blocks = std::min<long>((long)occ * cus, div_up(ntiles, STREAM_WAVES));
  kern<<<dim3((unsigned)blocks), dim3(64 * STREAM_WAVES), bytes, s>>>(phi, ld, Npad, C, nrow, ncol,
                                                                       kend, scale, nvalid, ntiles, out, sc);
  NAPMO_HIP(hipGetLastError());
}

template <int NB, int RB, int WAVES, int PF>
void launch_orb_stream_nb(const double *phi, int ld, long Npad, const double *C, int nrow, int ncol,
                          int kend, double scale, long nvalid, double *out, const DensityScreen &sc,
                          hipStream_t s) {
  if (sc.S > 0)
    launch_orb_stream_scr<NB, RB, WAVES, PF, true>(phi, ld, Npad, C, nrow, ncol, kend, scale, nvalid, out, sc, s);
  else
    launch_orb_stream_scr<NB, RB, WAVES, PF, false>(phi, ld, Npad, C, nrow, ncol, kend, scale, nvalid, out, sc, s);
}

template <int MODE, int DKS, int PF>
void launch_density_ks(int nb, dim3 tiles, const double *phi, int ld, long Npad, const double *B,
                       int ldb, int nreal, double scale, long nvalid, double *partial,
                       hipStream_t s) {
  // tiles = (point tiles, column tiles) -> one 1-D grid in the XCD-aware order
  const dim3 grid((unsigned)(round_up(tiles.x, 8) * tiles.y));
  // Phi columns and B rows from nreal on are zero: the K loop stops at the first stage
  // boundary past them (the skipped products are exact zeros)
  const int kend = (nreal > 0 && nreal < ld) ? (int)std::min<long>(ld, round_up(nreal, DKS)) : ld;
  switch (nb) {
    case 1: k_density_mfma<1, MODE, DKS, PF><<<grid, 256, 0, s>>>(phi, ld, Npad, B, ldb, kend, scale, nvalid, partial); break;
    case 2: k_density_mfma<2, MODE, DKS, PF><<<grid, 256, 0, s>>>(phi, ld, Npad, B, ldb, kend, scale, nvalid, partial); break;
    case 3: k_density_mfma<3, MODE, DKS, PF><<<grid, 256, 0, s>>>(phi, ld, Npad, B, ldb, kend, scale, nvalid, partial); break;
    default: k_density_mfma<4, MODE, DKS, PF><<<grid, 256, 0, s>>>(phi, ld, Npad, B, ldb, kend, scale, nvalid, partial); break;
  }
  NAPMO_HIP(hipGetLastError());
}

// ks: K per stage, 16 or 32 (32 needs ld % 32 == 0); pf: stages in flight, 1 or 2
template <int MODE>
void launch_density_nb(int nb, dim3 grid, const double *phi, int ld, long Npad, const double *B,
                       int ldb, double *partial, hipStream_t s, int ks = 16, int pf = 1,
                       int nreal = 0, double scale = 1.0, long nvalid = -1) {
  if (nvalid < 0) nvalid = Npad;
  if (ks == 32 && ld % 32 == 0) {
    if (pf == 2)
      launch_density_ks<MODE, 32, 2>(nb, grid, phi, ld, Npad, B, ldb, nreal, scale, nvalid, partial, s);
    else
      launch_density_ks<MODE, 32, 1>(nb, grid, phi, ld, Npad, B, ldb, nreal, scale, nvalid, partial, s);
  } else {
    if (pf == 2)
      launch_density_ks<MODE, 16, 2>(nb, grid, phi, ld, Npad, B, ldb, nreal, scale, nvalid, partial, s);
    else
      launch_density_ks<MODE, 16, 1>(nb, grid, phi, ld, Npad, B, ldb, nreal, scale, nvalid, partial, s);
  }
}

}  // namespace

long density_group_fill(DensityMember &d, const double *phi, int ld, long Npad, const double *P,
                        double *partial, int nreal, int ks, long block0) {
  NAPMO_REQUIRE(ld % TILE == 0 && Npad % DPT == 0 && block0 % 8 == 0, "grouped density: padding");
  const int dks = (ks == 32 && ld % 32 == 0) ? 32 : 16;
  d.phi = phi, d.B = P, d.partial = partial, d.Npad = Npad, d.nvalid = Npad, d.block0 = block0;
  d.ld = ld, d.ldb = ld, d.scale = 1.0;
  d.kend = (nreal > 0 && nreal < ld) ? (int)std::min<long>(ld, round_up(nreal, dks)) : ld;
  return round_up(Npad / DPT, 8) * (ld / TILE);
}

void launch_density_group(const DensityMember *d_mem, int nmem, long nblocks, int ks, int pf,
                          hipStream_t s) {
  if (nmem <= 0 || nblocks <= 0) return;
  const dim3 grid((unsigned)nblocks);
  if (ks == 32) {
    if (pf == 2)
      k_density_mfma_group<4, MODE_P, 32, 2><<<grid, 256, 0, s>>>(d_mem, nmem);
    else
      k_density_mfma_group<4, MODE_P, 32, 1><<<grid, 256, 0, s>>>(d_mem, nmem);
  } else {
    if (pf == 2)
      k_density_mfma_group<4, MODE_P, 16, 2><<<grid, 256, 0, s>>>(d_mem, nmem);
    else
      k_density_mfma_group<4, MODE_P, 16, 1><<<grid, 256, 0, s>>>(d_mem, nmem);
  }
  NAPMO_HIP(hipGetLastError());
}

void launch_density_f64(const double *phi, int ld, long Npad, const double *P, double *partial,
                        hipStream_t s, int ks, int pf, int nreal) {
  NAPMO_REQUIRE(ld % TILE == 0 && Npad % DPT == 0, "density: padding mismatch");
  launch_density_nb<MODE_P>(4, dim3(Npad / DPT, ld / TILE), phi, ld, Npad, P, ld, partial, s, ks,
                            pf, nreal);
}

int density_orb_cols(int nocc) {
  const int c = (int)round_up(nocc, 16);
  return c <= 64 ? c : (int)round_up(nocc, 64);
}

int density_orb_tiles(int ldc) { return ldc <= 64 ? 1 : ldc / 64; }

void launch_density_orb_f64(const double *phi, int ld, long Npad, const double *C, int ldc,
                            double *partial, hipStream_t s, int ks, int pf, int nreal, double scale,
                            long nvalid) {
  NAPMO_REQUIRE(Npad % DPT == 0 && ld % 16 == 0 && ldc % 16 == 0, "density_orb: padding");
  const int nb = ldc <= 64 ? ldc / 16 : 4;
  NAPMO_REQUIRE(ldc <= 64 || ldc % 64 == 0, "density_orb: orbital padding");
  NAPMO_REQUIRE(nvalid < 0 || density_orb_tiles(ldc) == 1,
                "density_orb: a direct (scaled) write needs one column tile");
  launch_density_nb<MODE_ORB>(nb, dim3(Npad / DPT, density_orb_tiles(ldc)), phi, ld, Npad, C, ldc,
                              partial, s, ks, pf, nreal, scale, nvalid);
}

bool density_orb_stream_supported(int ld, int ldc, int nreal) {
  if (ldc > 64 || ldc % 16 || ld % 16) return false;
  const int kend = (nreal > 0 && nreal < ld) ? (int)round_up(nreal, 16) : ld;
  return (size_t)kend * ldc * sizeof(double) <= max_lds_per_block();
}

template <int RB, int WAVES, int PF>
static void orb_stream_shape(int nb, const double *phi, int ld, long Npad, const double *C, int nrow,
                             int ncol, int kend, double scale, long nvalid, double *out,
                             const DensityScreen &sc, hipStream_t s) {
  switch (nb) {
    case 1: launch_orb_stream_nb<1, RB, WAVES, PF>(phi, ld, Npad, C, nrow, ncol, kend, scale, nvalid, out, sc, s); break;
    case 2: launch_orb_stream_nb<2, RB, WAVES, PF>(phi, ld, Npad, C, nrow, ncol, kend, scale, nvalid, out, sc, s); break;
    case 3: launch_orb_stream_nb<3, RB, WAVES, PF>(phi, ld, Npad, C, nrow, ncol, kend, scale, nvalid, out, sc, s); break;
    default: launch_orb_stream_nb<4, RB, WAVES, PF>(phi, ld, Npad, C, nrow, ncol, kend, scale, nvalid, out, sc, s); break;
  }
}

void launch_density_orb_stream(const double *phi, int ld, long Npad, const double *C, int nrow,
                               int ncol, double *out, hipStream_t s, double scale, long nvalid,
                               int shape, const DensityScreen &sc) {
  const int ldc = density_orb_cols(ncol);
  NAPMO_REQUIRE(nrow >= 1 && nrow <= ld && ncol >= 1, "density_orb_stream: bad C shape");
  NAPMO_REQUIRE(density_orb_stream_supported(ld, ldc, nrow) && Npad % 16 == 0,
                "density_orb_stream: C does not fit in LDS or bad padding");
  NAPMO_REQUIRE(sc.S <= 0 || (sc.perm && sc.eT && ld <= 1024),
                "density_orb_stream: screening needs the column order and block exponents");
  const int kend = (int)std::min<long>(ld, round_up(nrow, 16));
  if (nvalid < 0) nvalid = Npad;
  const int nb = ldc / 16;
  // shapes: rows per wave (16 RB), waves per workgroup, K groups in flight
  switch (shape) {
    case 2: orb_stream_shape<3, 8, 2>(nb, phi, ld, Npad, C, nrow, ncol, kend, scale, nvalid, out, sc, s); break;
    case 3: orb_stream_shape<3, 8, 3>(nb, phi, ld, Npad, C, nrow, ncol, kend, scale, nvalid, out, sc, s); break;
    default: orb_stream_shape<2, 12, 3>(nb, phi, ld, Npad, C, nrow, ncol, kend, scale, nvalid, out, sc, s); break;
  }
}

__global__ void k_phi_tile_exp(const double *__restrict__ phi, int ld, long ntile16,
                               signed char *__restrict__ out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ntile16 * ld) return;
  const long t = e / ld;
  const int f = (int)(e % ld);
  double m = 0.0;
  for (int r = 0; r < 16; ++r) m = fmax(m, fabs(phi[(16 * t + r) * ld + f]));
  out[e] = m > 0.0 ? (signed char)max(-127, min(127, dens_pow2_ceil_exp(m))) : (signed char)-128;
}

__global__ void k_permute_cols(const double *__restrict__ phi, int ld, long Npad, const int *__restrict__ perm,
                               double *__restrict__ out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Npad * ld) return;
  const long p = e / ld;
  const int k = (int)(e % ld);
  out[e] = phi[p * ld + perm[k]];
}

void launch_phi_tile_exp(const double *phi, int ld, long Npad, signed char *out, hipStream_t s) {
  NAPMO_REQUIRE(Npad % 16 == 0, "phi tile exponents: rows not a multiple of 16");
  const long ne = (Npad / 16) * ld;
  k_phi_tile_exp<<<div_up(ne, 256), 256, 0, s>>>(phi, ld, Npad / 16, out);
  NAPMO_HIP(hipGetLastError());
}

void launch_permute_cols(const double *phi, int ld, long Npad, const int *perm, double *out, hipStream_t s) {
  const long ne = Npad * ld;
  k_permute_cols<<<div_up(ne, 256), 256, 0, s>>>(phi, ld, Npad, perm, out);
  NAPMO_HIP(hipGetLastError());
}

void launch_phi_times(const double *phi, int ld, long Npad, const double *B, int ldb, double *X,
                      hipStream_t s) {
  NAPMO_REQUIRE(ld % TILE == 0 && ldb % TILE == 0 && Npad % DPT == 0, "phi_times: padding");
  launch_density_nb<MODE_STORE_X>(4, dim3(Npad / DPT, ldb / TILE), phi, ld, Npad, B, ldb, X, s);
}

}  // namespace napmo
