// gfx950 kernels of the nAPMO grid hot path: Becke partition, GTO collocation,
// Poisson projection / radial Green solve / synthesis, and the FP64 MFMA grid
// contractions rho = diag(Phi P Phi^T) and F = Phi^T diag(s) Phi.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "kernels.hpp"
#include "synth_kernel.hpp"

namespace napmo {

typedef double d4 __attribute__((ext_vector_type(4)));

static constexpr double INV_SQRT_4_PI = 0.28209479177387814347;

// ================================================================ Becke partition
// One thread per point. Pair tables are wave-uniform loads (scalar cache).
__global__ void k_becke(const double *__restrict__ x, const double *__restrict__ y,
                        const double *__restrict__ z, long N, const int *__restrict__ owner, int M,
                        const double *__restrict__ center, const double *__restrict__ R,
                        const double *__restrict__ As, double *__restrict__ bw) {
  // Same operation order as the reference loop (no FMA contraction): the weights pass
  // through 1 - s with s near 1, where contraction changes the last bits visibly.
#pragma clang fp contract(off)
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const double px = x[p], py = y[p], pz = z[p];
  const int own = owner[p];
  double sum = 0.0, mine = 0.0;
  for (int i = 0; i < M; ++i) {
    double dx = px - center[3 * i], dy = py - center[3 * i + 1], dz = pz - center[3 * i + 2];
    const double ri = sqrt(dx * dx + dy * dy + dz * dz);
    double prod = 1.0;
    for (int j = 0; j < M; ++j) {
      if (j == i) continue;
      double ex = px - center[3 * j], ey = py - center[3 * j + 1], ez = pz - center[3 * j + 2];
      const double rj = sqrt(ex * ex + ey * ey + ez * ez);
      const double mu = (ri - rj) / R[i * M + j];
      double sv = mu + As[i * M + j] * (1.0 - mu * mu);
      sv = 0.5 * sv * (3 - sv * sv);
      sv = 0.5 * sv * (3 - sv * sv);
      sv = 0.5 * sv * (3 - sv * sv);
      prod *= 0.5 * (1.0 - sv);
    }
    if (i == own) mine = prod;
    sum += prod;
  }
  bw[p] = mine / sum;
}

void launch_becke(const double *x, const double *y, const double *z, long N, const int *owner,
                  int M, const double *center, const double *R, const double *As, double *bw,
                  hipStream_t s) {
  if (N <= 0) return;
  k_becke<<<div_up(N, 256), 256, 0, s>>>(x, y, z, N, owner, M, center, R, As, bw);
  NAPMO_HIP(hipGetLastError());
}

// ==================================================================== collocation
// Thread per (point, function): lanes run over functions so row writes coalesce.
__device__ inline double ipow(double v, int e) {
  double r = 1.0;
  for (int k = 0; k < e; ++k) r *= v;
  return r;
}

__global__ void k_collocate(const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ z, long N, long Npad, int n, int ld,
                            const double *__restrict__ fo, const int *__restrict__ fl,
                            const int *__restrict__ ff, const int *__restrict__ fnp,
                            const double *__restrict__ fnorm, const double *__restrict__ pz,
                            const double *__restrict__ pc, double *__restrict__ out64,
                            float *__restrict__ out32) {
  const int mu = blockIdx.x * blockDim.x + threadIdx.x;
  const long p = (long)blockIdx.y * blockDim.y + threadIdx.y;
  if (mu >= ld || p >= Npad) return;
  double val = 0.0;
  if (mu < n && p < N) {
    const double dx = x[p] - fo[3 * mu], dy = y[p] - fo[3 * mu + 1], dz = z[p] - fo[3 * mu + 2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double ang = ipow(dx, fl[3 * mu]) * ipow(dy, fl[3 * mu + 1]) * ipow(dz, fl[3 * mu + 2]);
    double acc = 0.0;
    const int f0 = ff[mu], np = fnp[mu];
    for (int k = 0; k < np; ++k) acc += pc[f0 + k] * ang * exp(-pz[f0 + k] * r2);
    val = acc * fnorm[mu];
  }
  const size_t o = (size_t)p * ld + mu;
  if (out64) out64[o] = val;
  if (out32) out32[o] = (float)val;
}

void launch_collocate(const double *x, const double *y, const double *z, long N, long Npad, int n,
                      int ld, const double *fo, const int *fl, const int *ff, const int *fnp,
                      const double *fnorm, const double *pz, const double *pc, double *out64,
                      float *out32, hipStream_t s) {
  dim3 block(64, 4);
  dim3 grid(div_up(ld, 64), div_up(Npad, 4));
  k_collocate<<<grid, block, 0, s>>>(x, y, z, N, Npad, n, ld, fo, fl, ff, fnp, fnorm, pz, pc, out64,
                                     out32);
  NAPMO_HIP(hipGetLastError());
}

// ========================================================= Poisson: projection
// rho_lm[A][i][lm] = sum_j f_Aij Yhat[j][lm] with f = dens * bw, as a small GEMM per
// atom: one workgroup per (atom, block of PSB radial shells) so the Lebedev table
// Yhat[j][lm] = aw_j S_lm(u_j) sqrt(4 pi (2l+1)) / sum aw is read once per block; the
// shells' f values are staged in LDS in chunks of PJ angular points.
constexpr int PSB = 5, PJ = 256;

// Threads: 64 lm x PSL angular slices; slice k sums angular points j = k, k + PSL, ... (two
// independent loads in flight per step); slices are combined in order through LDS.
template <int PSL>
__global__ void __launch_bounds__(64 * PSL) k_project(AtomTables at, const double *__restrict__ dens,
                                                      const double *__restrict__ bw,
                                                      const double *__restrict__ yhat, int L,
                                                      double *__restrict__ rho_lm, long dstride,
                                                      long ostride) {
  __shared__ double sf[PSB][PJ];
  __shared__ double red[PSL][PSB][64];
  // blockIdx.z = o M + A: density o (dens + o dstride -> rho_lm + o ostride), atom A
  const int A = blockIdx.z % at.M, o = blockIdx.z / at.M;
  dens += o * dstride;
  rho_lm += o * ostride;
  const int i0 = blockIdx.x * PSB;
  const int nr = at.nrad[A];
  if (i0 >= nr) return;
  const int ns = min(PSB, nr - i0);
  const int na = at.nang[A];
  const long p0 = at.point_off[A] + (long)i0 * na;
  const double *Y = yhat + at.yhat_off[A];
  const int lml = threadIdx.x & 63, slice = threadIdx.x >> 6;
  const int lm = blockIdx.y * 64 + lml;
  // harmonics above the centre's cut are zero (written, not summed)
  const int Lc = at.lcut ? (at.lcut[A] + 1) * (at.lcut[A] + 1) : L;
  double acc[PSB];
#pragma unroll
  for (int k = 0; k < PSB; ++k) acc[k] = 0.0;
  for (int j0 = 0; j0 < na; j0 += PJ) {
    const int nj = min(PJ, na - j0);
    __syncthreads();
    for (int e = threadIdx.x; e < PSB * PJ; e += blockDim.x) {
      const int k = e / PJ, j = e % PJ;
      const long p = p0 + (long)k * na + j0 + j;
      sf[k][j] = (k < ns && j < nj) ? dens[p] * bw[p] : 0.0;
    }
    __syncthreads();
    if (lm < Lc) {
      // pairs (j, j + PSL) in order; the Yhat loads of the next two pairs are in flight
      // while one pair is summed
      int j = slice;
      const double *Yc = Y + (long)j0 * L + lm;
      double ya0 = 0, ya1 = 0, yb0 = 0, yb1 = 0;
      if (j + PSL < nj) ya0 = Yc[(long)j * L], ya1 = Yc[(long)(j + PSL) * L];
      if (j + 3 * PSL < nj) yb0 = Yc[(long)(j + 2 * PSL) * L], yb1 = Yc[(long)(j + 3 * PSL) * L];
      for (; j + PSL < nj; j += 2 * PSL) {
        const double y0 = ya0, y1 = ya1;
        ya0 = yb0, ya1 = yb1;
        if (j + 5 * PSL < nj) yb0 = Yc[(long)(j + 4 * PSL) * L], yb1 = Yc[(long)(j + 5 * PSL) * L];
#pragma unroll
        for (int k = 0; k < PSB; ++k) acc[k] += sf[k][j] * y0 + sf[k][j + PSL] * y1;
      }
      for (; j < nj; j += PSL) {
        const double y = Y[(long)(j0 + j) * L + lm];
#pragma unroll
        for (int k = 0; k < PSB; ++k) acc[k] += sf[k][j] * y;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < PSB; ++k) red[slice][k][lml] = acc[k];
  __syncthreads();
  if (slice == 0 && lm < L)
    for (int k = 0; k < ns; ++k) {
      double t = 0.0;
      for (int sl = 0; sl < PSL; ++sl) t += red[sl][k][lml];
      rho_lm[(at.rad_off[A] + i0 + k) * (long)L + lm] = lm < Lc ? t : 0.0;
    }
}

void launch_project(const AtomTables &at, const double *dens, const double *bw, const double *yhat,
                    int lmax, double *rho_lm, hipStream_t s, int nset, long dstride, long ostride) {
  const int L = (lmax + 1) * (lmax + 1);
  NAPMO_REQUIRE(nset >= 1 && (long)nset * at.M <= 65535, "projection: too many density sets");
  dim3 grid(div_up(at.max_nrad, PSB), div_up(L, 64), nset * at.M);
  // one slice count for every batch size, so a density's projections do not depend on
  // how many densities share the launch
  k_project<16><<<grid, 64 * 16, 0, s>>>(at, dens, bw, yhat, L, rho_lm, dstride, ostride);
  NAPMO_HIP(hipGetLastError());
}

// ====================================================== Poisson: radial solves
// U[A][i][lm] = sum_j G_{A,l}[i][j] rho_lm[A][j][lm] for y and dt; one workgroup per
// (atom, l, chunk of 256 outputs): the (nr x 2l+1) block of rho_lm is staged in LDS and
// each thread forms one output from a contiguous Green row. green block of an atom: for
// l = 0..lmax: Gy [nr x nr], then Gd [nr x nr].
constexpr int RMAX_STAGE = 4096;  // doubles of rho_lm staged per (atom, l)
constexpr int RADIAL_CHUNK = 256;

__global__ void __launch_bounds__(RADIAL_CHUNK) k_radial(AtomTables at, const double *__restrict__ rho_lm,
                                                const double *__restrict__ green, int L,
                                                double *__restrict__ Uy, double *__restrict__ Udt,
                                                int nchunk, long stride) {
  __shared__ double sr[RMAX_STAGE];
  const int A = blockIdx.x, l = blockIdx.y;
  if (at.lcut && l > at.lcut[A]) return;  // above the centre's cut: never read
  // blockIdx.z = set * nchunk + chunk: set k reads rho_lm + k stride, writes U + k stride
  const long set = blockIdx.z / nchunk;
  rho_lm += set * stride, Uy += set * stride, Udt += set * stride;
  const int nr = at.nrad[A], nm = 2 * l + 1;
  const int total = 2 * nr * nm;
  const int o0 = (blockIdx.z % nchunk) * RADIAL_CHUNK;
  if (o0 >= total) return;
  const double *G = green + at.green_off[A] + (long)l * 2 * nr * nr;
  const long roff = at.rad_off[A] * (long)L;
  const bool staged = nr * nm <= RMAX_STAGE;
  if (staged)
    for (int e = threadIdx.x; e < nr * nm; e += blockDim.x)
      sr[e] = rho_lm[roff + (long)(e / nm) * L + l * l + e % nm];
  __syncthreads();
  const int o = o0 + threadIdx.x;
  if (o >= total) return;
  const int m = o % nm, row = o / nm;  // row in [0, 2 nr): which * nr + i
  const int which = row / nr, i = row - which * nr;
  const double *g = G + (long)which * nr * nr + (long)i * nr;
  double acc = 0.0;
  if (staged) {
    for (int j = 0; j < nr; ++j) acc += g[j] * sr[j * nm + m];
  } else {
    const double *r = rho_lm + roff + l * l + m;
    for (int j = 0; j < nr; ++j) acc += g[j] * r[(long)j * L];
  }
  (which ? Udt : Uy)[roff + (long)i * L + l * l + m] = acc;
}

void launch_radial(const AtomTables &at, const double *rho_lm, const double *green, int lmax,
                   double *Uy, double *Udt, hipStream_t s, int nset, long stride) {
  const int L = (lmax + 1) * (lmax + 1);
  const int nchunk = (int)div_up(2L * at.max_nrad * (2 * lmax + 1), RADIAL_CHUNK);
  NAPMO_REQUIRE(nset >= 1 && (long)nset * nchunk <= 65535, "radial solves: too many sets");
  dim3 grid(at.M, lmax + 1, (unsigned)(nset * nchunk));
  k_radial<<<grid, RADIAL_CHUNK, 0, s>>>(at, rho_lm, green, L, Uy, Udt, nchunk, stride);
  NAPMO_HIP(hipGetLastError());
}

// ====================================================== Poisson: synthesis tables
// Synthesis order q: for m = 0..lmax, for l = m..lmax: C_lm (and S_lm when m > 0).
// Reference order index of (l, m, cs): l*l + (m ? 2m - 1 + cs : 0).
// Each coefficient carries sqrt(2l+1)/sqrt(4 pi) (evaluate.cpp:195-205) and the Racah
// factor sqrt(2 (l-m)!/(l+m)!) of moments.cpp:106-112, so that the kernel multiplies
// by (x + i y)^m times the PI_l^m(z) polynomial of the unit vector only.
__device__ inline void q_decode(int q, int lmax, int &l, int &m, int &cs) {
  int base = 0;
  for (m = 0; m <= lmax; ++m) {
    const int cnt = (lmax - m + 1) * (m ? 2 : 1);
    if (q < base + cnt) {
      const int r = q - base;
      if (m == 0) {
        l = r, cs = 0;
      } else {
        l = m + r / 2, cs = r % 2;
      }
      return;
    }
    base += cnt;
  }
  l = m = cs = 0;
}

__device__ inline double q_factor(int l, int m) {
  double f = sqrt(2.0 * l + 1.0) * INV_SQRT_4_PI;
  if (m > 0) {
    double r = M_SQRT2;
    for (int k = 1; k <= m; ++k) r /= sqrt((double)((l + k) * (l - k + 1)));
    f *= r;
  }
  return f;
}

void synth_order(int lmax, std::vector<int> &l, std::vector<int> &m, std::vector<int> &cs,
                 std::vector<double> &scale) {
  l.clear(), m.clear(), cs.clear(), scale.clear();
  for (int mm = 0; mm <= lmax; ++mm)
    for (int ll = mm; ll <= lmax; ++ll)
      for (int c = 0; c < (mm ? 2 : 1); ++c) {
        double f = std::sqrt(2.0 * ll + 1.0) * INV_SQRT_4_PI;
        if (mm > 0) {
          double r = M_SQRT2;
          for (int k = 1; k <= mm; ++k) r /= std::sqrt((double)((ll + k) * (ll - k + 1)));
          f *= r;
        }
        f *= pi_sigma(ll, mm);
        l.push_back(ll), m.push_back(mm), cs.push_back(c), scale.push_back(f);
      }
}

constexpr int SYNTH_TAB_J = 4;  // intervals per workgroup (the factors are formed once for them)

// Layout of the tables: coefficient c of (set, centre A, interval j, q) at
// T[set tset + (A (max_nrad - 1) + j) trow + q tq + c], amplitudes at
// amp[set aset + A arow + q aq]. Sets nvalid .. gridDim.z - 1 are written as zeros (the
// padding columns of the density-interleaved layout).
struct SynthTabLayout {
  long tset, trow, tq, aset, arow, aq;
  int nvalid;
};

__global__ void k_synth_tables(AtomTables at, const double *__restrict__ Uy,
                               const double *__restrict__ Udt, int lmax, int L,
                               int write_amps, double *__restrict__ T,
                               double *__restrict__ ampL, double *__restrict__ ampR,
                               long ustride, SynthTabLayout lay) {
  const int A = blockIdx.y, jb = blockIdx.x * SYNTH_TAB_J;
  // blockIdx.z = set: U + set ustride -> its tables
  const long set = blockIdx.z;
  const bool pad = set >= lay.nvalid;
  Uy += set * ustride, Udt += set * ustride, T += set * lay.tset;
  ampL += set * lay.aset, ampR += set * lay.aset;
  const int nr = at.nrad[A];
  if (jb >= nr - 1) return;
  const long roff = at.rad_off[A];
  // a centre with its own cut: its harmonics only, in its own synthesis order, in its own
  // rows (at.toff)
  const int la = at.lcut ? at.lcut[A] : lmax, La = (la + 1) * (la + 1);
  for (int q = threadIdx.x; q < La; q += blockDim.x) {
    int l, m, cs;
    q_decode(q, la, l, m, cs);
    const int ref = l * l + (m ? 2 * m - 1 + cs : 0);
    const double f = q_factor(l, m) * pi_sigma(l, m);
    for (int j = jb; j < min(jb + SYNTH_TAB_J, nr - 1); ++j) {
      double *t = at.toff ? T + at.toff[A] + ((long)j * La + q) * 4
                          : T + ((long)A * (at.max_nrad - 1) + j) * lay.trow + q * lay.tq;
      if (pad) {
        t[0] = t[1] = t[2] = t[3] = 0.0;
        if (j == 0 && write_amps) ampL[A * lay.arow + q * lay.aq] = ampR[A * lay.arow + q * lay.aq] = 0.0;
        continue;
      }
      const double y0 = Uy[(roff + j) * L + ref], y1 = Uy[(roff + j + 1) * L + ref];
      const double d0 = Udt[(roff + j) * L + ref], d1 = Udt[(roff + j + 1) * L + ref];
      const double zz = y1 - y0;
      t[0] = y0 * f;
      t[1] = d0 * f;
      t[2] = (3 * zz - 2 * d0 - d1) * f;
      t[3] = (-2 * zz + d0 + d1) * f;
      if (j == 0 && write_amps) {
        // PotentialExtrapolation::prepare (src/horton/extrapolation.cpp:156-159)
        const double r0 = at.r[roff], rn = at.r[roff + nr - 1];
        const double yn = Uy[(roff + nr - 1) * L + ref];
        ampL[A * lay.arow + q * lay.aq] = y0 / pow(r0, (double)l) * f;
        ampR[A * lay.arow + q * lay.aq] = yn * pow(rn, (double)(l + 1)) * f;
      }
    }
  }
}

long synth_table_size(const AtomTables &at, int lmax) {
  return (long)at.M * std::max(at.max_nrad - 1, 1) * (lmax + 1) * (lmax + 1) * 4;
}

void launch_synth_tables(const AtomTables &at, const double *Uy, const double *Udt, int lmax,
                         double *T, double *ampL, double *ampR, hipStream_t s, int write_amps,
                         int nset, long ustride) {
  const int L = (lmax + 1) * (lmax + 1);
  NAPMO_REQUIRE(nset >= 1 && nset <= 65535, "synthesis tables: too many sets");
  dim3 grid((unsigned)div_up(at.max_nrad - 1, SYNTH_TAB_J), at.M, nset);
  const SynthTabLayout lay{synth_table_size(at, lmax), (long)L * 4, 4, (long)at.M * L, L, 1, nset};
  k_synth_tables<<<grid, 256, 0, s>>>(at, Uy, Udt, lmax, L, write_amps, T, ampL, ampR, ustride, lay);
  NAPMO_HIP(hipGetLastError());
}

void launch_synth_tables_cols(const AtomTables &at, const double *Uy, const double *Udt, int lmax,
                              int nset, long ustride, int ndp, double *T, double *ampL,
                              double *ampR, hipStream_t s) {
  const int L = (lmax + 1) * (lmax + 1);
  NAPMO_REQUIRE(nset >= 1 && nset <= ndp && ndp % 16 == 0 && ndp <= 65535,
                "interleaved synthesis tables: 1 <= sets <= columns, columns a multiple of 16");
  dim3 grid((unsigned)div_up(at.max_nrad - 1, SYNTH_TAB_J), at.M, ndp);
  const SynthTabLayout lay{4, (long)L * ndp * 4, (long)ndp * 4, 1, (long)L * ndp, ndp, nset};
  k_synth_tables<<<grid, 256, 0, s>>>(at, Uy, Udt, lmax, L, 1, T, ampL, ampR, ustride, lay);
  NAPMO_HIP(hipGetLastError());
}

// ============================================== Poisson: the transposed operator
// z = V^T y for the discrete Becke-Poisson operator v = V rho (projection, radial Green
// solves, synthesis tables, synthesis over the (point, centre) pairs, sum over centres),
// each step transposed exactly, in the reverse order. Used for the derivative of
// E = 1/2 sum_p w bw rho v = 1/2 rho^T W V rho: dE/drho = 1/2 (W v + V^T (W rho)).

// synthesis^T: one workgroup per (centre, bucket), one thread per synthesis index q; the
// bucket's pairs (64 per slot, slots of a bucket contiguous) are staged in LDS slot by slot
// and every thread adds y_p harm_q(p) u^k (spline rows) or y_p harm_q(p) d^l / d^-(l+1)
// (extrapolation amplitudes) in pair order, so the result is deterministic.
__global__ void __launch_bounds__(1024) k_synth_adjoint(AtomTables at, int lmax, int L,
                                                        const double *__restrict__ geo,
                                                        const int *__restrict__ list,
                                                        const long *__restrict__ slot0,
                                                        const double *__restrict__ y,
                                                        double *__restrict__ Tt,
                                                        double *__restrict__ aLt,
                                                        double *__restrict__ aRt) {
  __shared__ double sg[5][64];
  const int b = blockIdx.x, nb = at.max_nrad + 1;
  const int A = b / nb, loc = b % nb;
  const int nr = at.nrad[A];
  if (loc > nr) return;
  const int reg = loc == 0 ? 0 : (loc == nr ? 2 : 1);
  const int j = reg == 1 ? loc - 1 : 0;
  const int q = threadIdx.x;
  int l = 0, m = 0, cs = 0;
  if (q < L) q_decode(q, lmax, l, m, cs);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  const long s0 = slot0[b], s1 = slot0[b + 1];
  for (long w = s0; w < s1; ++w) {
    __syncthreads();
    if (threadIdx.x < 64) {
      const int lane = threadIdx.x;
      const int p = list[w * 64 + lane];
      const double *g = geo + w * 256 + lane;
      sg[0][lane] = g[0], sg[1][lane] = g[64], sg[2][lane] = g[128], sg[3][lane] = g[192];
      sg[4][lane] = p >= 0 ? y[p] : 0.0;
    }
    __syncthreads();
    if (q >= L) continue;
    for (int k = 0; k < 64; ++k) {
      const double yy = sg[4][k];
      if (yy == 0.0) continue;  // padding slots and zero cotangents add nothing
      const double g0 = sg[0][k], ux = sg[1][k], uy = sg[2][k], uz = sg[3][k];
      double Am = 1.0, Bm = 0.0;  // (ux + i uy)^m as the forward synthesis forms it
      for (int t = 0; t < m; ++t) {
        const double re = ux * Am - uy * Bm;
        Bm = ux * Bm + uy * Am;
        Am = re;
      }
      double pil = 1.0, pi1 = 1.0, pi2 = 0.0;  // scaled PI_l^m(uz), the forward recurrence
      for (int ll = m + 1; ll <= l; ++ll) {
        pil = (ll == m + 1) ? uz : uz * pi1 - pi_beta(ll, m) * pi2;
        pi2 = pi1;
        pi1 = pil;
      }
      const double h = (cs ? Bm : Am) * pil * yy;
      if (reg == 1) {
        const double u = g0;
        a0 += h;
        a1 += h * u;
        a2 += h * u * u;
        a3 += h * u * u * u;
      } else {
        // d^l (left) or d^-(l+1) (right), by the forward's running products
        const double base = (reg == 0) ? g0 : 1.0 / g0;
        double pw = (reg == 2) ? base : 1.0;
        for (int t = 0; t < l; ++t) pw *= base;
        a0 += h * pw;
      }
    }
  }
  if (q >= L) return;
  if (reg == 1) {
    double *t = Tt + (((long)A * (at.max_nrad - 1) + j) * L + q) * 4;
    t[0] = a0, t[1] = a1, t[2] = a2, t[3] = a3;
  } else {
    (reg == 0 ? aLt : aRt)[(long)A * L + q] = a0;
  }
}

// synthesis tables^T: one workgroup per (radial node, centre), one thread per q; node i
// collects interval i (as its left end) and interval i - 1 (as its right end), and the
// extrapolation amplitudes at the first and last node.
__global__ void k_synth_tables_adjoint(AtomTables at, int lmax, int L, const double *__restrict__ Tt,
                                       const double *__restrict__ aLt, const double *__restrict__ aRt,
                                       double *__restrict__ Uyt, double *__restrict__ Udtt) {
  const int i = blockIdx.x, A = blockIdx.y;
  const int nr = at.nrad[A];
  if (i >= nr) return;
  const long roff = at.rad_off[A];
  for (int q = threadIdx.x; q < L; q += blockDim.x) {
    int l, m, cs;
    q_decode(q, lmax, l, m, cs);
    const int ref = l * l + (m ? 2 * m - 1 + cs : 0);
    const double f = q_factor(l, m) * pi_sigma(l, m);
    double gy = 0.0, gd = 0.0;
    if (i <= nr - 2) {
      const double *t = Tt + (((long)A * (at.max_nrad - 1) + i) * L + q) * 4;
      gy += f * (t[0] - 3 * t[2] + 2 * t[3]);
      gd += f * (t[1] - 2 * t[2] + t[3]);
    }
    if (i >= 1) {
      const double *t = Tt + (((long)A * (at.max_nrad - 1) + i - 1) * L + q) * 4;
      gy += f * (3 * t[2] - 2 * t[3]);
      gd += f * (t[3] - t[2]);
    }
    if (i == 0) gy += aLt[(long)A * L + q] * f / pow(at.r[roff], (double)l);
    if (i == nr - 1) gy += aRt[(long)A * L + q] * f * pow(at.r[roff + nr - 1], (double)(l + 1));
    Uyt[(roff + i) * L + ref] = gy;
    Udtt[(roff + i) * L + ref] = gd;
  }
}

// radial solves^T: rho_lm^T[A][j][lm] = sum_i Gy[i][j] Uy^T[A][i][lm] + Gd[i][j] Udt^T[A][i][lm].
__global__ void __launch_bounds__(256) k_radial_adjoint(AtomTables at, const double *__restrict__ green,
                                                        int L, const double *__restrict__ Uyt,
                                                        const double *__restrict__ Udtt,
                                                        double *__restrict__ rt) {
  const int A = blockIdx.z, lm = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int nr = at.nrad[A];
  if (j >= nr) return;
  const int l = (int)sqrt((double)lm + 0.5);
  const double *G = green + at.green_off[A] + (long)l * 2 * nr * nr;
  const long roff = at.rad_off[A];
  double acc = 0.0;
  for (int i = 0; i < nr; ++i)
    acc += G[(long)i * nr + j] * Uyt[(roff + i) * L + lm] + G[(long)nr * nr + (long)i * nr + j] * Udtt[(roff + i) * L + lm];
  rt[(roff + j) * L + lm] = acc;
}

// projection^T: z_p = bw_p sum_lm Yhat[j][lm] rho_lm^T[A][i][lm] on atom A's own points
// (p = point_off[A] + i na + j).
__global__ void __launch_bounds__(256) k_project_adjoint(AtomTables at, const int *__restrict__ owner,
                                                         const double *__restrict__ bw,
                                                         const double *__restrict__ yhat, int L, long N,
                                                         const double *__restrict__ rt,
                                                         double *__restrict__ z) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int A = owner[p];
  const int na = at.nang[A];
  const long loc = p - at.point_off[A];
  const long i = loc / na, j = loc - i * na;
  const double *Y = yhat + at.yhat_off[A] + j * L;
  const double *r = rt + (at.rad_off[A] + i) * L;
  double acc = 0.0;
  for (int lm = 0; lm < L; ++lm) acc += Y[lm] * r[lm];
  z[p] = bw[p] * acc;
}

__global__ void k_half_sum(const double *__restrict__ a, const double *__restrict__ b, long N, long Npad,
                           double *__restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < Npad) out[p] = p < N ? 0.5 * (a[p] + b[p]) : 0.0;
}

void launch_half_sum(const double *a, const double *b, long N, long Npad, double *out, hipStream_t st) {
  if (Npad <= 0) return;
  k_half_sum<<<div_up(Npad, 256), 256, 0, st>>>(a, b, N, Npad, out);
  NAPMO_HIP(hipGetLastError());
}

void launch_poisson_adjoint(const AtomTables &at, int lmax, const double *geo, const int *list,
                            const long *slot0, const int *owner, const double *bw,
                            const double *yhat, const double *green, long N, const double *y,
                            double *Tt, double *aLt, double *aRt, double *Uyt, double *Udtt,
                            double *rt, double *z, hipStream_t s) {
  NAPMO_REQUIRE(lmax >= 0 && lmax <= MAX_LMAX, "lmax out of range for the transposed synthesis");
  const int L = (lmax + 1) * (lmax + 1);
  const int nb = at.M * (at.max_nrad + 1);
  const int threads = (int)std::min<long>(1024, round_up(L, 64));
  k_synth_adjoint<<<nb, threads, 0, s>>>(at, lmax, L, geo, list, slot0, y, Tt, aLt, aRt);
  k_synth_tables_adjoint<<<dim3(at.max_nrad, at.M), 256, 0, s>>>(at, lmax, L, Tt, aLt, aRt, Uyt, Udtt);
  k_radial_adjoint<<<dim3(div_up(at.max_nrad, 256), L, at.M), 256, 0, s>>>(at, green, L, Uyt, Udtt, rt);
  if (N > 0) k_project_adjoint<<<div_up(N, 256), 256, 0, s>>>(at, owner, bw, yhat, L, N, rt, z);
  NAPMO_HIP(hipGetLastError());
}

// ================================================================ Poisson: synthesis
// (kernel template in synth_kernel.hpp, instantiated per lmax in synth_inst_*.hip)
template <int LM>
static void synth_dispatch(int lmax, dim3 g, dim3 b, hipStream_t s, const AtomTables &at,
                           const double *x, const double *y, const double *z, long N,
                           const double *T, const double *aL, const double *aR, int acc,
                           int ext_mode, double *v) {
  if constexpr (LM > MAX_LMAX) {
    throw Error("lmax above the supported maximum " + std::to_string(MAX_LMAX));
  } else {
    if (lmax == LM)
      launch_synth_t<LM>(g, b, s, at, x, y, z, N, T, aL, aR, acc, ext_mode, v);
    else
      synth_dispatch<LM + 1>(lmax, g, b, s, at, x, y, z, N, T, aL, aR, acc, ext_mode, v);
  }
}
void launch_synth(const AtomTables &at, const double *x, const double *y, const double *z, long N,
                  const double *T, const double *ampL, const double *ampR, int lmax, int accumulate,
                  double *v, hipStream_t s, int ext_mode) {
  NAPMO_REQUIRE(lmax >= 0 && lmax <= MAX_LMAX, "lmax out of range for synthesis");
  if (N <= 0) return;
  synth_dispatch<0>(lmax, dim3(div_up(N, 256)), dim3(256), s, at, x, y, z, N, T, ampL, ampR,
                    accumulate, ext_mode, v);
  NAPMO_HIP(hipGetLastError());
}

// ----------------------------------------------------- bucketed synthesis support
__device__ inline int pair_bucket(const AtomTables &at, int A, double X, double Y, double Z) {
  const double dx = X - at.center[3 * A], dy = Y - at.center[3 * A + 1], dz = Z - at.center[3 * A + 2];
  const double d = sqrt(dx * dx + dy * dy + dz * dz);
  const int nr = at.nrad[A];
  int loc;
  if (d < at.first_x[A]) {
    loc = 0;
  } else if (d <= at.last_x[A]) {
    const int kind = at.rkind[A];
    const double t = kind == 0 ? -(acos((d - at.rp0[A]) / (d + at.rp0[A])) / at.rp1[A]) + nr
                               : (kind == 1 ? pow(d / at.rp0[A], 1.0 / at.rp1[A]) - 1 : d);
    int j = (int)floor(t);
    if (j >= nr - 1) j = nr - 2;
    if (j < 0) j = 0;
    loc = 1 + j;
  } else {
    loc = nr;
  }
  return A * (at.max_nrad + 1) + loc;
}

// Per workgroup (256 points, one centre): LDS histogram of the centre's local buckets,
// then one global atomic per non-empty bin (pair_bucket is A * (max_nrad + 1) + loc).
constexpr int PAIR_BINS = 1024;  // local buckets per centre (max_nrad + 1 <= PAIR_BINS)

__global__ void __launch_bounds__(256) k_pair_count(AtomTables at, const double *__restrict__ x,
                                                    const double *__restrict__ y,
                                                    const double *__restrict__ z, long N,
                                                    int *__restrict__ count) {
  __shared__ int hist[PAIR_BINS];
  const int A = blockIdx.y, nb = at.max_nrad + 1;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) hist[b] = 0;
  __syncthreads();
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N) atomicAdd(&hist[pair_bucket(at, A, x[p], y[p], z[p]) - A * nb], 1);
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    if (hist[b]) atomicAdd(&count[A * nb + b], hist[b]);
}

__global__ void __launch_bounds__(256) k_pair_fill(AtomTables at, const double *__restrict__ x,
                                                   const double *__restrict__ y,
                                                   const double *__restrict__ z, long N,
                                                   const long *__restrict__ off,
                                                   int *__restrict__ cursor,
                                                   int *__restrict__ list) {
  __shared__ int hist[PAIR_BINS], base[PAIR_BINS];
  const int A = blockIdx.y, nb = at.max_nrad + 1;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) hist[b] = 0;
  __syncthreads();
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int loc = -1, rank = 0;
  if (p < N) {
    loc = pair_bucket(at, A, x[p], y[p], z[p]) - A * nb;
    rank = atomicAdd(&hist[loc], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    if (hist[b]) base[b] = atomicAdd(&cursor[A * nb + b], hist[b]);
  __syncthreads();
  if (loc >= 0) list[off[A * nb + loc] + base[loc] + rank] = (int)p;
}

void launch_pair_count(const AtomTables &at, const double *x, const double *y, const double *z,
                       long N, int *count, hipStream_t s) {
  NAPMO_REQUIRE(at.max_nrad + 1 <= PAIR_BINS, "too many radial points for the synthesis buckets");
  k_pair_count<<<dim3(div_up(N, 256), at.M), 256, 0, s>>>(at, x, y, z, N, count);
  NAPMO_HIP(hipGetLastError());
}

void launch_pair_fill(const AtomTables &at, const double *x, const double *y, const double *z,
                      long N, const long *off, int *cursor, int *list, hipStream_t s) {
  NAPMO_REQUIRE(at.max_nrad + 1 <= PAIR_BINS, "too many radial points for the synthesis buckets");
  k_pair_fill<<<dim3(div_up(N, 256), at.M), 256, 0, s>>>(at, x, y, z, N, off, cursor, list);
  NAPMO_HIP(hipGetLastError());
}

// One wavefront per 64-pair slot: the geometry each synthesis launch would otherwise form
// per pair (gathers of the point, distance, unit vector and the inverse radial transform),
// formed once per grid. Slots past a bucket's pairs (list entry -1) get a fixed finite
// geometry (unit distance along x, offset 0 in the interval) whose results are never used.
__global__ void __launch_bounds__(64) k_pair_geometry(AtomTables at, const double *__restrict__ px,
                                                      const double *__restrict__ py,
                                                      const double *__restrict__ pz,
                                                      const int *__restrict__ list,
                                                      const int *__restrict__ wave_bucket,
                                                      double *__restrict__ geo) {
  const long w = blockIdx.x;
  const int lane = threadIdx.x;
  const int gb = wave_bucket[w];
  const int A = gb / (at.max_nrad + 1), loc = gb - A * (at.max_nrad + 1);
  const int nr = at.nrad[A];
  const int reg = (loc == 0) ? 0 : (loc == nr ? 2 : 1);
  const int j = (reg == 1) ? loc - 1 : 0;
  const int p = list[w * 64 + lane];
  double *g = geo + w * 256 + lane;
  if (p < 0) {
    g[0] = (reg == 1) ? 0.0 : 1.0;
    g[64] = 1.0;
    g[128] = 0.0;
    g[192] = 0.0;
    return;
  }
  const double cx = at.center[3 * A], cy = at.center[3 * A + 1], cz = at.center[3 * A + 2];
  const double dx = px[p] - cx, dy = py[p] - cy, dz = pz[p] - cz;
  const double d = sqrt(dx * dx + dy * dy + dz * dz);
  const double inv_d = 1.0 / d;
  g[0] = (reg == 1) ? rtf_inv(at.rkind[A], at.rp0[A], at.rp1[A], nr, d) - j : d;
  g[64] = dx * inv_d;
  g[128] = dy * inv_d;
  g[192] = dz * inv_d;
}

void launch_pair_geometry(const AtomTables &at, const double *x, const double *y,
                          const double *z, const int *list, const int *wave_bucket, long nwaves,
                          double *geo, hipStream_t s) {
  if (nwaves <= 0) return;
  k_pair_geometry<<<dim3((unsigned)nwaves), 64, 0, s>>>(at, x, y, z, list, wave_bucket, geo);
  NAPMO_HIP(hipGetLastError());
}

template <int LM>
static void synth2_dispatch(int lmax, int np, hipStream_t s, const AtomTables &at,
                            const double *geo, long N,
                            const int *list, const int *wb, long nw, const double *T,
                            const double *aL, const double *aR, double *contrib) {
  if constexpr (LM > MAX_LMAX) {
    throw Error("lmax above the supported maximum " + std::to_string(MAX_LMAX));
  } else {
    if (lmax == LM)
      launch_synth2_t<LM>(np, s, at, geo, N, list, wb, nw, T, aL, aR, contrib);
    else
      synth2_dispatch<LM + 1>(lmax, np, s, at, geo, N, list, wb, nw, T, aL, aR, contrib);
  }
}

void launch_synth2(const AtomTables &at, const double *geo, long N,
                   const int *list, const int *wave_bucket, long nwaves, const double *T,
                   const double *ampL, const double *ampR, int lmax, double *contrib,
                   hipStream_t s, int np) {
  NAPMO_REQUIRE(lmax >= 0 && lmax <= MAX_LMAX, "lmax out of range for synthesis");
  NAPMO_REQUIRE(np == 1 || np == 2, "synthesis: 1 or 2 pairs per lane");
  NAPMO_REQUIRE(nwaves % (SYNTH_PAD / 64) == 0, "synthesis: pair lists must be padded to 128");
  if (nwaves <= 0) return;
  synth2_dispatch<0>(lmax, np, s, at, geo, N, list, wave_bucket, nwaves, T, ampL, ampR, contrib);
  NAPMO_HIP(hipGetLastError());
}

__global__ void k_contrib_reduce(const double *__restrict__ contrib, int M, long N, int accumulate,
                                 double scale, double *__restrict__ v) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  double s = 0.0;
  for (int A = 0; A < M; ++A) s += contrib[(long)A * N + p];
  v[p] = accumulate ? v[p] + scale * s : scale * s;
}

// v = sum_A contrib[A] and, in the same pass, sv = v wbw (zero for N <= p < Npad).
__global__ void k_contrib_reduce_scale(const double *__restrict__ contrib, int M, long N, long Npad,
                                       const double *__restrict__ wbw, double *__restrict__ v,
                                       double *__restrict__ sv) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Npad) return;
  if (p >= N) {
    sv[p] = 0.0;
    return;
  }
  double s = 0.0;
  for (int A = 0; A < M; ++A) s += contrib[(long)A * N + p];
  const double vp = 1.0 * s;
  v[p] = vp;
  sv[p] = vp * wbw[p];
}

void launch_contrib_reduce(const double *contrib, int M, long N, int accumulate, double *v,
                           hipStream_t s, double scale) {
  if (N <= 0) return;
  k_contrib_reduce<<<div_up(N, 256), 256, 0, s>>>(contrib, M, N, accumulate, scale, v);
  NAPMO_HIP(hipGetLastError());
}

// One thread per (point, density) of a tile, 16 densities of a point adjacent: the sum over
// centres in index order, as k_contrib_reduce.
__global__ void k_contrib_reduce_tile(const double *__restrict__ tile, int M, long N, int nvalid,
                                      double scale, double *__restrict__ out, long ostride) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long p = e >> 4;
  const int k = (int)(e & 15);
  if (p >= N || k >= nvalid) return;
  double s = 0.0;
  for (int A = 0; A < M; ++A) s += tile[((long)A * N + p) * 16 + k];
  out[k * ostride + p] = scale * s;
}

void launch_contrib_reduce_tile(const double *tile, int M, long N, int nvalid, double scale,
                                double *out, long ostride, hipStream_t s) {
  if (N <= 0 || nvalid <= 0) return;
  NAPMO_REQUIRE(nvalid <= 16, "tile reduction: at most 16 densities");
  k_contrib_reduce_tile<<<div_up(N * 16, 256), 256, 0, s>>>(tile, M, N, nvalid, scale, out, ostride);
  NAPMO_HIP(hipGetLastError());
}

template <int LM>
static void synth_mfma_dispatch(int lmax, hipStream_t s, const AtomTables &at, const double *geo,
                                long N, const int *list, const int *wb, long nw, int ntile,
                                const double *T, const double *aL, const double *aR,
                                double *contrib) {
  if constexpr (LM > MAX_LMAX) {
    throw Error("lmax above the supported maximum " + std::to_string(MAX_LMAX));
  } else {
    if (lmax == LM)
      launch_synth_mfma_t<LM>(s, at, geo, N, list, wb, nw, ntile, 16 * ntile, T, aL, aR,
                              (long)at.M * N * 16, contrib);
    else
      synth_mfma_dispatch<LM + 1>(lmax, s, at, geo, N, list, wb, nw, ntile, T, aL, aR, contrib);
  }
}

void launch_synth_mfma(const AtomTables &at, const double *geo, long N, const int *list,
                       const int *wave_bucket, long nwaves, int ntile, const double *T,
                       const double *ampL, const double *ampR, int lmax, double *contrib,
                       hipStream_t s) {
  NAPMO_REQUIRE(lmax >= 0 && lmax <= MAX_LMAX, "lmax out of range for synthesis");
  NAPMO_REQUIRE(ntile >= 1 && ntile <= 4096, "MFMA synthesis: 1 to 4096 tiles of 16 densities");
  if (nwaves <= 0) return;
  synth_mfma_dispatch<0>(lmax, s, at, geo, N, list, wave_bucket, nwaves, ntile, T, ampL, ampR,
                         contrib);
  NAPMO_HIP(hipGetLastError());
}

__global__ void k_contrib_reduce_scale_group(const ScaleMember *__restrict__ mem, int nmem,
                                             const double *__restrict__ contrib,
                                             const long *__restrict__ cbase,
                                             const double *__restrict__ wbw,
                                             double *__restrict__ v) {
  const long g = blockIdx.x;
  const ScaleMember &m = mem[group_member(mem, nmem, g)];
  const long q = (g - m.block0) * blockDim.x + threadIdx.x;  // the member's own point index
  if (q >= m.Npad) return;
  if (q >= m.N) {
    m.s[q] = 0.0;
    return;
  }
  const long p = m.p0 + q;
  double s = 0.0;
  for (int A = m.A0; A < m.A1; ++A) s += contrib[cbase[A] + p];
  const double vp = 1.0 * s;
  v[p] = vp;
  m.s[q] = vp * wbw[p];
}

void launch_contrib_reduce_scale_group(const ScaleMember *d_mem, int nmem, long nblocks,
                                       const double *contrib, const long *cbase,
                                       const double *wbw, double *v, hipStream_t st) {
  if (nmem <= 0 || nblocks <= 0) return;
  k_contrib_reduce_scale_group<<<dim3((unsigned)nblocks), 256, 0, st>>>(d_mem, nmem, contrib, cbase,
                                                                        wbw, v);
  NAPMO_HIP(hipGetLastError());
}

__global__ void k_apmo_reduce_group(const ScaleMember *__restrict__ mem, int nmem,
                                    const SpeciesSum *__restrict__ sp,
                                    const double *__restrict__ contrib, long Nu,
                                    const double *__restrict__ wbw, double *__restrict__ vown,
                                    double *__restrict__ veff) {
  const long g = blockIdx.x;
  const int ms = group_member(mem, nmem, g);
  const ScaleMember &m = mem[ms];
  const long q = (g - m.block0) * blockDim.x + threadIdx.x;
  if (q >= m.Npad) return;
  if (q >= m.N) {
    m.s[q] = 0.0;
    return;
  }
  const long p = m.p0 + q;
  double own = 0.0, vc = 0.0;
  bool first = true;
  for (int t = 0; t < nmem; ++t) {
    double s = 0.0;
    for (int A = sp[t].A0; A < sp[t].A1; ++A) s += contrib[(long)A * Nu + p];
    if (t == ms) {
      own = 1.0 * s;
    } else {
      vc = first ? sp[t].q * s : vc + sp[t].q * s;
      first = false;
    }
  }
  const double qs = sp[ms].q;
  const double ve = first ? qs * qs * own : qs * qs * own + qs * vc;
  vown[p] = own;
  veff[p] = ve;
  m.s[q] = ve * wbw[p];
}

void launch_apmo_reduce_group(const ScaleMember *d_mem, int nmem, long nblocks,
                              const SpeciesSum *sp, const double *contrib, long Nu,
                              const double *wbw, double *vown, double *veff, hipStream_t st) {
  if (nmem <= 0 || nblocks <= 0) return;
  k_apmo_reduce_group<<<dim3((unsigned)nblocks), 256, 0, st>>>(d_mem, nmem, sp, contrib, Nu, wbw,
                                                               vown, veff);
  NAPMO_HIP(hipGetLastError());
}

__global__ void k_apmo_energy_group(const ApmoEnergyMember *__restrict__ mem, int ns) {
  const int s = threadIdx.x;
  if (s >= ns) return;
  const ApmoEnergyMember &e = mem[s];
  e.E[2] = e.q2 * e.E2[1];
  e.E[3] = e.E[1] - e.E[2];
}

void launch_apmo_energy_group(const ApmoEnergyMember *d_mem, int ns, hipStream_t st) {
  NAPMO_REQUIRE(ns >= 1 && ns <= 256, "APMO energies: 1 to 256 species");
  k_apmo_energy_group<<<1, 256, 0, st>>>(d_mem, ns);
  NAPMO_HIP(hipGetLastError());
}

void launch_contrib_reduce_scale(const double *contrib, int M, long N, long Npad, const double *wbw,
                                 double *v, double *sv, hipStream_t s) {
  if (Npad <= 0) return;
  k_contrib_reduce_scale<<<div_up(Npad, 256), 256, 0, s>>>(contrib, M, N, Npad, wbw, v, sv);
  NAPMO_HIP(hipGetLastError());
}

// V[p] = sum_A q_A / |p - R_A|, charges in list order (napmo/wavefunction/nnuclear.py:15-37).
__global__ void k_nuclear(const double *__restrict__ x, const double *__restrict__ y,
                          const double *__restrict__ z, long N, const double *__restrict__ q,
                          const double *__restrict__ o, int nq, double *__restrict__ V) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  double s = 0.0;
  for (int a = 0; a < nq; ++a) {
    const double dx = o[3 * a] - x[p], dy = o[3 * a + 1] - y[p], dz = o[3 * a + 2] - z[p];
    s += q[a] / sqrt(dx * dx + dy * dy + dz * dz);
  }
  V[p] = s;
}

void launch_nuclear(const double *x, const double *y, const double *z, long N, const double *q,
                    const double *o, int nq, double *V, hipStream_t s) {
  if (N <= 0) return;
  k_nuclear<<<div_up(N, 256), 256, 0, s>>>(x, y, z, N, q, o, nq, V);
  NAPMO_HIP(hipGetLastError());
}

// Exchange-potential helpers (napmo/wavefunction/psi_numeric.py:186-215).
// chi[l][p] = sum_k D[k][l] psi[k][p] (psi, chi [ndim x N] function-major).
__global__ void k_exchange_chi(const double *__restrict__ psi, const double *__restrict__ D,
                               int ndim, long N, double *__restrict__ chi) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int l = blockIdx.y;
  if (p >= N) return;
  double s = 0.0;
  for (int k = 0; k < ndim; ++k) s += D[(long)k * ndim + l] * psi[(long)k * N + p];
  chi[(long)l * N + p] = s;
}

// out[p] = a[p] * b[p]
__global__ void k_mul(const double *__restrict__ a, const double *__restrict__ b, long N,
                      double *__restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N) out[p] = a[p] * b[p];
}

// K[v] += aux chi[s]; K[s] += aux chi[v] when s != v (aux = V[psi_s psi_v]).
__global__ void k_exchange_acc(const double *__restrict__ aux, const double *__restrict__ chi,
                               int s, int v, long N, double *__restrict__ K) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const double a = aux[p];
  K[(long)v * N + p] += a * chi[(long)s * N + p];
  if (s != v) K[(long)s * N + p] += a * chi[(long)v * N + p];
}

void launch_exchange_chi(const double *psi, const double *D, int ndim, long N, double *chi,
                         hipStream_t st) {
  if (N <= 0 || ndim <= 0) return;
  k_exchange_chi<<<dim3((unsigned)div_up(N, 256), ndim), 256, 0, st>>>(psi, D, ndim, N, chi);
  NAPMO_HIP(hipGetLastError());
}
void launch_mul(const double *a, const double *b, long N, double *out, hipStream_t st) {
  if (N <= 0) return;
  k_mul<<<div_up(N, 256), 256, 0, st>>>(a, b, N, out);
  NAPMO_HIP(hipGetLastError());
}
void launch_exchange_acc(const double *aux, const double *chi, int s, int v, long N, double *K,
                         hipStream_t st) {
  if (N <= 0) return;
  k_exchange_acc<<<div_up(N, 256), 256, 0, st>>>(aux, chi, s, v, N, K);
  NAPMO_HIP(hipGetLastError());
}

// v[p] = a x[p] + b y[p]
__global__ void k_axpby(double a, const double *__restrict__ x, double b,
                        const double *__restrict__ y, long N, double *__restrict__ v) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N) v[p] = a * x[p] + b * y[p];
}

void launch_axpby(double a, const double *x, double b, const double *y, long N, double *v,
                  hipStream_t s) {
  if (N <= 0) return;
  k_axpby<<<div_up(N, 256), 256, 0, s>>>(a, x, b, y, N, v);
  NAPMO_HIP(hipGetLastError());
}

// ============================================ FP64 MFMA layout; density partial sums
// v_mfma_f64_16x16x4f64 operand layout (lane i of 64), as measured on gfx950:
//   A[16 x 4]:  lane i holds A[i % 16][i / 16]
//   B[4 x 16]:  lane i holds B[i / 16][i % 16]
//   D[16 x 16]: lane i holds D[(i / 16) + 4 r][i % 16] in accumulator element r = 0..3
constexpr int KS = 16;  // K elements staged in LDS per step (FP64 Fock kernel)

__global__ void k_rho_reduce(const double *__restrict__ partial, int nt, long Npad, long N,
                             double scale, double *__restrict__ rho) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  double s = 0.0;
  for (int t = 0; t < nt; ++t) s += partial[(long)t * Npad + p];
  rho[p] = scale * s;
}

__global__ void k_rho_reduce_group(const RhoReduceMember *__restrict__ mem, int nmem) {
  const long g = blockIdx.x;
  const RhoReduceMember &r = mem[group_member(mem, nmem, g)];
  const long p = (g - r.block0) * blockDim.x + threadIdx.x;
  if (p >= r.N) return;
  double s = 0.0;
  for (int t = 0; t < r.ntile; ++t) s += r.partial[(long)t * r.Npad + p];
  r.rho[p] = 1.0 * s;
}

void launch_rho_reduce_group(const RhoReduceMember *d_mem, int nmem, long nblocks, hipStream_t st) {
  if (nmem <= 0 || nblocks <= 0) return;
  k_rho_reduce_group<<<dim3((unsigned)nblocks), 256, 0, st>>>(d_mem, nmem);
  NAPMO_HIP(hipGetLastError());
}

void launch_rho_reduce(const double *partial, int ntile, long Npad, long N, double *rho,
                       hipStream_t s, double scale) {
  if (N <= 0) return;
  k_rho_reduce<<<div_up(N, 256), 256, 0, s>>>(partial, ntile, Npad, N, scale, rho);
  NAPMO_HIP(hipGetLastError());
}

__global__ void k_scale3(const double *__restrict__ a, const double *__restrict__ b,
                         const double *__restrict__ c, long N, long Npad, double *__restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Npad) return;
  out[p] = (p < N) ? (c ? a[p] * b[p] * c[p] : a[p] * b[p]) : 0.0;
}

void launch_scale3(const double *a, const double *b, const double *c, long N, long Npad, double *s,
                   hipStream_t st) {
  k_scale3<<<div_up(Npad, 256), 256, 0, st>>>(a, b, c, N, Npad, s);
  NAPMO_HIP(hipGetLastError());
}

// =================================================== contractions: FP64 MFMA Fock
// F[mu][nu] = sum_p X[p][mu] s[p] Y[p][nu]: A[m][k] = X[p0+k][mu0+m], B[k][n] =
// s[p0+k] Y[p0+k][nu0+n]. Split over points (chunks); one workgroup per (tile, chunk).
FockShape fock_shape(int ld, long Npad, bool symmetric, long wg_target) {
  FockShape sh;
  sh.ld = ld;
  sh.nt = ld / TILE;
  sh.ntiles = symmetric ? sh.nt * (sh.nt + 1) / 2 : sh.nt * sh.nt;
  sh.symmetric = symmetric;
  sh.Npad = Npad;
  // aim at ~2048 workgroups (or wg_target); chunks a multiple of 256 points (the mixed
  // kernel's FP32 accumulation block)
  long want = std::max(1L, (wg_target > 0 ? wg_target : 2048L) / sh.ntiles);
  long chunk = round_up(div_up(Npad, want), 256);
  sh.chunk = std::max<long>(chunk, 256);
  sh.nchunk = (int)div_up(Npad, sh.chunk);
  return sh;
}

__device__ inline void tile_coords(int t, int nt, bool symmetric, int &ti, int &tj) {
  if (!symmetric) {
    ti = t / nt, tj = t % nt;
    return;
  }
  int i = 0, rem = t;
  while (rem >= nt - i) rem -= nt - i, ++i;
  ti = i, tj = i + rem;
}

// Points run in stages of KS: the next stage's rows (4 consecutive doubles of one point
// per thread for each operand) are loaded into registers while the current stage's MFMAs
// run, then stored to the other LDS buffer (one barrier per stage). LDS rows hold 64
// functions plus 16 doubles of padding, so a 16 x 4 fragment read touches each bank once
// per half-wave.
constexpr int FSTRIDE = TILE + 16;

__device__ __forceinline__ void fock_f64_body(const double *__restrict__ X,
                                              const double *__restrict__ Yv,
                                              const double *__restrict__ s, int ld, int nt,
                                              int symmetric, int ntiles, int nchunk, long Npad,
                                              long chunk, double *__restrict__ partial, int g) {
  __shared__ double As[2][KS][FSTRIDE];  // As[.][k][m] = X[p0 + k][mu0 + m]
  __shared__ double Bs[2][KS][FSTRIDE];  // Bs[.][k][n] = s[p0 + k] Y[p0 + k][nu0 + n]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // XCD-aware order: workgroup g runs on XCD g % 8; all tiles of one chunk of points get
  // ids congruent mod 8 (g = 8 i + x: tile i % ntiles, chunk (i / ntiles) * 8 + x), so the
  // chunk's Phi rows are shared through one L2
  const int xcd = g & 7, gi = g >> 3;
  const int tile = gi % ntiles, chunk_id = (gi / ntiles) * 8 + xcd;
  if (chunk_id >= nchunk) return;
  int ti, tj;
  tile_coords(tile, nt, symmetric, ti, tj);
  const int mu0 = ti * TILE, nu0 = tj * TILE;
  const bool same = (X == Yv) && (ti == tj);
  const long pbeg = (long)chunk_id * chunk;
  const long pend = min(Npad, pbeg + chunk);
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = tid >> 4, lc = (tid & 15) * 4;  // this thread's stage row and 4 columns
  d4 ra, rb;
  double rs = 0.0;
  auto load = [&](long k0) {
    const long p = k0 + lr;
    ra = *(const d4 *)&X[p * ld + mu0 + lc];
    if (!same) rb = *(const d4 *)&Yv[p * ld + nu0 + lc];
    rs = s[p];
  };
  auto store = [&](int buf) {
    const d4 b = same ? ra : rb;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      As[buf][lr][lc + e] = ra[e];
      Bs[buf][lr][lc + e] = rs * b[e];
    }
  };
  d4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = (d4){0, 0, 0, 0};
  if (pbeg < pend) {
    load(pbeg);
    store(0);
  }
  __syncthreads();
  int buf = 0;
  const int kr0 = lane >> 4, mi = lane & 15;
  for (long k0 = pbeg; k0 < pend; k0 += KS) {
    const bool more = k0 + KS < pend;
    if (more) load(k0 + KS);
#pragma unroll
    for (int ks = 0; ks < KS; ks += 4) {
      const int kr = ks + kr0;
      const double a0 = As[buf][kr][wr + mi], a1 = As[buf][kr][wr + 16 + mi];
      const double b0 = Bs[buf][kr][wc + mi], b1 = Bs[buf][kr][wc + 16 + mi];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  double *out = partial + (long)chunk_id * ((long)ntiles * TILE * TILE + 32) +
                (long)tile * TILE * TILE;
  const int col16 = lane & 15, rg = lane >> 4;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int r = 0; r < 4; ++r)
        out[(wr + a * 16 + rg + 4 * r) * TILE + wc + b * 16 + col16] = acc[a][b][r];
}

__global__ void __launch_bounds__(256) k_fock_f64(const double *__restrict__ X,
                                                  const double *__restrict__ Yv,
                                                  const double *__restrict__ s, int ld, int nt,
                                                  int symmetric, int ntiles, int nchunk, long Npad,
                                                  long chunk, double *__restrict__ partial) {
  fock_f64_body(X, Yv, s, ld, nt, symmetric, ntiles, nchunk, Npad, chunk, partial, blockIdx.x);
}

// symmetric F = Phi^T diag(s) Phi of many plans in one launch
__global__ void __launch_bounds__(256) k_fock_f64_group(const FockMember *__restrict__ mem, int nmem) {
  const long g = blockIdx.x;
  const FockMember &f = mem[group_member(mem, nmem, g)];
  fock_f64_body(f.X, f.X, f.s, f.ld, f.nt, 1, f.ntiles, f.nchunk, f.Npad, f.chunk, f.partial,
                (int)(g - f.block0));
}

long fock_group_fill(FockMember &f, const FockShape &sh, const double *X, const double *s,
                     double *partial, long block0) {
  NAPMO_REQUIRE(sh.symmetric && block0 % 8 == 0, "grouped Fock: symmetric shapes, aligned blocks");
  f.X = X, f.s = s, f.partial = partial, f.Npad = sh.Npad, f.chunk = sh.chunk, f.block0 = block0;
  f.ld = sh.ld, f.nt = sh.nt, f.ntiles = sh.ntiles, f.nchunk = sh.nchunk;
  return (long)sh.ntiles * round_up(sh.nchunk, 8);
}

void launch_fock_group(const FockMember *d_mem, int nmem, long nblocks, hipStream_t st) {
  if (nmem <= 0 || nblocks <= 0) return;
  k_fock_f64_group<<<dim3((unsigned)nblocks), 256, 0, st>>>(d_mem, nmem);
  NAPMO_HIP(hipGetLastError());
}

void launch_fock_f64(const FockShape &sh, const double *X, const double *Y, const double *s,
                     double *partial, hipStream_t st) {
  const dim3 grid((unsigned)(sh.ntiles * round_up(sh.nchunk, 8)));
  k_fock_f64<<<grid, 256, 0, st>>>(X, Y, s, sh.ld, sh.nt, sh.symmetric ? 1 : 0, sh.ntiles,
                                   sh.nchunk, sh.Npad, sh.chunk, partial);
  NAPMO_HIP(hipGetLastError());
}

// Sum the chunk partials of each computed tile element: 4 slices of chunks per element
// (slice k takes chunks k, k + 4, ...), combined in slice order (deterministic). Threads
// walk the tile-element space of the partials, so every load is coalesced; the mirror
// image of an off-diagonal upper tile is written by the same thread. Partials are FP64
// (per-tile kernels) or FP32 (the one-pass mixed kernel, whose accumulators are FP32).
template <class T>
__device__ __forceinline__ void fock_reduce_body(const T *__restrict__ partial, int nt, int ntiles,
                                                 int nchunk, int symmetric, int n,
                                                 int upper_mirror, double scale, int accumulate,
                                                 double *__restrict__ F, long g) {
  __shared__ double red[4][64];
  const int el = threadIdx.x & 63, slice = threadIdx.x >> 6;
  const long u = g * 64 + el;
  const long nel = (long)ntiles * TILE * TILE;
  double sum = 0.0;
  if (u < nel) {
    const long cstride = nel + 32;
    int c = slice;
    // 8 loads in flight, added in chunk order
    for (; c + 28 < nchunk; c += 32) {
      double v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = (double)partial[(long)(c + 4 * k) * cstride + u];
#pragma unroll
      for (int k = 0; k < 8; ++k) sum += v[k];
    }
    for (; c < nchunk; c += 4) sum += (double)partial[(long)c * cstride + u];
  }
  red[slice][el] = sum;
  __syncthreads();
  if (slice != 0 || u >= nel) return;
  const double tot = ((red[0][el] + red[1][el]) + red[2][el] + red[3][el]) * scale;
  const int t = (int)(u / (TILE * TILE)), r = (int)(u % (TILE * TILE)) / TILE, cc = (int)(u % TILE);
  int ti, tj;
  if (symmetric) {
    // upper tiles in row order: (0,0), (0,1), ..., (0,nt-1), (1,1), ...
    int rem = t;
    ti = 0;
    while (rem >= nt - ti) rem -= nt - ti, ++ti;
    tj = ti + rem;
  } else {
    ti = t / nt, tj = t % nt;
  }
  const int i = ti * TILE + r, j = tj * TILE + cc;
  if (i >= n || j >= n) return;
  auto put = [&](int a, int b) {
    const long e = (long)a * n + b;
    F[e] = accumulate ? F[e] + tot : tot;
  };
  if (symmetric) {
    // the upper triangle is mirrored, so F is exactly symmetric (a diagonal tile's own
    // lower half, equal to rounding, is not used)
    if (i > j) return;
    put(i, j);
    if (i != j) put(j, i);
  } else if (upper_mirror) {
    if (i <= j) {
      put(i, j);
      if (i != j) put(j, i);
    }
  } else {
    put(i, j);
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_fock_reduce(const T *__restrict__ partial, int nt,
                                                     int ntiles, int nchunk, int symmetric, int n,
                                                     int upper_mirror, double scale, int accumulate,
                                                     double *__restrict__ F) {
  fock_reduce_body<T>(partial, nt, ntiles, nchunk, symmetric, n, upper_mirror, scale, accumulate,
                      F, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_fock_reduce_group(const FockReduceMember *__restrict__ mem,
                                                           int nmem) {
  const long g = blockIdx.x;
  const FockReduceMember &f = mem[group_member(mem, nmem, g)];
  fock_reduce_body<double>(f.partial, f.nt, f.ntiles, f.nchunk, 1, f.n, 0, 1.0, 0, f.F, g - f.block0);
}

long fock_reduce_group_fill(FockReduceMember &f, const FockShape &sh, const double *partial, int n,
                            double *F, long block0) {
  NAPMO_REQUIRE(sh.symmetric, "grouped Fock reduction: symmetric shapes");
  f.partial = partial, f.F = F, f.block0 = block0;
  f.nt = sh.nt, f.ntiles = sh.ntiles, f.nchunk = sh.nchunk, f.n = n;
  return div_up((long)sh.ntiles * TILE * TILE, 64);
}

void launch_fock_reduce_group(const FockReduceMember *d_mem, int nmem, long nblocks, hipStream_t st) {
  if (nmem <= 0 || nblocks <= 0) return;
  k_fock_reduce_group<<<dim3((unsigned)nblocks), 256, 0, st>>>(d_mem, nmem);
  NAPMO_HIP(hipGetLastError());
}

void launch_fock_reduce(const FockShape &sh, const double *partial, int n, int upper_mirror,
                        double scale, int accumulate, double *F, hipStream_t st) {
  const long nel = (long)sh.ntiles * TILE * TILE;
  k_fock_reduce<double><<<div_up(nel, 64), 256, 0, st>>>(partial, sh.nt, sh.ntiles, sh.nchunk,
                                                         sh.symmetric ? 1 : 0, n, upper_mirror,
                                                         scale, accumulate, F);
  NAPMO_HIP(hipGetLastError());
}

void launch_fock_reduce_f32(const FockShape &sh, const float *partial, int n, double *F,
                            hipStream_t st) {
  const long nel = (long)sh.ntiles * TILE * TILE;
  k_fock_reduce<float><<<div_up(nel, 64), 256, 0, st>>>(partial, sh.nt, sh.ntiles, sh.nchunk,
                                                        sh.symmetric ? 1 : 0, n, 0, 1.0, 0, F);
  NAPMO_HIP(hipGetLastError());
}

// ============================================================ FP64 energy traces
// Two passes with a fixed reduction order (deterministic): per-block partial sums of
// P o F and w bw rho v, then one block adds the partials.
constexpr int TRACE_BLOCKS = 512;

// block g of nb blocks (256 threads)
__device__ __forceinline__ void traces_partial_body(const double *__restrict__ P,
                                                    const double *__restrict__ F, long nn,
                                                    const double *__restrict__ wbw,
                                                    const double *__restrict__ rho,
                                                    const double *__restrict__ v, long N,
                                                    double *__restrict__ part, int g, int nb) {
  __shared__ double sa[256], sb[256];
  const int tid = threadIdx.x;
  const long stride = (long)nb * 256;
  double a = 0.0, b = 0.0;
  if (P && F)
    for (long e = (long)g * 256 + tid; e < nn; e += stride) a += P[e] * F[e];
  if (wbw && rho && v)
    for (long p = (long)g * 256 + tid; p < N; p += stride) b += wbw[p] * rho[p] * v[p];
  sa[tid] = a;
  sb[tid] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) sa[tid] += sa[tid + w], sb[tid] += sb[tid + w];
    __syncthreads();
  }
  if (tid == 0) part[g] = sa[0], part[nb + g] = sb[0];
}

__global__ void __launch_bounds__(256) k_traces_partial(const double *__restrict__ P,
                                                        const double *__restrict__ F, long nn,
                                                        const double *__restrict__ wbw,
                                                        const double *__restrict__ rho,
                                                        const double *__restrict__ v, long N,
                                                        double *__restrict__ part) {
  traces_partial_body(P, F, nn, wbw, rho, v, N, part, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void traces_final_body(const double *__restrict__ part, int nb,
                                                  double *__restrict__ E) {
  __shared__ double sa[512], sb[512];
  const int tid = threadIdx.x;
  sa[tid] = tid < nb ? part[tid] : 0.0;
  sb[tid] = tid < nb ? part[nb + tid] : 0.0;
  __syncthreads();
  for (int w = 256; w > 0; w >>= 1) {
    if (tid < w) sa[tid] += sa[tid + w], sb[tid] += sb[tid + w];
    __syncthreads();
  }
  if (tid == 0) E[0] = 0.5 * sa[0], E[1] = 0.5 * sb[0];
}

__global__ void __launch_bounds__(512) k_traces_final(const double *__restrict__ part, int nb,
                                                      double *__restrict__ E) {
  traces_final_body(part, nb, E);
}

// grouped: member = blockIdx.y, TRACE_BLOCKS partial blocks each (the same sums as one plan)
__global__ void __launch_bounds__(256) k_traces_partial_group(const TraceMember *__restrict__ mem) {
  const TraceMember &t = mem[blockIdx.y];
  traces_partial_body(t.P, t.F, t.nn, t.wbw, t.rho, t.v, t.N, t.work, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(512) k_traces_final_group(const TraceMember *__restrict__ mem,
                                                            int nb) {
  const TraceMember &t = mem[blockIdx.x];
  traces_final_body(t.work, nb, t.E);
}

void launch_traces(const double *P, const double *F, int n, const double *wbw, const double *rho,
                   const double *v, long N, double *E, double *work, hipStream_t st) {
  k_traces_partial<<<TRACE_BLOCKS, 256, 0, st>>>(P, F, (long)n * n, wbw, rho, v, N, work);
  k_traces_final<<<1, 512, 0, st>>>(work, TRACE_BLOCKS, E);
  NAPMO_HIP(hipGetLastError());
}

void launch_traces_group(const TraceMember *d_mem, int nmem, hipStream_t st) {
  if (nmem <= 0) return;
  NAPMO_REQUIRE(nmem <= 65535, "grouped traces: too many members");
  k_traces_partial_group<<<dim3(TRACE_BLOCKS, nmem), 256, 0, st>>>(d_mem);
  k_traces_final_group<<<dim3(nmem), 512, 0, st>>>(d_mem, TRACE_BLOCKS);
  NAPMO_HIP(hipGetLastError());
}

}  // namespace napmo
