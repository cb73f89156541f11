// Synthesis of many densities at once on FP64 MFMA (instantiated per lmax in
// synth_inst_*.hip). The same sum as k_synth2n, v_s(p) = sum_A sum_q S^{A,s}_q(d) Y_q(p - R_A),
// for up to 16 densities s per wavefront: inside one (centre, radial interval) bucket every
// pair shares the interval's spline coefficients, so with the pair's features f_qc = u^c Y_q
// the sum is the product F [pairs x 4 L] . T [4 L x densities]
// (src/horton/evaluate.cpp:129-223 evaluated for every density).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace napmo {

typedef double smf_d4 __attribute__((ext_vector_type(4)));

constexpr int SYNTH_MFMA_WPB = 4;  // waves per workgroup (independent, no LDS)
constexpr int SYNTH_MFMA_PD = 4;   // coefficient loads in flight, in synthesis steps q

// One wavefront: 64 pairs of one bucket (four 16-pair row blocks b) x 16 densities.
// v_mfma_f64_16x16x4f64 per (q, b): A[i][k] from lane (i = lane % 16, k = lane / 16),
// B[k][j] from lane (k = lane / 16, j = lane % 16).
//   REG 1 (interval): k = Horner power c, A = u^c Y_q, B = T_j[q][c].
//   REG 0 / 2 (extrapolation): only k = 0 carries a term, A = d^l Y_q or d^-(l+1) Y_q,
//   B = amplitude_j[q] (the other three k rows of A are zero).
// coef points at this lane's B element of q = 0; consecutive q are qstride apart. The
// harmonics are formed by the recurrences of synth_harmonics_np (Re / Im (x + i y)^m and
// the scaled PI_l^m(z)), the (x + i y)^m factor entering the feature before the product.
template <int LMAX, int REG>
__device__ __forceinline__ void synth_mfma_body(const double (&g0)[4], const double (&ux)[4],
                                                const double (&uy)[4], const double (&uz)[4],
                                                int c, const double *__restrict__ coef,
                                                long qstride, smf_d4 (&acc)[4]) {
  constexpr int L = (LMAX + 1) * (LMAX + 1);
  constexpr int PD = SYNTH_MFMA_PD < L ? SYNTH_MFMA_PD : L;
  double w[4], base[4], pwm[4], Am[4], Bm[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    if (REG == 1) {
      const double u = g0[b], u2 = u * u;
      w[b] = c == 0 ? 1.0 : (c == 1 ? u : (c == 2 ? u2 : u2 * u));
      base[b] = 1.0, pwm[b] = 1.0;
    } else {
      w[b] = c == 0 ? 1.0 : 0.0;
      base[b] = (REG == 0) ? g0[b] : 1.0 / g0[b];
      pwm[b] = (REG == 2) ? base[b] : 1.0;
    }
    Am[b] = 1.0, Bm[b] = 0.0;
  }
  double ring[PD];
#pragma unroll
  for (int t = 0; t < PD; ++t) ring[t] = coef[t * qstride];
  int q = 0;
#pragma unroll
  for (int m = 0; m <= LMAX; ++m) {
    double fa[4], fb[4], pi1[4], pi2[4], pw[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (m > 0) {
        const double a2 = fma(ux[b], Am[b], -(uy[b] * Bm[b]));
        Bm[b] = fma(ux[b], Bm[b], uy[b] * Am[b]);
        Am[b] = a2;
        if (REG != 1) pwm[b] *= base[b];
      }
      fa[b] = w[b] * Am[b], fb[b] = w[b] * Bm[b];
      pi2[b] = 0.0, pi1[b] = 1.0, pw[b] = pwm[b];
    }
#pragma unroll
    for (int l = m; l <= LMAX; ++l) {
      double pil[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        if (l == m) {
          pil[b] = 1.0;
        } else if (l == m + 1) {
          pil[b] = uz[b];
          pi2[b] = pi1[b];
          pi1[b] = pil[b];
        } else {
          pil[b] = fma(uz[b], pi1[b], -(pi_beta(l, m) * pi2[b]));
          pi2[b] = pi1[b];
          pi1[b] = pil[b];
        }
        if (REG != 1 && l > m) pw[b] *= base[b];
      }
#pragma unroll
      for (int cs = 0; cs < (m ? 2 : 1); ++cs) {
        const double bv = ring[q % PD];
        if (q + PD < L) ring[q % PD] = coef[(long)(q + PD) * qstride];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const double f = (cs ? fb[b] : fa[b]) * pil[b];
          const double a = (REG == 1) ? f : f * pw[b];
          acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[b], 0, 0, 0);
        }
        ++q;
      }
    }
  }
}

// Work unit = (64-pair slot, 16-density tile), the tile index fastest; workgroups are dealt
// to the 8 XCDs round-robin, so unit ranges are remapped to be contiguous per XCD and the
// waves of one bucket share that XCD's L2 copy of the bucket's coefficient rows.
// Tables (density-interleaved, ndp = 16 ntile columns):
//   T [A][interval][q][ndp][4], ampL / ampR [A][q][ndp].
// Output: contrib tile t at contrib + t * tstride, layout [A][N][16] (one 128-byte row of
// 16 densities per (centre, point)); padding pairs (list entry < 0) are not written.
template <int LMAX>
__global__ void __launch_bounds__(64 * SYNTH_MFMA_WPB) k_synth_mfma(
    AtomTables at, const double *__restrict__ geo, long N, const int *__restrict__ list,
    const int *__restrict__ wave_bucket, long nslots, int ntile, int ndp,
    const double *__restrict__ T, const double *__restrict__ ampL,
    const double *__restrict__ ampR, long tstride, double *__restrict__ contrib) {
  constexpr int L = (LMAX + 1) * (LMAX + 1);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long units = nslots * ntile;
  const long per_xcd = (long)(gridDim.x >> 3);
  const long wg = (long)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  const long unit = wg * SYNTH_MFMA_WPB + wave;
  if (unit >= units) return;
  const long slot = unit / ntile;
  const int tile = (int)(unit - slot * ntile);
  const int gb = __builtin_amdgcn_readfirstlane(wave_bucket[slot]);
  const int A = gb / (at.max_nrad + 1), loc = gb - A * (at.max_nrad + 1);
  const int nr = at.nrad[A];
  const int reg = (loc == 0) ? 0 : (loc == nr ? 2 : 1);
  const int j = (reg == 1) ? loc - 1 : 0;
  const int i = lane & 15, c = lane >> 4;
  const double *gs = geo + slot * 256;
  double g0[4], ux[4], uy[4], uz[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    g0[b] = gs[16 * b + i];
    ux[b] = gs[64 + 16 * b + i];
    uy[b] = gs[128 + 16 * b + i];
    uz[b] = gs[192 + 16 * b + i];
  }
  const int pown = list[slot * 64 + lane];
  const int dcol = tile * 16 + i;
  smf_d4 acc[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) acc[b] = (smf_d4){0.0, 0.0, 0.0, 0.0};
  if (reg == 1) {
    const double *coef = T + (((long)A * (at.max_nrad - 1) + j) * L * ndp + dcol) * 4 + c;
    synth_mfma_body<LMAX, 1>(g0, ux, uy, uz, c, coef, (long)ndp * 4, acc);
  } else if (reg == 0) {
    synth_mfma_body<LMAX, 0>(g0, ux, uy, uz, c, ampL + (long)A * L * ndp + dcol, ndp, acc);
  } else {
    synth_mfma_body<LMAX, 2>(g0, ux, uy, uz, c, ampR + (long)A * L * ndp + dcol, ndp, acc);
  }
  // D element r of lane: pair 16 b + lane / 16 + 4 r, density column lane % 16
  double *out = contrib + (long)tile * tstride + (long)A * N * 16 + i;
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int p = __shfl(pown, 16 * b + c + 4 * r);
      if (p >= 0) out[(long)p * 16] = acc[b][r];
    }
}

template <int LMAX>
void launch_synth_mfma_t(hipStream_t s, const AtomTables &at, const double *geo, long N,
                         const int *list, const int *wave_bucket, long nslots, int ntile,
                         int ndp, const double *T, const double *aL, const double *aR,
                         long tstride, double *contrib) {
  const long units = nslots * ntile;
  const long wgs = div_up(units, SYNTH_MFMA_WPB);
  const long per_xcd = div_up(wgs, 8);
  k_synth_mfma<LMAX><<<dim3((unsigned)(8 * per_xcd)), dim3(64 * SYNTH_MFMA_WPB), 0, s>>>(
      at, geo, N, list, wave_bucket, nslots, ntile, ndp, T, aL, aR, tstride, contrib);
}

}  // namespace napmo
