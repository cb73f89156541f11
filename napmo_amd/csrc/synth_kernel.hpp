// Synthesis kernel template (instantiated per lmax in synth_inst_*.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace napmo {

// One thread per point, loop over centres. Per (point, centre): one inverse radial
// transform, then the m-major recurrence of the scaled PI_l^m(z) polynomials of the unit
// vector (r^2 = 1) and (x + i y)^m, each harmonic contracted with its Horner spline.
static __device__ inline double rtf_inv(int kind, double p0, double p1, int n, double r) {
  if (kind == 0) return -(acos((r - p0) / (r + p0)) / p1) + n;
  if (kind == 1) return pow(r / p0, 1.0 / p1) - 1;
  return r;
}

template <int LMAX>
__global__ void __launch_bounds__(256) k_synth(AtomTables at, const double *__restrict__ px,
                                               const double *__restrict__ py,
                                               const double *__restrict__ pz, long N,
                                               const double *__restrict__ T,
                                               const double *__restrict__ ampL,
                                               const double *__restrict__ ampR, int accumulate,
                                               int ext_mode, double *__restrict__ v) {
  constexpr int L = (LMAX + 1) * (LMAX + 1);
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const double X = px[p], Y = py[p], Z = pz[p];
  double total = accumulate ? v[p] : 0.0;
  const int rowstride = (at.max_nrad - 1) * L * 4;
  for (int A = 0; A < at.M; ++A) {
    const double dx = X - at.center[3 * A], dy = Y - at.center[3 * A + 1],
                 dz = Z - at.center[3 * A + 2];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    const double inv_d = 1.0 / d;
    const double ux = dx * inv_d, uy = dy * inv_d, uz = dz * inv_d;
    const int nr = at.nrad[A];
    // ext_mode 1 (extrapolation without tail): centres beyond the spline domain add
    // nothing (src/horton/evaluate.cpp:185 'd < rcut || has_tail').
    if (ext_mode == 1 && !(d < at.last_x[A])) continue;
    int reg = 1, j = 0;
    double u = 0.0;
    if (d < at.first_x[A]) {
      reg = 0;
    } else if (d <= at.last_x[A]) {
      const double t = rtf_inv(at.rkind[A], at.rp0[A], at.rp1[A], nr, d);
      j = (int)floor(t);
      if (j >= nr - 1) j = nr - 2;
      if (j < 0) j = 0;
      u = t - j;
    } else {
      reg = 2;
    }
    const double *row = T + (long)A * rowstride + (long)j * L * 4;
    const double *aL = ampL + (long)A * L, *aR = ampR + (long)A * L;
    const double base = (reg == 0) ? d : inv_d;  // running power for extrapolation
    double acc = 0.0;
    int q = 0;
    double Am = 1.0, Bm = 0.0;   // Re, Im of (ux + i uy)^m
    double pwm = (reg == 2) ? inv_d : 1.0;  // d^m (left) or d^-(m+1) (right)
#pragma unroll
    for (int m = 0; m <= LMAX; ++m) {
      if (m > 0) {
        const double a2 = ux * Am - uy * Bm;
        Bm = ux * Bm + uy * Am;
        Am = a2;
        pwm *= base;
      }
      double pi2 = 0.0, pi1 = 1.0, pw = pwm;
#pragma unroll
      for (int l = m; l <= LMAX; ++l) {
        // scaled recurrence (pi_sigma folded into the tables)
        double pil;
        if (l == m) {
          pil = 1.0;
        } else if (l == m + 1) {
          pil = uz;
          pi2 = pi1;
          pi1 = pil;
        } else {
          pil = uz * pi1 - pi_beta(l, m) * pi2;
          pi2 = pi1;
          pi1 = pil;
        }
        if (l > m) pw *= base;
#pragma unroll
        for (int cs = 0; cs < (m ? 2 : 1); ++cs) {
          const double harm = (cs ? Bm : Am) * pil;
          double sv;
          if (reg == 1) {
            const double *c = row + q * 4;
            sv = c[0] + u * (c[1] + u * (c[2] + u * c[3]));
          } else if (ext_mode == 0) {
            sv = (reg == 0 ? aL[q] : aR[q]) * pw;  // PotentialExtrapolation
          } else {
            sv = aL[q] * exp(aR[q] * (d - at.first_x[A]));  // CuspExtrapolation (a0, b0)
          }
          acc += sv * harm;
          ++q;
        }
      }
    }
    total += acc;
  }
  v[p] = total;
}


// Harmonic sum of one (point, centre) pair: sum_q s_q(u or d) * Re/Im[(x+iy)^m] PI_l^m(z)
// over the synthesis order q. REG 1: Horner spline rows c[q][0..3] at u; REG 0 / 2:
// PotentialExtrapolation amplitudes amp[q] * d^l (left) or amp[q] / d^(l+1) (right).
template <int LMAX, int REG>
__device__ __forceinline__ double synth_harmonics(const double *c, double u, double ux, double uy,
                                                  double uz, double d, double inv_d) {
  // Harmonic sum of one (point, centre) pair in the synthesis order q (m-major): the
  // spline value of every (l, m, cos|sin) by Horner in u (REG 1) or the extrapolation
  // amplitude times d^l / d^-(l+1) (REG 0 / 2), times the scaled PI_l^m(z) recurrence
  // (pi_sigma folded into the tables) and Re / Im (x + i y)^m.
  const double base = (REG == 0) ? d : inv_d;
  double acc = 0.0;
  int q = 0;
  double Am = 1.0, Bm = 0.0;
  double pwm = (REG == 2) ? inv_d : 1.0;
#pragma unroll
  for (int m = 0; m <= LMAX; ++m) {
    if (m > 0) {
      const double a2 = ux * Am - uy * Bm;
      Bm = ux * Bm + uy * Am;
      Am = a2;
      if (REG != 1) pwm *= base;
    }
    double pi2 = 0.0, pi1 = 1.0, pw = pwm;
#pragma unroll
    for (int l = m; l <= LMAX; ++l) {
      double pil;
      if (l == m) {
        pil = 1.0;
      } else if (l == m + 1) {
        pil = uz;
        pi2 = pi1;
        pi1 = pil;
      } else {
        pil = uz * pi1 - pi_beta(l, m) * pi2;
        pi2 = pi1;
        pi1 = pil;
      }
      if (REG != 1 && l > m) pw *= base;
#pragma unroll
      for (int cs = 0; cs < (m ? 2 : 1); ++cs) {
        const double harm = (cs ? Bm : Am) * pil;
        double sv;
        if (REG == 1) {
          const double *cq = c + q * 4;
          sv = cq[0] + u * (cq[1] + u * (cq[2] + u * cq[3]));
        } else {
          sv = c[q] * pw;
        }
        acc += sv * harm;
        ++q;
      }
    }
  }
  return acc;
}

// Bucketed variant: the (point, centre) pairs are sorted by (centre, spline interval)
// once per geometry; each wavefront handles 64 pairs of one bucket, copies that bucket's
// spline row into its own LDS slice once, and every lane then reads the coefficients as
// LDS broadcasts; only the Horner variable u differs per lane.
// bucket = A * (max_nrad + 1) + loc, loc 0 = left extrapolation, 1 + j = interval j,
// nrad_A = right extrapolation. Writes contrib[A * N + p].
template <int LMAX> struct Synth2Cfg {
  static constexpr int L = (LMAX + 1) * (LMAX + 1);
  static constexpr int WPB = (L * 4 * 8 * 4 <= 65536) ? 4 : 2;  // waves per block (LDS <= 64 KB)
};

template <int LMAX>
__global__ void __launch_bounds__(256) k_synth2(AtomTables at, const double *__restrict__ px,
                                                const double *__restrict__ py,
                                                const double *__restrict__ pz, long N,
                                                const int *__restrict__ list,
                                                const int *__restrict__ wave_bucket, long nwaves,
                                                const double *__restrict__ T,
                                                const double *__restrict__ ampL,
                                                const double *__restrict__ ampR,
                                                double *__restrict__ contrib) {
  constexpr int L = Synth2Cfg<LMAX>::L;
  constexpr int WPB = Synth2Cfg<LMAX>::WPB;
  __shared__ double srow[WPB][L * 4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long wid = (long)blockIdx.x * WPB + wave;
  const bool active = wid < nwaves;
  const int gb = active ? __builtin_amdgcn_readfirstlane(wave_bucket[wid]) : 0;
  const int A = gb / (at.max_nrad + 1), loc = gb - A * (at.max_nrad + 1);
  const int nr = at.nrad[A];
  const int reg = (loc == 0) ? 0 : (loc == nr ? 2 : 1);
  const int j = (reg == 1) ? loc - 1 : 0;
  if (active && reg == 1) {
    const double *row = T + ((long)A * (at.max_nrad - 1) + j) * (long)L * 4;
    for (int e = lane; e < L * 4; e += 64) srow[wave][e] = row[e];
  }
  __syncthreads();
  if (!active) return;
  const int p = list[wid * 64 + lane];
  const long pp = p >= 0 ? p : 0;
  const double dx = px[pp] - at.center[3 * A], dy = py[pp] - at.center[3 * A + 1],
               dz = pz[pp] - at.center[3 * A + 2];
  const double d = sqrt(dx * dx + dy * dy + dz * dz);
  const double inv_d = 1.0 / d;
  const double ux = dx * inv_d, uy = dy * inv_d, uz = dz * inv_d;
  double acc;
  if (reg == 1) {
    const double u = rtf_inv(at.rkind[A], at.rp0[A], at.rp1[A], nr, d) - j;
    acc = synth_harmonics<LMAX, 1>(srow[wave], u, ux, uy, uz, d, inv_d);
  } else if (reg == 0) {
    acc = synth_harmonics<LMAX, 0>(ampL + (long)A * L, 0.0, ux, uy, uz, d, inv_d);
  } else {
    acc = synth_harmonics<LMAX, 2>(ampR + (long)A * L, 0.0, ux, uy, uz, d, inv_d);
  }
  if (p >= 0) contrib[(long)A * N + p] = acc;
}

template <int LMAX>
void launch_synth2_t(dim3 g, dim3 b, hipStream_t s, const AtomTables &at, const double *x,
                     const double *y, const double *z, long N, const int *list,
                     const int *wave_bucket, long nwaves, const double *T, const double *aL,
                     const double *aR, double *contrib) {
  (void)g, (void)b;
  constexpr int WPB = Synth2Cfg<LMAX>::WPB;
  k_synth2<LMAX><<<dim3(div_up(nwaves, WPB)), dim3(64 * WPB), 0, s>>>(at, x, y, z, N, list,
                                                                      wave_bucket, nwaves, T, aL,
                                                                      aR, contrib);
}

template <int LMAX>
void launch_synth_t(dim3 g, dim3 b, hipStream_t s, const AtomTables &at, const double *x,
                    const double *y, const double *z, long N, const double *T, const double *aL,
                    const double *aR, int acc, int ext_mode, double *v) {
  k_synth<LMAX><<<g, b, 0, s>>>(at, x, y, z, N, T, aL, aR, acc, ext_mode, v);
}

}  // namespace napmo
