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

// Same sum for NP pairs of one bucket at once: every coefficient read (from LDS or the
// amplitude table) serves NP pairs. Per pair the operations and their order are those of
// synth_harmonics, so the result is bit-identical to it.
template <int LMAX, int REG, int NP>
__device__ __forceinline__ void synth_harmonics_np(const double *c, const double (&u)[NP],
                                                   const double (&ux)[NP], const double (&uy)[NP],
                                                   const double (&uz)[NP], const double (&d)[NP],
                                                   const double (&inv_d)[NP], double (&acc)[NP]) {
  // every multiply-add is an explicit fma and nothing else is contracted, so the result
  // does not depend on how the compiler schedules NP pairs
#pragma clang fp contract(off)
  double Am[NP], Bm[NP], pwm[NP], pi1[NP], pi2[NP], pw[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    acc[k] = 0.0;
    Am[k] = 1.0, Bm[k] = 0.0;
    pwm[k] = (REG == 2) ? inv_d[k] : 1.0;
  }
  int q = 0;
#pragma unroll
  for (int m = 0; m <= LMAX; ++m) {
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      if (m > 0) {
        const double a2 = fma(ux[k], Am[k], -(uy[k] * Bm[k]));
        Bm[k] = fma(ux[k], Bm[k], uy[k] * Am[k]);
        Am[k] = a2;
        if (REG != 1) pwm[k] *= (REG == 0) ? d[k] : inv_d[k];
      }
      pi2[k] = 0.0, pi1[k] = 1.0, pw[k] = pwm[k];
    }
    // per m, the cos and sin parts are summed over l first and multiplied by Re / Im
    // (x + i y)^m once
    double accc[NP], accs[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) accc[k] = 0.0, accs[k] = 0.0;
#pragma unroll
    for (int l = m; l <= LMAX; ++l) {
      double pil[NP];
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        if (l == m) {
          pil[k] = 1.0;
        } else if (l == m + 1) {
          pil[k] = uz[k];
          pi2[k] = pi1[k];
          pi1[k] = pil[k];
        } else {
          pil[k] = fma(uz[k], pi1[k], -(pi_beta(l, m) * pi2[k]));
          pi2[k] = pi1[k];
          pi1[k] = pil[k];
        }
        if (REG != 1 && l > m) pw[k] *= (REG == 0) ? d[k] : inv_d[k];
      }
#pragma unroll
      for (int cs = 0; cs < (m ? 2 : 1); ++cs) {
        if (REG == 1) {
          const double *cq = c + q * 4;
          const double c0 = cq[0], c1 = cq[1], c2 = cq[2], c3 = cq[3];
#pragma unroll
          for (int k = 0; k < NP; ++k) {
            const double sv = fma(u[k], fma(u[k], fma(u[k], c3, c2), c1), c0);
            if (cs)
              accs[k] = fma(sv, pil[k], accs[k]);
            else
              accc[k] = fma(sv, pil[k], accc[k]);
          }
        } else {
          const double cq = c[q];
#pragma unroll
          for (int k = 0; k < NP; ++k) {
            if (cs)
              accs[k] = fma(cq * pw[k], pil[k], accs[k]);
            else
              accc[k] = fma(cq * pw[k], pil[k], accc[k]);
          }
        }
        ++q;
      }
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      acc[k] = fma(Am[k], accc[k], acc[k]);
      if (m > 0) acc[k] = fma(Bm[k], accs[k], acc[k]);
    }
  }
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

// NP pairs per lane: a wavefront takes NP consecutive 64-pair slots of one bucket (the
// lists are padded to 64 NP per bucket) and reads each staged coefficient once for all.
template <int LMAX, int NP>
__global__ void __launch_bounds__(256, NP == 1 ? 5 : 4) k_synth2n(AtomTables at,
                                                 const double *__restrict__ geo, long N,
                                                 const int *__restrict__ list,
                                                 const int *__restrict__ wave_bucket, long ngroups,
                                                 const double *__restrict__ T,
                                                 const double *__restrict__ ampL,
                                                 const double *__restrict__ ampR,
                                                 double *__restrict__ contrib) {
  constexpr int L = Synth2Cfg<LMAX>::L;
  constexpr int WPB = Synth2Cfg<LMAX>::WPB;
  __shared__ __align__(16) double srow[WPB][L * 4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long gid = (long)blockIdx.x * WPB + wave;
  const bool active = gid < ngroups;
  const int gb = active ? __builtin_amdgcn_readfirstlane(wave_bucket[gid * NP]) : 0;
  const int A = gb / (at.max_nrad + 1), loc = gb - A * (at.max_nrad + 1);
  const int nr = at.nrad[A];
  const int reg = (loc == 0) ? 0 : (loc == nr ? 2 : 1);
  const int j = (reg == 1) ? loc - 1 : 0;
  if (!active) return;
  // per-pair geometry formed once per grid (launch_pair_geometry): u or d, unit vector; its
  // loads and the spline row's are all issued before any of them is waited for
  int p[NP];
  double g0[NP], ux[NP], uy[NP], uz[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const long slot = gid * NP + k;
    p[k] = list[slot * 64 + lane];
    const double *g = geo + slot * 256 + lane;
    g0[k] = g[0], ux[k] = g[64], uy[k] = g[128], uz[k] = g[192];
  }
  // each wave stages its own row into its own LDS slice (16-byte loads, all in flight), so
  // a wavefront-scope sync suffices and the waves of a workgroup start independently
  if (reg == 1) {
    constexpr int NV = L * 2;  // double2 per row
    constexpr int NLD = (NV + 63) / 64;
    const double2 *row = (const double2 *)(at.toff ? T + at.toff[A] + (long)j * L * 4
                                                   : T + ((long)A * (at.max_nrad - 1) + j) * (long)L * 4);
    double2 rv[NLD];
#pragma unroll
    for (int t = 0; t < NLD; ++t) rv[t] = row[min(lane + 64 * t, NV - 1)];
#pragma unroll
    for (int t = 0; t < NLD; ++t)
      if (lane + 64 * t < NV) ((double2 *)srow[wave])[lane + 64 * t] = rv[t];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  double u[NP], d[NP], inv_d[NP], acc[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    u[k] = (reg == 1) ? g0[k] : 0.0;
    d[k] = (reg == 1) ? 0.0 : g0[k];
    inv_d[k] = (reg == 2) ? 1.0 / g0[k] : 0.0;
  }
  const long arow = (long)A * (at.amp_stride ? at.amp_stride : L);
  if (reg == 1)
    synth_harmonics_np<LMAX, 1, NP>(srow[wave], u, ux, uy, uz, d, inv_d, acc);
  else if (reg == 0)
    synth_harmonics_np<LMAX, 0, NP>(ampL + arow, u, ux, uy, uz, d, inv_d, acc);
  else
    synth_harmonics_np<LMAX, 2, NP>(ampR + arow, u, ux, uy, uz, d, inv_d, acc);
  const long cb = at.cbase ? at.cbase[A] : (long)A * N;
#pragma unroll
  for (int k = 0; k < NP; ++k)
    if (p[k] >= 0) contrib[cb + p[k]] = acc[k];
}

template <int LMAX>
void launch_synth2_t(int np, hipStream_t s, const AtomTables &at, const double *geo, long N,
                     const int *list, const int *wave_bucket, long nwaves, const double *T,
                     const double *aL, const double *aR, double *contrib) {
  constexpr int WPB = Synth2Cfg<LMAX>::WPB;
  if (np == 2) {
    const long ng = nwaves / 2;
    k_synth2n<LMAX, 2><<<dim3(div_up(ng, WPB)), dim3(64 * WPB), 0, s>>>(
        at, geo, N, list, wave_bucket, ng, T, aL, aR, contrib);
  } else {
    k_synth2n<LMAX, 1><<<dim3(div_up(nwaves, WPB)), dim3(64 * WPB), 0, s>>>(
        at, geo, N, list, wave_bucket, nwaves, T, aL, aR, contrib);
  }
}

template <int LMAX>
void launch_synth_t(dim3 g, dim3 b, hipStream_t s, const AtomTables &at, const double *x,
                    const double *y, const double *z, long N, const double *T, const double *aL,
                    const double *aR, int acc, int ext_mode, double *v) {
  k_synth<LMAX><<<g, b, 0, s>>>(at, x, y, z, N, T, aL, aR, acc, ext_mode, v);
}

}  // namespace napmo
