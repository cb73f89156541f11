// Synthesis kernel template (instantiated per lmax in synth_inst_*.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace napmo {

// One thread per point, loop over centres. Per (point, centre): one inverse radial
// transform, then the m-major recurrence of the PI_l^m(z) polynomials of the unit
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
    double pimm = 1.0;           // PI_m^m = (2m-1)!!
    double pwm = (reg == 2) ? inv_d : 1.0;  // d^m (left) or d^-(m+1) (right)
#pragma unroll
    for (int m = 0; m <= LMAX; ++m) {
      if (m > 0) {
        const double a2 = ux * Am - uy * Bm;
        Bm = ux * Bm + uy * Am;
        Am = a2;
        pimm *= (2 * m - 1);
        pwm *= base;
      }
      double pi2 = 0.0, pi1 = pimm, pw = pwm;
#pragma unroll
      for (int l = m; l <= LMAX; ++l) {
        double pil;
        if (l == m) {
          pil = pimm;
        } else if (l == m + 1) {
          pil = uz * (2 * m + 1) * pimm;
          pi2 = pi1;
          pi1 = pil;
        } else {
          pil = (uz * (2 * l - 1) * pi1 - (l + m - 1) * pi2) / (l - m);
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


template <int LMAX>
void launch_synth_t(dim3 g, dim3 b, hipStream_t s, const AtomTables &at, const double *x,
                    const double *y, const double *z, long N, const double *T, const double *aL,
                    const double *aR, int acc, int ext_mode, double *v) {
  k_synth<LMAX><<<g, b, 0, s>>>(at, x, y, z, N, T, aL, aR, acc, ext_mode, v);
}

}  // namespace napmo
