// Becke-Dickson orbital correction on the GPU (napmo/wavefunction/ndpsi.py:15-89 and the
// per-orbital steps of napmo/wavefunction/psi_optimization.py:175-203).
//
// Per orbital o, atom A and angular momentum l the correction solves, on A's radial grid,
//   u'' + (2/r) u' + a u = c mu_lm,  a = c (l(l+1)/r^2 + Vsph - oi - doi),  c = -2/m_inv,
// with u(r_0) = u(r_max) = 0, for the 2l+1 projected sources mu_lm of (phi doi - ri) bw.
// a depends on the orbital, so unlike the Poisson solve the matrix cannot be factored once
// per geometry: one wavefront per (A, l, o) assembles the Hermite finite-element matrix
// from per-grid tables, factors it as a band matrix with partial pivoting and solves the
// 2l+1 right-hand sides, one lane per right-hand side.
#include "common.hpp"
#include "kernels.hpp"

namespace napmo {

namespace {

// out[o][p] = a[o][p] s[o] - b[o][p]
__global__ void k_scaled_diff(const double *__restrict__ a, const double *__restrict__ s,
                              const double *__restrict__ b, long N, double *__restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int o = blockIdx.y;
  if (p < N) out[(long)o * N + p] = a[(long)o * N + p] * s[o] - b[(long)o * N + p];
}

// One block per (shell i, atom A, orbital o): sum_j (f bw) w / sum_j w.
__global__ void __launch_bounds__(64) k_shell_average(AtomTables at, const double *__restrict__ V,
                                                      long N, const double *__restrict__ w,
                                                      const double *__restrict__ bw,
                                                      long nrad_total, double *__restrict__ Vsph) {
  const int i = blockIdx.x, A = blockIdx.y, o = blockIdx.z;
  if (i >= at.nrad[A]) return;
  const int na = at.nang[A];
  const long p0 = at.point_off[A] + (long)i * na;
  const double *f = V + (long)o * N;
  double num = 0.0, den = 0.0;
  for (int j = threadIdx.x; j < na; j += 64) {
    const long p = p0 + j;
    num += (f[p] * bw[p]) * w[p];
    den += w[p];
  }
  for (int k = 32; k > 0; k >>= 1) {
    num += __shfl_xor(num, k);
    den += __shfl_xor(den, k);
  }
  if (threadIdx.x == 0) Vsph[(long)o * nrad_total + at.rad_off[A] + i] = num / den;
}

// Slopes of the C2 cubic spline through y (the (2, 4, ..., 4, 2 | 1) system of
// src/horton/cubic_spline.cpp:54-103, same elimination order); up, rhs: n scratch each.
__device__ void spline_slopes(const double *y, double *dt, int n, double *up, double *rhs) {
  for (int i = 0; i < n - 1; ++i) up[i] = 1.0;
  rhs[0] = 3.0 * (y[1] - y[0]);
  for (int i = 1; i < n - 1; ++i) rhs[i] = 3.0 * (y[i + 1] - y[i - 1]);
  rhs[n - 1] = 3.0 * (y[n - 1] - y[n - 2]);
  double tmp1 = up[0], tmp2 = 0.0;
  up[0] /= 2.0;
  rhs[0] /= 2.0;
  for (int i = 1; i < n; ++i) {
    const double den = (i == n - 1 ? 2.0 : 4.0) - up[i - 1] * tmp1;
    if (i < n - 1) {
      tmp2 = up[i];
      up[i] /= den;
    }
    rhs[i] = (rhs[i] - rhs[i - 1] * tmp1) / den;
    tmp1 = tmp2;
  }
  dt[n - 1] = rhs[n - 1];
  for (int i = n - 2; i >= 0; --i) dt[i] = rhs[i] - up[i] * dt[i + 1];
}

// Band storage: row k holds columns k - 3 .. k + 6 in slots 0 .. 9 (slot = col - row + 3);
// after pivoting the upper band grows to k + 6. L multipliers of column k in Lf[k][0..2].
constexpr int WSLOT = 10;

__global__ void __launch_bounds__(64) k_helmholtz_radial(
    AtomTables at, HelmholtzTables ht, int L, long nrad_total, const double *__restrict__ Vsph,
    const double *__restrict__ oi, const double *__restrict__ doi, double mass_inv,
    const double *__restrict__ mu, int mc, double *__restrict__ Uy, double *__restrict__ Udt,
    int *__restrict__ status) {
  extern __shared__ double sm[];
  const int A = blockIdx.x, l = blockIdx.y, o = blockIdx.z;
  const int lane = threadIdx.x;
  const int nr = at.nrad[A], nfn = 2 * nr, nm = 2 * l + 1;
  const int nfn_max = 2 * at.max_nrad;
  double *W = sm;                          // [nfn][WSLOT]
  double *Lf = W + nfn_max * WSLOT;        // [nfn][3]
  double *sinv = Lf + nfn_max * 3;         // [nr] 1 / pivots of the spline system
  double *supm = sinv + at.max_nrad;       // [nr] its eliminated super-diagonal
  int *ipiv = (int *)(supm + at.max_nrad); // [nfn]
  double *X = (double *)(ipiv + nfn_max);  // [nfn][mc] columns; ay, ad, an, scratch first
  double *ay = X, *ad = X + nr, *an = X + nfn, *scr = X + 2 * nfn;
  const long r0 = at.rad_off[A];
  const double c = -2.0 / mass_inv;
  const double e_o = oi[o], d_o = doi[o];
  const double *rr = at.r + r0, *j1 = ht.j1 + r0, *j2 = ht.j2 + r0;

  // a = c (l(l+1)/r^2 + Vsph - oi - doi), its spline slopes, then the t-axis transform
  // a_y j1^2, (a_x j1^2 + 2 a_y j2) j1 (napmo/utilities/ode2.py:35-118)
  for (int i = lane; i < nr; i += 64) {
    const double r2 = rr[i] * rr[i];
    const double r2inv = 1.0 / r2;
    ay[i] = c * ((l * (l + 1.0) * r2inv) + (Vsph[(long)o * nrad_total + r0 + i] - e_o - d_o));
  }
  __syncthreads();
  if (lane == 0) spline_slopes(ay, ad, nr, scr, scr + nr);
  __syncthreads();
  for (int i = lane; i < nr; i += 64) {
    const double j1sq = j1[i] * j1[i];
    const double ax = ad[i] / j1[i];
    an[2 * i] = ay[i] * j1sq;
    an[2 * i + 1] = (ax * j1sq + 2 * ay[i] * j2[i]) * j1[i];
  }
  __syncthreads();

  // Galerkin matrix: c0 + sum_k a_k Oa[k] on the band; rows 0 and nfn - 2 are u = 0
  const double *C0 = ht.c0 + ht.c0_off[A], *OA = ht.oa + ht.oa_off[A];
  for (int row = lane; row < nfn; row += 64) {
    double *wr = W + row * WSLOT;
    for (int s = 0; s < WSLOT; ++s) wr[s] = 0.0;
    if (row == 0 || row == nfn - 2) {
      wr[3] = 1.0;
      continue;
    }
    const int begin = max(2 * (row / 2 - 1), 0), end = min(2 * (row / 2 + 2), nfn);
    for (int off = -3; off <= 3; ++off) {
      const int col = row + off;
      if (col < begin || col >= end) continue;
      double v = C0[row * 7 + off + 3];
      const double *oa = OA + (row * 7 + off + 3) * 6;
      for (int kk = 0; kk < 6 && begin + kk < end; ++kk) v += an[begin + kk] * oa[kk];
      wr[off + 3] = v;
    }
  }
  // Elimination factors of the spline-slope system (2, 4, ..., 4, 2 | 1), shared by every
  // column: r_i = (rhs_i - r_{i-1}) / den_i forward, dt_i = r_i - upm_i dt_{i+1} back
  // (the order of src/horton/cubic_spline.cpp:54-79 with unit off-diagonals)
  if (lane == 0) {
    double upm = 0.5;
    sinv[0] = 0.5;
    supm[0] = upm;
    for (int i = 1; i < nr; ++i) {
      const double den = (i == nr - 1 ? 2.0 : 4.0) - upm;
      sinv[i] = 1.0 / den;
      upm = (i < nr - 1) ? 1.0 / den : 0.0;
      supm[i] = upm;
    }
  }
  __syncthreads();

  // band LU with partial pivoting (pivot among rows k .. k + 3)
  for (int k = 0; k < nfn; ++k) {
    int p = k;
    double best = fabs(W[k * WSLOT + 3]);
    for (int d = 1; d <= 3 && k + d < nfn; ++d) {
      const double v = fabs(W[(k + d) * WSLOT + 3 - d]);
      if (v > best) best = v, p = k + d;
    }
    if (lane == 0) {
      ipiv[k] = p;
      if (best == 0.0) status[0] = 1;
    }
    if (p != k && lane < 7) {
      const int j = k + lane;
      if (j < nfn) {
        double *a = W + k * WSLOT + lane + 3, *b = W + p * WSLOT + (j - p + 3);
        const double t = *a;
        *a = *b;
        *b = t;
      }
    }
    __syncthreads();
    const double piv = W[k * WSLOT + 3];
    if (lane < 18) {
      const int d = 1 + lane / 6, jj = 1 + lane % 6, i = k + d, j = k + jj;
      if (i < nfn && j < nfn) {
        const double f = W[i * WSLOT + 3 - d] / piv;
        W[i * WSLOT + j - i + 3] -= f * W[k * WSLOT + jj + 3];
      }
    } else if (lane < 21) {
      const int d = lane - 17;
      Lf[k * 3 + d - 1] = (k + d < nfn) ? W[(k + d) * WSLOT + 3 - d] / piv : 0.0;
    }
    __syncthreads();
  }

  // right-hand sides in chunks of mc columns; lane q owns column q of X, which holds in
  // turn the merged source f (value at even rows, slope at odd rows), the Galerkin
  // right-hand side, and the forward-substituted vector
  const double *O2 = ht.o2 + ht.o2_off[A];
  const long ubase = ((long)o * nrad_total + r0) * L + l * l;
  for (int m0 = 0; m0 < nm; m0 += mc) {
    const int cnt = min(mc, nm - m0);
    if (lane < cnt) {
      const int q = lane;
      double *x = X + q;  // x[k * mc]: row k of this column
      // mu and the slopes of its spline on the t-axis
      for (int i = 0; i < nr; ++i) x[2 * i * mc] = mu[ubase + (long)i * L + m0 + q];
      double rprev = 0.0;
#pragma unroll 4
      for (int i = 0; i < nr; ++i) {
        const double ya = x[2 * (i == 0 ? 0 : i - 1) * mc];
        const double yb = x[2 * (i == nr - 1 ? nr - 1 : i + 1) * mc];
        const double rhs = 3.0 * (yb - ya);
        rprev = (i == 0) ? rhs * sinv[0] : (rhs - rprev) * sinv[i];
        x[(2 * i + 1) * mc] = rprev;
      }
      double dtn = rprev;
#pragma unroll 4
      for (int i = nr - 2; i >= 0; --i) {
        dtn = x[(2 * i + 1) * mc] - supm[i] * dtn;
        x[(2 * i + 1) * mc] = dtn;
      }
      // f = c (mu, mu_x) on the t-axis, (fy j1^2, (fx j1^2 + 2 fy j2) j1); the right-hand
      // side of rows 2i, 2i+1 is the sum of f over nodes i-1, i, i+1 times the node
      // integrals O2 (zero rows for the value conditions), written in place through a
      // register window of the three nodes
      auto fnode = [&](int i, double &f0, double &f1) {
        const double j1i = j1[i], j1sq = j1i * j1i;
        const double fy = c * x[2 * i * mc], fx = c * (x[(2 * i + 1) * mc] / j1i);
        f0 = fy * j1sq;
        f1 = (fx * j1sq + 2 * fy * j2[i]) * j1i;
      };
      double fa0 = 0.0, fa1 = 0.0, fb0, fb1, fc0 = 0.0, fc1 = 0.0;
      fnode(0, fb0, fb1);
#pragma unroll 2
      for (int i = 0; i < nr; ++i) {
        if (i + 1 < nr)
          fnode(i + 1, fc0, fc1);
        else
          fc0 = 0.0, fc1 = 0.0;
        for (int r = 0; r < 2; ++r) {
          const double *w6 = O2 + (2 * i + r) * 6;
          x[(2 * i + r) * mc] = fa0 * w6[0] + fa1 * w6[1] + fb0 * w6[2] + fb1 * w6[3] +
                                fc0 * w6[4] + fc1 * w6[5];
        }
        fa0 = fb0, fa1 = fb1, fb0 = fc0, fb1 = fc1;
      }
      // forward: row interchanges and L, window w = x[k .. k + 3]
      double w0 = x[0], w1 = x[mc], w2 = x[2 * mc], w3 = (3 < nfn) ? x[3 * mc] : 0.0;
#pragma unroll 4
      for (int k = 0; k < nfn; ++k) {
        const int d = ipiv[k] - k;
        if (d == 1) {
          const double t = w0; w0 = w1; w1 = t;
        } else if (d == 2) {
          const double t = w0; w0 = w2; w2 = t;
        } else if (d == 3) {
          const double t = w0; w0 = w3; w3 = t;
        }
        const double xk = w0;
        x[k * mc] = xk;
        w0 = w1 - Lf[k * 3 + 0] * xk;
        w1 = w2 - Lf[k * 3 + 1] * xk;
        w2 = w3 - Lf[k * 3 + 2] * xk;
        w3 = (k + 4 < nfn) ? x[(k + 4) * mc] : 0.0;
      }
      // backward: U, window u = x[k + 1 .. k + 6]; the solution's value y = u at even
      // rows, slope dt = u at odd rows (written as (u / j1) j1, the reference's dx round trip)
      double u1 = 0.0, u2 = 0.0, u3 = 0.0, u4 = 0.0, u5 = 0.0, u6 = 0.0;
#pragma unroll 4
      for (int k = nfn - 1; k >= 0; --k) {
        const double *wr = W + k * WSLOT;
        const double s = x[k * mc] - wr[4] * u1 - wr[5] * u2 - wr[6] * u3 - wr[7] * u4 -
                         wr[8] * u5 - wr[9] * u6;
        const double xs = s / wr[3];
        const int i = k >> 1;
        if (k & 1)
          Udt[ubase + (long)i * L + m0 + q] = (xs / j1[i]) * j1[i];
        else
          Uy[ubase + (long)i * L + m0 + q] = xs;
        u6 = u5, u5 = u4, u4 = u3, u3 = u2, u2 = u1, u1 = xs;
      }
    }
    __syncthreads();
  }
}

// R = T + V - e psi; per-block partial sums of R psi w bw, then a fixed-order final sum.
__global__ void __launch_bounds__(256) k_residual(const double *__restrict__ psi,
                                                  const double *__restrict__ T,
                                                  const double *__restrict__ V,
                                                  const double *__restrict__ e,
                                                  const double *__restrict__ wbw, long N,
                                                  double *__restrict__ R,
                                                  double *__restrict__ work) {
  __shared__ double red[4];
  const int o = blockIdx.y;
  const long base = (long)o * N;
  const double eo = e[o];
  double acc = 0.0;
  for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < N; p += (long)gridDim.x * 256) {
    const double r = T[base + p] + V[base + p] - eo * psi[base + p];
    R[base + p] = r;
    acc += (r * psi[base + p]) * wbw[p];
  }
  for (int k = 32; k > 0; k >>= 1) acc += __shfl_xor(acc, k);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) work[(long)o * gridDim.x + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ void k_residual_final(const double *__restrict__ work, int nblk, int norb,
                                 double *__restrict__ de) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= norb) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += work[(long)o * nblk + b];
  de[o] = s;
}

}  // namespace

void launch_scaled_diff(const double *a, const double *s, const double *b, int norb, long N,
                        double *out, hipStream_t st) {
  if (norb <= 0 || N <= 0) return;
  k_scaled_diff<<<dim3((unsigned)div_up(N, 256), norb), 256, 0, st>>>(a, s, b, N, out);
  NAPMO_HIP(hipGetLastError());
}

void launch_shell_average(const AtomTables &at, const double *V, int norb, long N, const double *w,
                          const double *bw, long nrad_total, double *Vsph, hipStream_t st) {
  if (norb <= 0) return;
  k_shell_average<<<dim3(at.max_nrad, at.M, norb), 64, 0, st>>>(at, V, N, w, bw, nrad_total, Vsph);
  NAPMO_HIP(hipGetLastError());
}

void launch_helmholtz_radial(const AtomTables &at, const HelmholtzTables &ht, int lmax, int norb,
                             long nrad_total, const double *Vsph, const double *oi,
                             const double *doi, double mass_inv, const double *mu, double *Uy,
                             double *Udt, int *status, hipStream_t st) {
  if (norb <= 0) return;
  NAPMO_REQUIRE(at.max_nrad <= HELMHOLTZ_MAX_NRAD,
                "orbital correction: radial grids of more than " +
                    std::to_string(HELMHOLTZ_MAX_NRAD) + " points are not supported");
  NAPMO_REQUIRE(at.max_nrad >= 3, "orbital correction: radial grids need at least 3 points");
  const int L = (lmax + 1) * (lmax + 1);
  const long nr = at.max_nrad, nfn = 2L * nr;
  // W, Lf, spline factors, ipiv, then nfn x mc columns (at least 3 nfn for the assembly);
  // mc sized so the block stays near 40 KB and several blocks share a CU
  const long fixed = nfn * (WSLOT + 3) + 2 * nr + nfn / 2;  // doubles (ipiv: nfn ints)
  const long target = 40 * 1024 / 8;
  const long mc = std::max<long>(1, std::min<long>((target - fixed) / nfn, 2L * lmax + 1));
  const size_t bytes = (size_t)(fixed + nfn * std::max<long>(mc, 3)) * sizeof(double);
  NAPMO_REQUIRE(bytes <= max_lds_per_block(), "orbital correction: radial system does not fit in LDS");
  // past 64 KB (about 230 radial points) the kernel needs the device's full LDS per block
  static int configured = -1;
  int dev = 0;
  NAPMO_HIP(hipGetDevice(&dev));
  if (configured != dev) {
    NAPMO_HIP(hipFuncSetAttribute((const void *)k_helmholtz_radial, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)max_lds_per_block()));
    configured = dev;
  }
  k_helmholtz_radial<<<dim3(at.M, lmax + 1, norb), 64, bytes, st>>>(
      at, ht, L, nrad_total, Vsph, oi, doi, mass_inv, mu, (int)mc, Uy, Udt, status);
  NAPMO_HIP(hipGetLastError());
}

void launch_residual(const double *psi, const double *T, const double *V, const double *e,
                     const double *wbw, int norb, long N, double *R, double *de, double *work,
                     hipStream_t st) {
  if (norb <= 0) return;
  const int nblk = (int)std::max<long>(1, std::min<long>(RESID_BLOCKS, div_up(N, 256)));
  k_residual<<<dim3(nblk, norb), 256, 0, st>>>(psi, T, V, e, wbw, N, R, work);
  k_residual_final<<<div_up(norb, 64), 64, 0, st>>>(work, nblk, norb, de);
  NAPMO_HIP(hipGetLastError());
}

}  // namespace napmo
