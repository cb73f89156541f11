// Device kernels behind two legacy exports that the Python layer binds but does not call on
// its main path: dot_multi_moments (src/horton/dot_multi.cpp:148-186) and
// nwavefunction_compute_density_from_dm (src/wavefunction/nwavefunction.cpp:109-159).
// Both keep the reference's summation order, so with the same inputs they agree with it to
// the last bit (dot_multi_moments) or to rounding of the basis values (density).
#include "legacy_kernels.hpp"

#include <algorithm>

namespace napmo {

#pragma clang fp contract(off)

int moment_count(int lmax, int mtype) {
  if (lmax <= 0) return 1;
  switch (mtype) {
    case 1: return (lmax + 1) * (lmax + 2) * (lmax + 3) / 6;
    case 3: return lmax + 1;
    default: return (lmax + 1) * (lmax + 1);
  }
}

// ----------------------------------------------------------------- moment polynomials
// work[0 .. nmoment-2] of one point as fill_polynomials_wrapper (src/horton/dot_multi.cpp:55-80)
// with fill_{cartesian,pure,radial}_polynomials (src/horton/moments.cpp:24-57, 59-125, 138-155).
__device__ void moment_polynomials(double *work, double dx, double dy, double dz, int lmax,
                                   int mtype) {
  if (mtype == 1) {
    work[0] = dx, work[1] = dy, work[2] = dz;
    int old_off = 0, old_n = 3;
    for (int l = 2; l <= lmax; ++l) {
      const int new_n = old_n + l + 1, new_off = old_off + old_n;
      for (int i = 0; i < old_n; ++i) work[new_off + i] = work[0] * work[old_off + i];
      for (int i = 0; i < l; ++i) work[new_off + old_n + i] = work[1] * work[new_off - l + i];
      work[new_off + new_n - 1] = work[2] * work[new_off - 1];
      old_n = new_n, old_off = new_off;
    }
    return;
  }
  if (mtype == 3) {
    work[0] = sqrt(dx * dx + dy * dy + dz * dz);
    for (int c = 0; c + 1 < lmax; ++c) work[c + 1] = work[0] * work[c];
    return;
  }
  double z = dz, x = dx, y = dy;
  if (mtype == 4) {
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    z = dz / r, x = dx / r, y = dy / r;
  }
  work[0] = z, work[1] = x, work[2] = y;
  if (lmax <= 1) return;
  const double r2 = x * x + y * y + z * z;
  double pold[MOMENT_LMAX + 1], pnew[MOMENT_LMAX + 1], a[MOMENT_LMAX + 1], b[MOMENT_LMAX + 1];
  pold[0] = 1, pnew[0] = z, pnew[1] = 1, a[1] = x, b[1] = y;
  int old_off = 0, old_n = 3;
  for (int l = 2; l <= lmax; ++l) {
    const int new_n = old_n + 2, new_off = old_off + old_n;
    double factor = 2 * l - 1;
    for (int m = 0; m <= l - 2; ++m) {
      const double tmp = pold[m];
      pold[m] = pnew[m];
      pnew[m] = (z * factor * pold[m] - r2 * (l + m - 1) * tmp) / (l - m);
    }
    pold[l - 1] = pnew[l - 1];
    pnew[l] = factor * pold[l - 1];
    pnew[l - 1] = z * pnew[l];
    a[l] = x * a[l - 1] - y * b[l - 1];
    b[l] = x * b[l - 1] + y * a[l - 1];
    work[new_off] = pnew[0];
    factor = 1.4142135623730950488;  // sqrt(2)
    for (int m = 1; m <= l; ++m) {
      factor /= sqrt((double)((l + m) * (l - m + 1)));
      work[new_off + 2 * m - 1] = factor * a[m] * pnew[m];
      work[new_off + 2 * m] = factor * b[m] * pnew[m];
    }
    old_n = new_n, old_off = new_off;
  }
}

// One workgroup per segment. Points are staged CH at a time in LDS as
// term * [1, work...]; thread t then owns moments t, t + 256, ... and adds the staged
// rows in point order, which is the reference's order of accumulation per moment.
__global__ void __launch_bounds__(256) k_dot_multi_moments(
    const long *__restrict__ seg_start, const long *__restrict__ seg_stop,
    const long *__restrict__ seg_index, long npoint, long nvector,
    const double *__restrict__ data, const double *__restrict__ points, double cx, double cy,
    double cz, int lmax, int mtype, int nmoment, int ch, double *__restrict__ output) {
  extern __shared__ double rows[];  // [ch][nmoment]
  const long start = seg_start[blockIdx.x], stop = seg_stop[blockIdx.x];
  double *out = output + seg_index[blockIdx.x] * nmoment;
  double acc[MOMENT_PER_THREAD];
  int nmine = 0;
  for (int m = threadIdx.x; m < nmoment; m += 256) acc[nmine++] = out[m];
  for (long p0 = start; p0 < stop; p0 += ch) {
    const int cnt = (int)min((long)ch, stop - p0);
    if (threadIdx.x < cnt) {
      const long p = p0 + threadIdx.x;
      double term = data[(nvector - 1) * npoint + p];
      for (long v = nvector - 2; v >= 0; --v) term *= data[v * npoint + p];
      double *row = rows + (long)threadIdx.x * nmoment;
      row[0] = term;
      if (lmax > 0) {
        // polynomials built in place in the staged row, then scaled by the integrand
        moment_polynomials(row + 1, points[3 * p] - cx, points[3 * p + 1] - cy,
                           points[3 * p + 2] - cz, lmax, mtype);
        for (int m = 1; m < nmoment; ++m) row[m] = term * row[m];
      }
    }
    __syncthreads();
    int k = 0;
    for (int m = threadIdx.x; m < nmoment; m += 256, ++k) {
      double s = acc[k];
      for (int t = 0; t < cnt; ++t) s += rows[(long)t * nmoment + m];
      acc[k] = s;
    }
    __syncthreads();
  }
  int k = 0;
  for (int m = threadIdx.x; m < nmoment; m += 256) out[m] = acc[k++];
}

void launch_dot_multi_moments(const long *seg_start, const long *seg_stop, const long *seg_index,
                              int nseg, long npoint, long nvector, const double *data,
                              const double *points, const double *center, int lmax, int mtype,
                              int nmoment, double *output, hipStream_t s) {
  NAPMO_REQUIRE(nmoment == moment_count(lmax, mtype) && nmoment <= 256 * MOMENT_PER_THREAD,
                "dot_multi_moments: nmoment does not match lmax and mtype");
  const int ch = (int)std::min<long>(256, std::max<long>(1, 8192 / nmoment));
  const size_t lds = sizeof(double) * (size_t)ch * nmoment;
  k_dot_multi_moments<<<nseg, 256, lds, s>>>(seg_start, seg_stop, seg_index, npoint, nvector,
                                             data, points, center[0], center[1], center[2], lmax,
                                             mtype, nmoment, ch, output);
  NAPMO_HIP(hipGetLastError());
}

// ------------------------------------------------------------ screened density from a DM
// Per point p (Phi point-major [Npad x ld]): absmax_p = max_i |phi_i|,
// upper_p = (sum_i |phi_i| dmmaxrow_i) * (n * absmax_p)  (nwavefunction.cpp:117-128).
__global__ void k_dm_bounds(const double *__restrict__ phi, int ld, long N, int n,
                            const double *__restrict__ dmmaxrow, double *__restrict__ absmax,
                            double *__restrict__ upper) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const double *row = phi + p * ld;
  double amax = 0.0, up = 0.0;
  for (int i = 0; i < n; ++i) amax = fmax(fabs(row[i]), amax);
  for (int i = 0; i < n; ++i) up += fabs(row[i]) * dmmaxrow[i];
  absmax[p] = amax;
  upper[p] = up * (n * amax);
}

// The reference walks the points in order with one running epsilon: it stops at the first
// point whose bound is below it, and otherwise divides it by absmax_p * n * n before the
// point's own screening (nwavefunction.cpp:130-135). This sequential recurrence runs in one
// lane; thr[p] receives the threshold point p screens with and *stop the first point left
// unwritten.
__global__ void k_dm_epsilon_scan(const double *__restrict__ absmax,
                                  const double *__restrict__ upper, long N, int n, double eps,
                                  double *__restrict__ thr, long *__restrict__ stop) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  long p = 0;
  for (; p < N; ++p) {
    // once epsilon is no longer positive (underflow to 0, or NaN from 0 / 0) the reference
    // skips both the bound and the screening for every later point
    if (eps > 0) {
      if (upper[p] < eps) break;
      eps /= absmax[p] * n * n;
    }
    thr[p] = eps;
  }
  stop[0] = p;
}

// rho_p = sum_i [2 sum_{j<i} phi_j D_ij + D_ii phi_i] phi_i over the unscreened i, with the
// inner sum running j = i-1 .. 0 (nwavefunction.cpp:139-155). screen = 0: thresholds unused.
__global__ void k_dm_density(const double *__restrict__ phi, int ld, long N, int n,
                             const double *__restrict__ dm, const double *__restrict__ dmmaxrow,
                             const double *__restrict__ thr, const long *__restrict__ stop,
                             int screen, double *__restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long end = screen ? stop[0] : N;
  if (p >= end) return;
  const double *row = phi + p * ld;
  const double t = screen ? thr[p] : 0.0;
  double rho = 0.0;
  for (int i = 0; i < n; ++i) {
    if (screen && t > 0 && fabs(row[i]) * dmmaxrow[i] < t) continue;
    double tmp = 0.0;
    for (int j = i - 1; j >= 0; --j) tmp += row[j] * dm[(long)i * n + j];
    rho += (2.0 * tmp + dm[(long)i * (n + 1)] * row[i]) * row[i];
  }
  out[p] = rho;
}

void launch_dm_density(const double *phi, int ld, long N, int n, const double *dm,
                       const double *dmmaxrow, double eps, double *absmax, double *upper,
                       double *thr, long *stop, double *out, hipStream_t s) {
  const int screen = eps > 0 ? 1 : 0;
  const unsigned grid = (unsigned)div_up(N, 256);
  if (screen) {
    k_dm_bounds<<<grid, 256, 0, s>>>(phi, ld, N, n, dmmaxrow, absmax, upper);
    k_dm_epsilon_scan<<<1, 64, 0, s>>>(absmax, upper, N, n, eps, thr, stop);
  }
  k_dm_density<<<grid, 256, 0, s>>>(phi, ld, N, n, dm, dmmaxrow, thr, stop, screen, out);
  NAPMO_HIP(hipGetLastError());
}

// ----------------------------------------------------------- basis gradient
// For phi = N x^a y^b z^c sum_k c_k exp(-zeta_k r^2) (origin-shifted coordinates):
//   d phi / dx = N y^b z^c (a x^(a-1) g + x^(a+1) h),  g = sum_k c_k e_k,  h = -2 sum_k c_k zeta_k e_k,
// and likewise for y and z. One thread per (function, point).
__device__ inline double ipow_b(double v, int e) {
  double r = 1.0;
  for (int k = 0; k < e; ++k) r *= v;
  return r;
}

__global__ void k_basis_gradient(const double *__restrict__ x, const double *__restrict__ y,
                                 const double *__restrict__ z, long N, int n,
                                 const double *__restrict__ fo, const int *__restrict__ fl,
                                 const int *__restrict__ ff, const int *__restrict__ fnp,
                                 const double *__restrict__ fnorm, const double *__restrict__ pz,
                                 const double *__restrict__ pc, double *__restrict__ out) {
  const int mu = blockIdx.x * blockDim.x + threadIdx.x;
  const long p = (long)blockIdx.y * blockDim.y + threadIdx.y;
  if (mu >= n || p >= N) return;
  const double dx = x[p] - fo[3 * mu], dy = y[p] - fo[3 * mu + 1], dz = z[p] - fo[3 * mu + 2];
  const double r2 = dx * dx + dy * dy + dz * dz;
  const int a = fl[3 * mu], b = fl[3 * mu + 1], c = fl[3 * mu + 2];
  double g = 0.0, h = 0.0;
  const int f0 = ff[mu], np = fnp[mu];
  for (int k = 0; k < np; ++k) {
    const double e = pc[f0 + k] * exp(-pz[f0 + k] * r2);
    g += e;
    h -= 2.0 * pz[f0 + k] * e;
  }
  const double px = ipow_b(dx, a), py = ipow_b(dy, b), pzv = ipow_b(dz, c);
  const double ax = (a ? a * ipow_b(dx, a - 1) * g : 0.0) + px * dx * h;
  const double ay = (b ? b * ipow_b(dy, b - 1) * g : 0.0) + py * dy * h;
  const double az = (c ? c * ipow_b(dz, c - 1) * g : 0.0) + pzv * dz * h;
  const double nm = fnorm[mu];
  double *o = out + ((size_t)p * n + mu) * 3;
  o[0] = nm * ax * py * pzv;
  o[1] = nm * px * ay * pzv;
  o[2] = nm * px * py * az;
}

void launch_basis_gradient(const double *x, const double *y, const double *z, long N, int n,
                           const double *fo, const int *fl, const int *ff, const int *fnp,
                           const double *fnorm, const double *pz, const double *pc, double *out,
                           hipStream_t s) {
  if (N <= 0 || n <= 0) return;
  dim3 block(64, 4);
  dim3 grid(div_up(n, 64), div_up(N, 4));
  k_basis_gradient<<<grid, block, 0, s>>>(x, y, z, N, n, fo, fl, ff, fnp, fnorm, pz, pc, out);
  NAPMO_HIP(hipGetLastError());
}

}  // namespace napmo
