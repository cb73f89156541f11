// Becke grid device mirror and Poisson plans (host-side setup in FP64).
#include "poisson_plan.hpp"

#include <algorithm>
#include <cmath>
#include <string>
#include <cstring>

namespace napmo {

// ------------------------------------------------------ Hermite element integrals
// Cubic Hermite shape functions on a unit segment u in [0, 1], integer coefficients of
// 1, u, u^2, u^3: node at the segment's left end (value, slope), right end (value, slope).
static const int SHAPE[2][2][4] = {{{1, 0, -3, 2}, {0, 1, -2, 1}},   // left node
                                   {{0, 0, 3, -2}, {0, 0, -1, 1}}};  // right node

// Integral over [0,1] of a product of shape functions (or their u-derivatives), exact in
// rationals (denominator 2520 = lcm(1..10) covers degree <= 9) and rounded once.
static double seg_integral(const int (*f)[3], int nf) {
  long long poly[10] = {1};
  int deg = 0;
  for (int k = 0; k < nf; ++k) {
    const int node = f[k][0], kind = f[k][1], der = f[k][2];
    long long q[4] = {0, 0, 0, 0};
    int qd = 3;
    for (int c = 0; c < 4; ++c) q[c] = SHAPE[node][kind][c];
    if (der) {
      for (int c = 0; c < 3; ++c) q[c] = (c + 1) * q[c + 1];
      q[3] = 0;
      qd = 2;
    }
    long long out[10] = {0};
    for (int a = 0; a <= deg; ++a)
      for (int b = 0; b <= qd; ++b) out[a + b] += poly[a] * q[b];
    deg += qd;
    std::memcpy(poly, out, sizeof(out));
  }
  long long num = 0;
  for (int c = 0; c <= deg; ++c) num += poly[c] * (2520 / (c + 1));
  return (double)num / 2520.0;
}

// Overlap of basis functions (index i = 2*node + kind, derivative flag) summed over the
// segments around the first function's node inside [0, xmax]: left segment first.
static double fe_overlap(long xmax, const long *idx, const int *der, int nf) {
  const long c0 = idx[0] / 2;
  for (int k = 1; k < nf; ++k)
    if (std::labs(idx[k] / 2 - c0) >= 2) return 0.0;
  double res = 0.0;
  for (int side = 0; side < 2; ++side) {
    const long seg = side == 0 ? c0 - 1 : c0;
    if (side == 0 && !(c0 >= 1 && c0 <= xmax)) continue;
    if (side == 1 && !(c0 >= 0 && c0 <= xmax - 1)) continue;
    int f[3][3];
    bool zero = false;
    for (int k = 0; k < nf; ++k) {
      const long rel = idx[k] / 2 - seg;
      if (rel < 0 || rel > 1) {
        zero = true;
        break;
      }
      f[k][0] = (int)rel, f[k][1] = (int)(idx[k] % 2), f[k][2] = der[k];
    }
    if (!zero) res += seg_integral(f, nf);
  }
  return res;
}

static double fe_node(long x, long i, int deriv) {
  return ((i % 2) == deriv) ? (double)(x == i / 2) : 0.0;
}

// Dense LU with partial pivoting, solving A X = B in place (A n x n, B n x m).
static void lu_solve(std::vector<double> &A, std::vector<double> &B, int n, int m) {
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[(size_t)i * n + k]) > std::fabs(A[(size_t)piv * n + k])) piv = i;
    NAPMO_REQUIRE(A[(size_t)piv * n + k] != 0.0 && std::isfinite(A[(size_t)piv * n + k]),
                  "singular radial Poisson system (zero or non-finite pivot)");
    if (piv != k) {
      for (int c = 0; c < n; ++c) std::swap(A[(size_t)k * n + c], A[(size_t)piv * n + c]);
      for (int c = 0; c < m; ++c) std::swap(B[(size_t)k * m + c], B[(size_t)piv * m + c]);
    }
    const double inv = 1.0 / A[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double f = A[(size_t)i * n + k] * inv;
      if (f == 0.0) continue;
      A[(size_t)i * n + k] = f;
      for (int c = k + 1; c < n; ++c) A[(size_t)i * n + c] -= f * A[(size_t)k * n + c];
      for (int c = 0; c < m; ++c) B[(size_t)i * m + c] -= f * B[(size_t)k * m + c];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    for (int c = 0; c < m; ++c) {
      double s = B[(size_t)k * m + c];
      for (int j = k + 1; j < n; ++j) s -= A[(size_t)k * n + j] * B[(size_t)j * m + c];
      B[(size_t)k * m + c] = s / A[(size_t)k * n + k];
    }
  }
}

void fe_assemble(const double *b, const double *a, const double *f, const double *bcs,
                 double *coeffs, double *rhs, long npoint) {
  const long nfn = 2 * npoint, xmax = npoint - 1;
  long skip[2] = {-1, -1};
  int cnt = 0;
  const long rows[4] = {0, 1, nfn - 2, nfn - 1};
  for (int k = 0; k < 4; ++k)
    if (bcs[k] == bcs[k]) {  // NaN marks a free end
      coeffs[rows[k] * nfn + rows[k]] = 1;
      rhs[rows[k]] = bcs[k];
      if (cnt < 2) skip[cnt] = rows[k];
      ++cnt;
    }
  for (long irow = 0; irow < nfn; ++irow) {
    if (irow == skip[0] || irow == skip[1]) continue;
    const long begin = std::max(2 * (irow / 2 - 1), 0L), end = std::min(2 * (irow / 2 + 2), nfn);
    for (long icol = begin; icol < end; ++icol) {
      double c = coeffs[irow * nfn + icol];
      {
        long idx[2] = {irow, icol};
        int der[2] = {1, 1};
        c -= fe_overlap(xmax, idx, der, 2);
      }
      for (long k = begin; k < end; ++k) {
        long idx[3] = {irow, icol, k};
        if (b[k] != 0) {
          int der[3] = {0, 1, 0};
          c += b[k] * fe_overlap(xmax, idx, der, 3);
        }
        if (a[k] != 0) {
          int der[3] = {0, 0, 0};
          c += a[k] * fe_overlap(xmax, idx, der, 3);
        }
      }
      c += fe_node(xmax, irow, 0) * fe_node(xmax, icol, 1);
      c -= fe_node(0, irow, 0) * fe_node(0, icol, 1);
      coeffs[irow * nfn + icol] = c;
    }
    double acc = rhs[irow];
    for (long k = begin; k < end; ++k) {
      if (f[k] == 0) continue;
      long idx[2] = {irow, k};
      int der[2] = {0, 0};
      acc += f[k] * fe_overlap(xmax, idx, der, 2);
    }
    rhs[irow] = acc;
  }
}

// Right-hand-side map of build_ode2 for f = fscale (rho, rho_x) with rho_x the slopes of
// rho's cubic spline (dt / j1), transformed to the t-axis (fy j1^2, (fx j1^2 + 2 fy j2) j1):
// B [2n x n], column jc = unit rho_jc; rows 0 and 2n-2 (the value conditions) left zero.
static void fe_rhs_map(int n, const std::vector<double> &j1, const std::vector<double> &j2,
                       double fscale, std::vector<double> &B) {
  const long nfn = 2L * n, xmax = n - 1;
  B.assign((size_t)nfn * n, 0.0);
  std::vector<double> e(n), dt(n), f(nfn);
  for (int jc = 0; jc < n; ++jc) {
    std::fill(e.begin(), e.end(), 0.0);
    e[jc] = 1.0;
    spline_derivatives(e.data(), dt.data(), n);
    for (int i = 0; i < n; ++i) {
      const double fy = fscale * e[i], fdx = fscale * (dt[i] / j1[i]), j1sq = j1[i] * j1[i];
      f[2 * i] = fy * j1sq;
      f[2 * i + 1] = (fdx * j1sq + 2 * fy * j2[i]) * j1[i];
    }
    for (long irow = 0; irow < nfn; ++irow) {
      if (irow == 0 || irow == nfn - 2) continue;
      long begin = std::max(2 * (irow / 2 - 1), 0L), end = std::min(2 * (irow / 2 + 2), nfn);
      double acc = 0.0;
      for (long k = begin; k < end; ++k) {
        if (f[k] == 0) continue;
        long idx[2] = {irow, k};
        int der[2] = {0, 0};
        acc += f[k] * fe_overlap(xmax, idx, der, 2);
      }
      B[(size_t)irow * n + jc] = acc;
    }
  }
}

// r(t) and its t-derivatives at the nodes; b = 2/r (b_x = -2/r^2) merged [value, slope]
// on the t-axis (napmo/utilities/ode2.py:75-90).
static void radial_coeffs(const RTransform &rtf, std::vector<double> &r, std::vector<double> &j1,
                          std::vector<double> &j2, std::vector<double> &j3, std::vector<double> &b) {
  const int n = rtf.npoint;
  r.resize(n), j1.resize(n), j2.resize(n), j3.resize(n), b.resize(2 * n);
  for (int i = 0; i < n; ++i) {
    r[i] = rtf.radius(i);
    j1[i] = rtf.deriv(i);
    j2[i] = rtf.deriv2(i);
    j3[i] = rtf.deriv3(i);
  }
  // the radial ODE divides by r and by dr/dt: a node at r = 0 (the identity transform's
  // first point, a zero Chebyshev radius) or a non-increasing grid makes the system singular
  for (int i = 0; i < n; ++i)
    NAPMO_REQUIRE(std::isfinite(r[i]) && r[i] > 0.0 && std::isfinite(j1[i]) && j1[i] > 0.0 &&
                      (i == 0 || r[i] > r[i - 1]),
                  "singular radial Poisson system: radial point " + std::to_string(i) + " at r = " +
                      std::to_string(r[i]) + " (the points must be positive and strictly increasing)");
  for (int i = 0; i < n; ++i) {
    const double r2 = r[i] * r[i], j1sq = j1[i] * j1[i];
    const double by = 2 / r[i], bd = -2 / r2;
    b[2 * i] = j1[i] * by - j2[i] / j1[i];
    b[2 * i + 1] = j2[i] * by + j1sq * bd + (j2[i] * j2[i] - j1[i] * j3[i]) / j1sq;
  }
}

void radial_green(const RTransform &rtf, const std::vector<double> &rw, int l,
                  std::vector<double> &Gy, std::vector<double> &Gd) {
  const int n = rtf.npoint, nfn = 2 * n;
  std::vector<double> r, j1, j2, j3, b;
  radial_coeffs(rtf, r, j1, j2, j3, b);
  // a = -l(l+1)/r^2 (dx 2l(l+1)/r^3) transformed to the t-axis, merged [value, slope]
  const double l2l = (double)l * (l + 1);
  std::vector<double> a(nfn);
  for (int i = 0; i < n; ++i) {
    const double r2 = r[i] * r[i], r3 = r2 * r[i], j1sq = j1[i] * j1[i];
    const double ay = -l2l / r2, ad = 2 * l2l / r3;
    a[2 * i] = ay * j1sq;
    a[2 * i + 1] = (ad * j1sq + 2 * ay * j2[i]) * j1[i];
  }
  // Galerkin matrix; rows 0 and nfn-2 are the value boundary conditions at rmin, rmax.
  std::vector<double> A((size_t)nfn * nfn, 0.0), zf(nfn, 0.0), zr(nfn, 0.0);
  const double bcs[4] = {0.0, NAN, 0.0, NAN};
  fe_assemble(b.data(), a.data(), zf.data(), bcs, A.data(), zr.data(), n);
  // B: rho_j -> right-hand side, f = -4 pi (rho, rho_x); then the boundary values
  // (napmo/grids/poisson_solver.py:92-105)
  std::vector<double> B;
  fe_rhs_map(n, j1, j2, -4.0 * M_PI, B);
  for (int jc = 0; jc < n; ++jc) {
    const double e = 1.0;
    B[jc] = (e * std::pow(r[jc], -l - 1.0) * rw[jc]) * (std::pow(r[0], (double)l) / (2 * l + 1));
    B[(size_t)(nfn - 2) * n + jc] =
        (e * std::pow(r[jc], (double)l) * rw[jc]) * (1.0 / std::pow(r[n - 1], l + 1.0) / (2 * l + 1));
  }
  lu_solve(A, B, nfn, n);
  Gy.assign((size_t)n * n, 0.0);
  Gd.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int jc = 0; jc < n; ++jc) {
      Gy[(size_t)i * n + jc] = B[(size_t)(2 * i) * n + jc];
      Gd[(size_t)i * n + jc] = B[(size_t)(2 * i + 1) * n + jc];
    }
}

void helmholtz_tables(const RTransform &rtf, std::vector<double> &c0, std::vector<double> &oa,
                      std::vector<double> &o2, std::vector<double> &j1, std::vector<double> &j2) {
  const int n = rtf.npoint, nfn = 2 * n;
  const long xmax = n - 1;
  std::vector<double> r, j3, b;
  radial_coeffs(rtf, r, j1, j2, j3, b);
  std::vector<double> A((size_t)nfn * nfn, 0.0), za(nfn, 0.0), zr(nfn, 0.0);
  const double bcs[4] = {0.0, NAN, 0.0, NAN};
  fe_assemble(b.data(), za.data(), za.data(), bcs, A.data(), zr.data(), n);
  c0.assign((size_t)nfn * 7, 0.0);
  oa.assign((size_t)nfn * 7 * 6, 0.0);
  o2.assign((size_t)nfn * 6, 0.0);
  for (long row = 0; row < nfn; ++row) {
    const bool cond = row == 0 || row == nfn - 2;
    const long begin = std::max(2 * (row / 2 - 1), 0L), end = std::min(2 * (row / 2 + 2), (long)nfn);
    // right-hand side integrals of f against the row's test function, f_k for
    // k = 2 (row / 2 - 1) + kk (zero where k falls off the grid)
    for (int kk = 0; kk < 6 && !cond; ++kk) {
      const long k = 2 * (row / 2 - 1) + kk;
      if (k < 0 || k >= nfn) continue;
      long idx[2] = {row, k};
      int der[2] = {0, 0};
      o2[(size_t)row * 6 + kk] = fe_overlap(xmax, idx, der, 2);
    }
    for (int off = -3; off <= 3; ++off) {
      const long col = row + off;
      if (col < 0 || col >= nfn) continue;
      c0[(size_t)row * 7 + off + 3] = A[(size_t)row * nfn + col];
      if (cond || col < begin || col >= end) continue;
      for (long k = begin; k < end; ++k) {
        long idx[3] = {row, col, k};
        int der[3] = {0, 0, 0};
        oa[((size_t)row * 7 + off + 3) * 6 + (k - begin)] = fe_overlap(xmax, idx, der, 3);
      }
    }
  }
}

void kinetic_green(const RTransform &rtf, double radii, const std::vector<double> &r, int l,
                   std::vector<double> &Gy, std::vector<double> &Gd) {
  const int n = rtf.npoint;
  NAPMO_REQUIRE(rtf.kind == 0, "kinetic operator: the radial grids must be Chebyshev");
  Gy.assign((size_t)n * n, 0.0);
  Gd.assign((size_t)n * n, 0.0);
  std::vector<double> e(n), y(n), dt(n);
  const double l2l = (double)l * (l + 1);
  for (int jc = 0; jc < n; ++jc) {
    std::fill(e.begin(), e.end(), 0.0);
    e[jc] = 1.0;
    radial_deriv2(radii, r.data(), e.data(), y.data(), n);
    y[jc] -= l2l / (r[jc] * r[jc]);
    spline_derivatives(y.data(), dt.data(), n);
    for (int i = 0; i < n; ++i) {
      Gy[(size_t)i * n + jc] = y[i];
      Gd[(size_t)i * n + jc] = dt[i];
    }
  }
}

std::vector<double> projection_table(const AngularGrid &ang, int lmax) {
  const int L = (lmax + 1) * (lmax + 1), na = ang.lorder;
  std::vector<double> t((size_t)na * L), s(L);
  double sw = 0.0;
  for (int j = 0; j < na; ++j) sw += ang.weights[j];
  for (int j = 0; j < na; ++j) {
    solid_harmonics(ang.points[3 * j], ang.points[3 * j + 1], ang.points[3 * j + 2], lmax, s.data());
    int lm = 0;
    for (int l = 0; l <= lmax; ++l) {
      const double nrm = std::sqrt(4.0 * M_PI * (2.0 * l + 1));
      for (int m = -l; m <= l; ++m, ++lm) t[(size_t)j * L + lm] = ang.weights[j] * s[lm] * nrm / sw;
    }
  }
  return t;
}

// ------------------------------------------------------------------ DeviceGrid

DeviceGrid::~DeviceGrid() {
  plans.clear();
  for (auto e : prof_pool) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

void DeviceGrid::build(const std::vector<AtomInfo> &in) {
  NAPMO_REQUIRE(!in.empty(), "BeckeGrid needs at least one centre");
  // the Becke cell functions divide by the distance of every centre pair
  // (src/grid/becke.cpp:150): checked before any device work
  for (size_t i = 0; i < in.size(); ++i)
    for (size_t j = 0; j < i; ++j) {
      const double dx = in[i].origin[0] - in[j].origin[0], dy = in[i].origin[1] - in[j].origin[1],
                   dz = in[i].origin[2] - in[j].origin[2];
      NAPMO_REQUIRE(dx * dx + dy * dy + dz * dz > 0.0,
                    "coincident centres in BeckeGrid (" + std::to_string(j) + " and " + std::to_string(i) + ")");
    }
  build_geometry(in);
  becke_weights();
}

void DeviceGrid::build_geometry(const std::vector<AtomInfo> &in) {
  require_device();
  atoms = in;
  M = (int)atoms.size();
  if (!stream) NAPMO_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  offsets.assign(M + 1, 0);
  for (int a = 0; a < M; ++a) offsets[a + 1] = offsets[a] + (long)atoms[a].nrad * atoms[a].nang;
  N = offsets[M];
  points.resize(3 * N);
  weights.resize(N);
  origin.resize(3 * M);
  radii.resize(M);
  std::vector<double> hx(N), hy(N), hz(N);
  std::vector<int> own(N);
  std::vector<long> rad_off(M);
  std::vector<int> nrad(M), nang(M), rkind(M);
  std::vector<double> rp0(M), rp1(M), first(M), last(M), allr;
  max_nrad = 0;
  for (int a = 0; a < M; ++a) {
    const AtomInfo &at = atoms[a];
    for (int c = 0; c < 3; ++c) origin[3 * a + c] = at.origin[c];
    radii[a] = at.radii;
    rad_off[a] = (long)allr.size();
    allr.insert(allr.end(), at.r.begin(), at.r.end());
    nrad[a] = at.nrad, nang[a] = at.nang, rkind[a] = at.rtf.kind;
    rp0[a] = at.rtf.p0, rp1[a] = at.rtf.p1;
    first[a] = at.rtf.radius(0);
    last[a] = at.rtf.radius(at.nrad - 1);
    max_nrad = std::max(max_nrad, at.nrad);
    for (int i = 0; i < at.nrad; ++i)
      for (int j = 0; j < at.nang; ++j) {
        const long p = offsets[a] + (long)i * at.nang + j;
        for (int c = 0; c < 3; ++c) points[3 * p + c] = at.u[3 * j + c] * at.r[i] + at.origin[c];
        weights[p] = at.aw[j] * at.rw[i];
        hx[p] = points[3 * p], hy[p] = points[3 * p + 1], hz[p] = points[3 * p + 2];
        own[p] = a;
      }
  }
  x.upload(hx, stream), y.upload(hy, stream), z.upload(hz, stream), w.upload(weights, stream);
  owner.upload(own, stream);
  d_point_off.upload(offsets, stream);
  d_rad_off.upload(rad_off, stream);
  d_nrad.upload(nrad, stream), d_nang.upload(nang, stream), d_rkind.upload(rkind, stream);
  d_rp0.upload(rp0, stream), d_rp1.upload(rp1, stream);
  d_first.upload(first, stream), d_last.upload(last, stream);
  d_center.upload(origin, stream);
  d_r.upload(allr, stream);
}

void DeviceGrid::becke_weights() {
  // Becke pair tables (src/grid/becke.cpp:83-115) with the evaluation sign folded in.
  std::vector<double> R((size_t)M * M, 0.0), As((size_t)M * M, 0.0);
  for (int i = 0; i < M; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) {
        const double d = origin[3 * i + c] - origin[3 * j + c];
        s += d * d;
      }
      R[(size_t)i * M + j] = R[(size_t)j * M + i] = std::sqrt(s);
      double u = (radii[i] - radii[j]) / (radii[i] + radii[j]);
      u = u / (u * u - 1.0);
      u = std::min(std::max(u, -0.45), 0.45);
      As[(size_t)i * M + j] = u;    // i >= j: sign +1
      As[(size_t)j * M + i] = -u;   // j < i evaluated as (j, i): sign -1
      if (i == j) As[(size_t)i * M + i] = 0.0;
    }
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < M; ++j)
      if (i != j) NAPMO_REQUIRE(R[(size_t)i * M + j] > 0.0, "coincident centres in BeckeGrid");
  DeviceBuffer<double> dR, dA;
  dR.upload(R, stream);
  dA.upload(As, stream);
  bw.resize(N);
  wbw.resize(N);
  launch_becke(x.get(), y.get(), z.get(), N, owner.get(), M, d_center.get(), dR.get(), dA.get(),
               bw.get(), stream);
  launch_scale3(w.get(), bw.get(), nullptr, N, N, wbw.get(), stream);
  becke.resize(N);
  bw.download(becke.data(), N, stream);
  NAPMO_HIP(hipStreamSynchronize(stream));
}

void DeviceGrid::build_points(const double *pw, long n, int ncenter) {
  require_device();
  NAPMO_REQUIRE(n > 0, "BeckeGrid_from_points: empty point set");
  NAPMO_REQUIRE(ncenter >= 0, "BeckeGrid_from_points: negative centre count");
  if (!stream) NAPMO_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  atoms.clear();
  M = ncenter;
  N = n;
  offsets.assign(1, 0);
  points.resize(3 * N);
  weights.assign(N, 1.0);
  becke.resize(N);
  origin.clear();
  radii.clear();
  std::vector<double> hx(N), hy(N), hz(N);
  for (long p = 0; p < N; ++p) {
    for (int c = 0; c < 3; ++c) points[3 * p + c] = pw[4 * p + c];
    becke[p] = pw[4 * p + 3];
    hx[p] = pw[4 * p], hy[p] = pw[4 * p + 1], hz[p] = pw[4 * p + 2];
  }
  x.upload(hx, stream), y.upload(hy, stream), z.upload(hz, stream), w.upload(weights, stream);
  bw.upload(becke, stream);
  wbw.resize(N);
  launch_scale3(w.get(), bw.get(), nullptr, N, N, wbw.get(), stream);
  NAPMO_HIP(hipStreamSynchronize(stream));
}

void DeviceGrid::require_atoms(const char *what) const {
  NAPMO_REQUIRE(!atoms.empty(), std::string(what) +
                                    ": this grid was given as points only (no atomic grids)");
}

AtomTables DeviceGrid::tables(const PoissonPlan &plan) const {
  AtomTables t;
  t.M = M;
  t.point_off = d_point_off.get();
  t.nrad = d_nrad.get();
  t.nang = d_nang.get();
  t.rad_off = d_rad_off.get();
  t.yhat_off = plan.yhat_off.get();
  t.green_off = plan.green_off.get();
  t.rkind = d_rkind.get();
  t.rp0 = d_rp0.get();
  t.rp1 = d_rp1.get();
  t.first_x = d_first.get();
  t.last_x = d_last.get();
  t.center = d_center.get();
  t.r = d_r.get();
  t.max_nrad = max_nrad;
  return t;
}

void DeviceGrid::alloc_work(int lmax, PoissonWork &w) const {
  const int L = (lmax + 1) * (lmax + 1);
  long nrad_total = 0;
  for (auto &at : atoms) nrad_total += at.nrad;
  w.rho_lm.resize((size_t)nrad_total * L);
  w.Uy.resize((size_t)nrad_total * L);
  w.Udt.resize((size_t)nrad_total * L);
  w.T.resize((size_t)M * std::max(max_nrad - 1, 1) * L * 4);
  w.ampL.resize((size_t)M * L);
  w.ampR.resize((size_t)M * L);
}

PoissonPlan &DeviceGrid::plan(int lmax) { return build_plan(lmax, false); }
PoissonPlan &DeviceGrid::kinetic_plan(int lmax) { return build_plan(lmax, true); }

PoissonPlan &DeviceGrid::build_plan(int lmax, bool kinetic) {
  require_atoms(kinetic ? "kinetic operator" : "Poisson solver");
  NAPMO_REQUIRE(lmax >= 0 && lmax <= MAX_LMAX,
                "lmax must be in [0, " + std::to_string(MAX_LMAX) + "]");
  auto &cache = kinetic ? kplans : plans;
  auto it = cache.find(lmax);
  if (it != cache.end()) return *it->second;
  auto p = std::make_unique<PoissonPlan>();
  p->lmax = lmax;
  const int L = (lmax + 1) * (lmax + 1);
  // projection tables, one per distinct (Lebedev rule, harmonic cut): with a cut c below
  // lmax (lcut, grids of several molecules) the columns of l > c are zero, so those
  // projections, radial solutions and synthesis terms are exact zeros
  std::map<std::pair<int, int>, long> yoff;
  std::vector<double> yh;
  std::vector<long> yh_off(M);
  for (int a = 0; a < M; ++a) {
    const int na = atoms[a].nang;
    const int cut = lcut.empty() ? lmax : std::min(lcut[a], lmax);
    const auto key = std::make_pair(na, cut);
    if (!yoff.count(key)) {
      AngularGrid ag;
      ag.lorder = na;
      ag.points = atoms[a].u;
      ag.weights = atoms[a].aw;
      yoff[key] = (long)yh.size();
      auto t = projection_table(ag, lmax);
      const int Lc = (cut + 1) * (cut + 1);
      for (size_t e = 0; e < t.size(); ++e)
        if ((int)(e % L) >= Lc) t[e] = 0.0;
      yh.insert(yh.end(), t.begin(), t.end());
    }
    yh_off[a] = yoff[key];
  }
  // Green matrices, one set per distinct radial grid
  std::vector<double> gr;
  std::vector<long> g_off(M);
  std::vector<std::pair<std::vector<double>, long>> seen;  // (key, offset)
  for (int a = 0; a < M; ++a) {
    const AtomInfo &at = atoms[a];
    std::vector<double> key = {(double)at.rtf.kind, at.rtf.p0, at.rtf.p1, (double)at.nrad};
    long off = -1;
    for (auto &s : seen)
      if (s.first == key) off = s.second;
    if (off < 0) {
      off = (long)gr.size();
      for (int l = 0; l <= lmax; ++l) {
        std::vector<double> Gy, Gd;
        if (kinetic)
          kinetic_green(at.rtf, at.radii, at.r, l, Gy, Gd);
        else
          radial_green(at.rtf, at.rw, l, Gy, Gd);
        gr.insert(gr.end(), Gy.begin(), Gy.end());
        gr.insert(gr.end(), Gd.begin(), Gd.end());
      }
      seen.push_back({key, off});
    }
    g_off[a] = off;
  }
  p->yhat.upload(yh, stream);
  p->yhat_off.upload(yh_off, stream);
  p->green.upload(gr, stream);
  p->green_off.upload(g_off, stream);
  alloc_work(lmax, p->work);
  NAPMO_HIP(hipStreamSynchronize(stream));
  auto &ref = *p;
  cache[lmax] = std::move(p);
  return ref;
}

void DeviceGrid::kinetic_expand(const double *d_psi, int lmax) {
  PoissonPlan &pl = kinetic_plan(lmax);
  AtomTables at = tables(pl);
  PoissonWork &w = pl.work;
  launch_project(at, d_psi, bw.get(), pl.yhat.get(), lmax, w.rho_lm.get(), stream);
  launch_radial(at, w.rho_lm.get(), pl.green.get(), lmax, w.Uy.get(), w.Udt.get(), stream);
}

void DeviceGrid::kinetic(const double *d_psi, int lmax, double *d_out) {
  kinetic_expand(d_psi, lmax);
  PoissonPlan &pl = kinetic_plan(lmax);
  AtomTables at = tables(pl);
  PoissonWork &w = pl.work;
  SynthPairs &sp = synth_pairs();
  launch_synth_tables(at, w.Uy.get(), w.Udt.get(), lmax, w.T.get(), w.ampL.get(), w.ampR.get(),
                      stream);
  launch_synth2(at, sp.geo.get(), N, sp.list.get(), sp.wave_bucket.get(), sp.nwaves,
                w.T.get(), w.ampL.get(), w.ampR.get(), lmax, sp.contrib.get(), stream, synth_np);
  launch_contrib_reduce(sp.contrib.get(), M, N, 0, d_out, stream, -0.5);
}

void DeviceGrid::tic(int phase) {
  if (!prof_on) return;
  if (prof_next + 2 > prof_pool.size()) {
    hipEvent_t a, b;
    NAPMO_HIP(hipEventCreate(&a));
    NAPMO_HIP(hipEventCreate(&b));
    prof_pool.push_back(a), prof_pool.push_back(b);
  }
  prof_start_ev = prof_pool[prof_next++];
  prof_phase = phase;
  NAPMO_HIP(hipEventRecord(prof_start_ev, stream));
}

void DeviceGrid::toc() {
  if (!prof_on) return;
  hipEvent_t e = prof_pool[prof_next++];
  NAPMO_HIP(hipEventRecord(e, stream));
  prof_marks.push_back({prof_phase, {prof_start_ev, e}});
}

void DeviceGrid::profile_start() {
  NAPMO_HIP(hipStreamSynchronize(stream));
  prof_on = true;
  prof_calls = 0;
  prof_next = 0;
  prof_marks.clear();
}

int DeviceGrid::profile_read(double *avg_ms, int nphase) {
  NAPMO_HIP(hipStreamSynchronize(stream));
  const int k = std::min(nphase, DPSI_NPHASE);
  for (int p = 0; p < k; ++p) avg_ms[p] = 0.0;
  for (auto &m : prof_marks) {
    float ms = 0.f;
    NAPMO_HIP(hipEventElapsedTime(&ms, m.second.first, m.second.second));
    if (m.first < k) avg_ms[m.first] += ms;
  }
  if (prof_calls)
    for (int p = 0; p < k; ++p) avg_ms[p] /= prof_calls;
  return prof_calls;
}

HelmholtzTables HelmholtzPlan::tables() const {
  HelmholtzTables t;
  t.c0 = c0.get(), t.oa = oa.get(), t.o2 = o2.get(), t.j1 = j1.get(), t.j2 = j2.get();
  t.c0_off = c0_off.get(), t.oa_off = oa_off.get(), t.o2_off = o2_off.get();
  return t;
}

HelmholtzPlan &DeviceGrid::helmholtz() {
  if (hplan) return *hplan;
  require_atoms("orbital correction");
  auto h = std::make_unique<HelmholtzPlan>();
  std::vector<double> c0, oa, o2, j1all, j2all;
  std::vector<long> c0o(M), oao(M), o2o(M);
  std::vector<std::pair<std::vector<double>, int>> seen;  // (radial grid key, first atom)
  for (int a = 0; a < M; ++a) {
    const AtomInfo &at = atoms[a];
    NAPMO_REQUIRE(at.nrad >= 3 && at.nrad <= HELMHOLTZ_MAX_NRAD,
                  "orbital correction: radial grids need 3 to " +
                      std::to_string(HELMHOLTZ_MAX_NRAD) + " points");
    std::vector<double> tc0, toa, to2, tj1, tj2;
    helmholtz_tables(at.rtf, tc0, toa, to2, tj1, tj2);
    j1all.insert(j1all.end(), tj1.begin(), tj1.end());
    j2all.insert(j2all.end(), tj2.begin(), tj2.end());
    std::vector<double> key = {(double)at.rtf.kind, at.rtf.p0, at.rtf.p1, (double)at.nrad};
    int first = -1;
    for (auto &s : seen)
      if (s.first == key) first = s.second;
    if (first >= 0) {
      c0o[a] = c0o[first], oao[a] = oao[first], o2o[a] = o2o[first];
      continue;
    }
    seen.push_back({key, a});
    c0o[a] = (long)c0.size(), oao[a] = (long)oa.size(), o2o[a] = (long)o2.size();
    c0.insert(c0.end(), tc0.begin(), tc0.end());
    oa.insert(oa.end(), toa.begin(), toa.end());
    o2.insert(o2.end(), to2.begin(), to2.end());
  }
  h->c0.upload(c0, stream), h->oa.upload(oa, stream), h->o2.upload(o2, stream);
  h->j1.upload(j1all, stream), h->j2.upload(j2all, stream);
  h->c0_off.upload(c0o, stream), h->oa_off.upload(oao, stream), h->o2_off.upload(o2o, stream);
  NAPMO_HIP(hipStreamSynchronize(stream));
  hplan = std::move(h);
  return *hplan;
}

void DeviceGrid::orbital_correction(int lmax, int norb, const double *d_phi, const double *d_doi,
                                    const double *d_oi, const double *d_ri, const double *d_V,
                                    double mass_inv, double *d_out) {
  NAPMO_REQUIRE(norb >= 0, "orbital correction: negative orbital count");
  NAPMO_REQUIRE(mass_inv != 0.0, "orbital correction: mass_inv must be non-zero");
  if (norb == 0) return;
  PoissonPlan &pl = plan(lmax);
  HelmholtzPlan &hp = helmholtz();
  AtomTables at = tables(pl);
  const int L = (lmax + 1) * (lmax + 1);
  long S = 0;
  for (auto &a : atoms) S += a.nrad;
  if (!owork) owork = std::make_unique<OrbitalWork>();
  OrbitalWork &ow = *owork;
  ow.src.resize((size_t)norb * N);
  ow.mu.resize((size_t)norb * S * L);
  ow.Uy.resize((size_t)norb * S * L);
  ow.Udt.resize((size_t)norb * S * L);
  ow.Vsph.resize((size_t)norb * S);
  ow.status.resize(1);
  if (ow.lmax != lmax) {
    alloc_work(lmax, ow.pw);
    ow.lmax = lmax;
  }
  PoissonWork &wk = ow.pw;
  NAPMO_HIP(hipMemsetAsync(ow.status.get(), 0, sizeof(int), stream));
  if (prof_on) ++prof_calls;
  // sources (phi doi - ri) bw projected per atom, V_sph = spherical average of V bw
  tic(1);
  launch_scaled_diff(d_phi, d_doi, d_ri, norb, N, ow.src.get(), stream);
  for (int o0 = 0; o0 < norb; o0 += SET_BATCH)
    launch_project(at, ow.src.get() + (long)o0 * N, bw.get(), pl.yhat.get(), lmax,
                   ow.mu.get() + (long)o0 * S * L, stream, std::min(SET_BATCH, norb - o0), N, S * L);
  toc();
  tic(2);
  launch_shell_average(at, d_V, norb, N, w.get(), bw.get(), S, ow.Vsph.get(), stream);
  toc();
  tic(3);
  launch_helmholtz_radial(at, hp.tables(), lmax, norb, S, ow.Vsph.get(), d_oi, d_doi, mass_inv,
                          ow.mu.get(), ow.Uy.get(), ow.Udt.get(), ow.status.get(), stream);
  toc();
  synth_sets(lmax, norb, ow.Uy.get(), ow.Udt.get(), S * L, d_out, N);
  int h_status = 0;
  NAPMO_HIP(hipMemcpyAsync(&h_status, ow.status.get(), sizeof(int), hipMemcpyDeviceToHost, stream));
  NAPMO_HIP(hipStreamSynchronize(stream));
  NAPMO_REQUIRE(h_status == 0, "orbital correction: singular radial system (zero pivot)");
}

void DeviceGrid::exchange_potential(int ndim, const double *d_psi, const double *d_D, int lmax,
                                    double *d_K) {
  if (ndim <= 0) return;
  PoissonPlan &pl = plan(lmax);
  AtomTables at = tables(pl);
  // orbital pairs (s, v), v <= s, in order, SET_BATCH at a time: pair densities, their
  // projections and radial solves in one launch each, the synthesis of the batch
  // (synth_sets), then per pair the accumulation into K
  std::vector<std::pair<int, int>> prs;
  for (int s = 0; s < ndim; ++s)
    for (int v = 0; v <= s; ++v) prs.push_back({s, v});
  const int npair = (int)prs.size(), nb = std::min(SET_BATCH, npair);
  const int L = (lmax + 1) * (lmax + 1);
  long S = 0;
  for (auto &a : atoms) S += a.nrad;
  xchi.resize((size_t)ndim * N);
  xdens.resize((size_t)nb * N);
  xaux.resize((size_t)nb * N);
  xrho.resize((size_t)nb * S * L);
  xUy.resize((size_t)nb * S * L);
  xUdt.resize((size_t)nb * S * L);
  if (prof_on) ++prof_calls;
  launch_exchange_chi(d_psi, d_D, ndim, N, xchi.get(), stream);
  launch_axpby(0.0, xchi.get(), 0.0, xchi.get(), (long)ndim * N, d_K, stream);  // K = 0
  for (int c0 = 0; c0 < npair; c0 += SET_BATCH) {
    const int cnt = std::min(SET_BATCH, npair - c0);
    tic(1);
    for (int k = 0; k < cnt; ++k)
      launch_mul(d_psi + (long)prs[c0 + k].first * N, d_psi + (long)prs[c0 + k].second * N, N,
                 xdens.get() + (long)k * N, stream);
    launch_project(at, xdens.get(), bw.get(), pl.yhat.get(), lmax, xrho.get(), stream, cnt, N, S * L);
    launch_radial(at, xrho.get(), pl.green.get(), lmax, xUy.get(), xUdt.get(), stream, cnt, S * L);
    toc();
    synth_sets(lmax, cnt, xUy.get(), xUdt.get(), S * L, xaux.get(), N);
    tic(6);
    for (int k = 0; k < cnt; ++k)
      launch_exchange_acc(xaux.get() + (long)k * N, xchi.get(), prs[c0 + k].first,
                          prs[c0 + k].second, N, d_K, stream);
    toc();
  }
}

template <class T> static void grow(DeviceBuffer<T> &b, size_t n) {
  if (b.size() < n) b.resize(n);
}

void DeviceGrid::synth_sets(int lmax, int nset, const double *Uy, const double *Udt, long ustride,
                            double *out, long ostride, double scale) {
  if (nset <= 0) return;
  PoissonPlan &pl = plan(lmax);
  AtomTables at = tables(pl);
  SynthPairs &sp = synth_pairs();
  const long L = (long)(lmax + 1) * (lmax + 1);
  const long tsz = synth_table_size(at, lmax);
  int nmfma = synth_mfma_min > 16 ? 0 : (nset / 16) * 16;
  if (synth_mfma_min <= 16 && nset - nmfma >= std::max(synth_mfma_min, 1)) nmfma = nset;
  const long tile = (long)M * N * 16;
  const int per = (int)std::max<long>(1, SYNTH_MFMA_BUDGET / (tile * (long)sizeof(double)));
  for (int s0 = 0; s0 < nmfma; s0 += 16 * per) {
    const int cnt = std::min(16 * per, nmfma - s0);
    const int ntile = (cnt + 15) / 16, ndp = 16 * ntile;
    grow(mT, (size_t)tsz * ndp);
    grow(maL, (size_t)M * L * ndp);
    grow(maR, (size_t)M * L * ndp);
    grow(mcontrib, (size_t)tile * ntile);
    tic(4);
    launch_synth_tables_cols(at, Uy + s0 * ustride, Udt + s0 * ustride, lmax, cnt, ustride, ndp,
                             mT.get(), maL.get(), maR.get(), stream);
    toc();
    tic(7);
    launch_synth_mfma(at, sp.geo.get(), N, sp.list.get(), sp.wave_bucket.get(), sp.nwaves, ntile,
                      mT.get(), maL.get(), maR.get(), lmax, mcontrib.get(), stream);
    toc();
    tic(6);
    for (int t = 0; t < ntile; ++t)
      launch_contrib_reduce_tile(mcontrib.get() + t * tile, M, N, std::min(16, cnt - 16 * t), scale,
                                 out + (long)(s0 + 16 * t) * ostride, ostride, stream);
    toc();
  }
  const int rest = nset - nmfma;
  if (rest <= 0) return;
  grow(sT, (size_t)tsz * rest);
  grow(saL, (size_t)M * L * rest);
  grow(saR, (size_t)M * L * rest);
  tic(4);
  launch_synth_tables(at, Uy + nmfma * ustride, Udt + nmfma * ustride, lmax, sT.get(), saL.get(),
                      saR.get(), stream, 1, rest, ustride);
  toc();
  for (int k = 0; k < rest; ++k) {
    tic(5);
    launch_synth2(at, sp.geo.get(), N, sp.list.get(), sp.wave_bucket.get(), sp.nwaves,
                  sT.get() + k * tsz, saL.get() + k * M * L, saR.get() + k * M * L, lmax,
                  sp.contrib.get(), stream, synth_np);
    toc();
    tic(6);
    launch_contrib_reduce(sp.contrib.get(), M, N, 0, out + (long)(nmfma + k) * ostride, stream, scale);
    toc();
  }
}

void DeviceGrid::becke_dickson_step(int lmax, int norb, const double *d_psi, const double *d_T,
                                    const double *d_V, const double *d_e, double mass_inv,
                                    double *d_dpsi, double *d_de, double *d_R) {
  NAPMO_REQUIRE(norb >= 0, "Becke-Dickson step: negative orbital count");
  if (norb == 0) return;
  if (!owork) owork = std::make_unique<OrbitalWork>();
  owork->resid.resize((size_t)norb * RESID_BLOCKS);
  tic(0);
  launch_residual(d_psi, d_T, d_V, d_e, wbw.get(), norb, N, d_R, d_de, owork->resid.get(), stream);
  toc();
  orbital_correction(lmax, norb, d_psi, d_de, d_e, d_R, d_V, mass_inv, d_dpsi);
}

void DeviceGrid::poisson_expand(const double *d_dens, int lmax) {
  PoissonPlan &pl = plan(lmax);
  AtomTables at = tables(pl);
  launch_project(at, d_dens, bw.get(), pl.yhat.get(), lmax, pl.work.rho_lm.get(), stream);
  launch_radial(at, pl.work.rho_lm.get(), pl.green.get(), lmax, pl.work.Uy.get(), pl.work.Udt.get(), stream);
}

void DeviceGrid::poisson_synthesize(int lmax, double *d_v) {
  PoissonPlan &pl = plan(lmax);
  AtomTables at = tables(pl);
  SynthPairs &sp = synth_pairs();
  launch_synth_tables(at, pl.work.Uy.get(), pl.work.Udt.get(), lmax, pl.work.T.get(), pl.work.ampL.get(),
                      pl.work.ampR.get(), stream);
  launch_synth2(at, sp.geo.get(), N, sp.list.get(), sp.wave_bucket.get(), sp.nwaves,
                pl.work.T.get(), pl.work.ampL.get(), pl.work.ampR.get(), lmax, sp.contrib.get(), stream, synth_np);
  launch_contrib_reduce(sp.contrib.get(), M, N, 0, d_v, stream);
}

AtomTables DeviceGrid::geometry_tables() const {
  require_atoms("synthesis");
  AtomTables at;
  at.M = M;
  at.point_off = d_point_off.get();
  at.nrad = d_nrad.get();
  at.nang = d_nang.get();
  at.rad_off = d_rad_off.get();
  at.yhat_off = nullptr;
  at.green_off = nullptr;
  at.rkind = d_rkind.get();
  at.rp0 = d_rp0.get();
  at.rp1 = d_rp1.get();
  at.first_x = d_first.get();
  at.last_x = d_last.get();
  at.center = d_center.get();
  at.r = d_r.get();
  at.max_nrad = max_nrad;
  return at;
}

std::unique_ptr<SynthPairs> build_synth_pairs(const DeviceGrid &src, const double *x,
                                              const double *y, const double *z, long N,
                                              hipStream_t stream) {
  // bucket geometry does not depend on lmax: only the atom arrays are needed
  const AtomTables at = src.geometry_tables();
  const int M = src.M;
  const int nb = M * (src.max_nrad + 1);
  DeviceBuffer<int> count(nb), cursor(nb);
  NAPMO_HIP(hipMemsetAsync(count.get(), 0, sizeof(int) * nb, stream));
  launch_pair_count(at, x, y, z, N, count.get(), stream);
  std::vector<int> hc(nb);
  count.download(hc.data(), nb, stream);
  NAPMO_HIP(hipStreamSynchronize(stream));
  std::vector<long> off(nb);
  std::vector<int> wb;
  long total = 0;
  for (int b = 0; b < nb; ++b) {
    off[b] = total;
    // padded to SYNTH_PAD pairs, so a wavefront may take two 64-pair slots of one bucket
    const long waves = div_up(hc[b], SYNTH_PAD) * (SYNTH_PAD / 64);
    for (long w = 0; w < waves; ++w) wb.push_back(b);
    total += waves * 64;
  }
  auto p = std::make_unique<SynthPairs>();
  p->nwaves = (long)wb.size();
  NAPMO_REQUIRE(total < (1L << 31), "too many (point, centre) pairs for the synthesis lists");
  p->list.resize(std::max<long>(total, 1));
  NAPMO_HIP(hipMemsetAsync(p->list.get(), 0xFF, sizeof(int) * std::max<long>(total, 1), stream));
  DeviceBuffer<long> doff;
  doff.upload(off, stream);
  NAPMO_HIP(hipMemsetAsync(cursor.get(), 0, sizeof(int) * nb, stream));
  launch_pair_fill(at, x, y, z, N, doff.get(), cursor.get(), p->list.get(), stream);
  p->wave_bucket.upload(wb, stream);
  std::vector<long> s0(nb + 1);
  for (int b = 0; b < nb; ++b) s0[b] = off[b] / 64;
  s0[nb] = total / 64;
  p->slot0.upload(s0, stream);
  p->geo.resize((size_t)std::max<long>(p->nwaves, 1) * 256);
  launch_pair_geometry(at, x, y, z, p->list.get(), p->wave_bucket.get(), p->nwaves, p->geo.get(),
                       stream);
  p->contrib.resize((size_t)M * N);
  NAPMO_HIP(hipStreamSynchronize(stream));
  return p;
}

SynthPairs &DeviceGrid::synth_pairs() {
  if (!pairs) pairs = build_synth_pairs(*this, x.get(), y.get(), z.get(), N, stream);
  return *pairs;
}

void DeviceGrid::synthesize_at(int lmax, const PoissonWork &w, const SynthPairs &sp,
                               const double *tx, const double *ty, const double *tz, long tN,
                               double scale, int accumulate, double *d_v, hipStream_t st) {
  PoissonPlan &pl = plan(lmax);
  AtomTables at = tables(pl);
  launch_synth2(at, sp.geo.get(), tN, sp.list.get(), sp.wave_bucket.get(), sp.nwaves, w.T.get(),
                w.ampL.get(), w.ampR.get(), lmax, sp.contrib.get(), st, synth_np);
  launch_contrib_reduce(sp.contrib.get(), M, tN, accumulate, d_v, st, scale);
}

}  // namespace napmo

namespace napmo {

void decomposition_on_points(CubicSpline **sp, int nspline, int lmax, const double *center,
                             const double *points, long npoint, double *output) {
  require_device();
  NAPMO_REQUIRE(lmax <= MAX_LMAX, "lmax above the supported maximum");
  const RTransform &rtf = *sp[0]->rtf;
  const int n = sp[0]->n;
  for (int s = 1; s < nspline; ++s)
    NAPMO_REQUIRE(sp[s]->n == n && sp[s]->rtf->kind == rtf.kind && sp[s]->rtf->p0 == rtf.p0 &&
                      sp[s]->rtf->p1 == rtf.p1,
                  "eval_decomposition_grid: splines must share one radial transform");
  const bool tail = sp[0]->ext->has_tail();
  const int L = nspline;
  // U arrays in reference lm order [n x L]
  std::vector<double> Uy((size_t)n * L), Ud((size_t)n * L);
  for (int s = 0; s < L; ++s)
    for (int i = 0; i < n; ++i) Uy[(size_t)i * L + s] = sp[s]->y[i], Ud[(size_t)i * L + s] = sp[s]->dt[i];
  // extrapolation parameters in synthesis order, scaled like the table coefficients
  std::vector<int> ql, qm, qc;
  std::vector<double> qs;
  synth_order(lmax, ql, qm, qc, qs);
  std::vector<double> aL(L), aR(L);
  for (int q = 0; q < L; ++q) {
    const int ref = ql[q] * ql[q] + (qm[q] ? 2 * qm[q] - 1 + qc[q] : 0);
    const Extrapolation &e = *sp[ref]->ext;
    if (tail) {
      aL[q] = e.amp_left * qs[q], aR[q] = e.amp_right * qs[q];
    } else if (e.kind == EX_CUSP) {
      aL[q] = e.a0 * qs[q], aR[q] = e.b0;
    } else {
      aL[q] = 0.0, aR[q] = 0.0;
    }
  }
  DeviceGrid one;  // single-centre tables (no grid points needed beyond the radial nodes)
  hipStream_t st;
  NAPMO_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  one.stream = st;
  std::vector<long> z0 = {0, 0}, roff = {0};
  std::vector<int> nr = {n}, na = {1}, rk = {rtf.kind};
  std::vector<double> p0 = {rtf.p0}, p1 = {rtf.p1}, fx = {sp[0]->first_x}, lx = {sp[0]->last_x};
  std::vector<double> c = {center[0], center[1], center[2]}, rr(n);
  for (int i = 0; i < n; ++i) rr[i] = rtf.radius(i);
  one.d_point_off.upload(z0, st), one.d_rad_off.upload(roff, st);
  one.d_nrad.upload(nr, st), one.d_nang.upload(na, st), one.d_rkind.upload(rk, st);
  one.d_rp0.upload(p0, st), one.d_rp1.upload(p1, st), one.d_first.upload(fx, st);
  one.d_last.upload(lx, st), one.d_center.upload(c, st), one.d_r.upload(rr, st);
  one.M = 1;
  one.max_nrad = n;
  PoissonPlan pl;
  pl.work.Uy.upload(Uy, st), pl.work.Udt.upload(Ud, st);
  pl.work.ampL.upload(aL, st), pl.work.ampR.upload(aR, st);
  pl.work.T.resize((size_t)(n - 1) * L * 4);
  AtomTables at = one.tables(pl);
  launch_synth_tables(at, pl.work.Uy.get(), pl.work.Udt.get(), lmax, pl.work.T.get(), pl.work.ampL.get(), pl.work.ampR.get(),
                      st, 0);
  std::vector<double> hx(npoint), hy(npoint), hz(npoint);
  for (long i = 0; i < npoint; ++i)
    hx[i] = points[3 * i], hy[i] = points[3 * i + 1], hz[i] = points[3 * i + 2];
  DeviceBuffer<double> dx, dy, dz, dv;
  dx.upload(hx, st), dy.upload(hy, st), dz.upload(hz, st), dv.upload(output, npoint, st);
  launch_synth(at, dx.get(), dy.get(), dz.get(), npoint, pl.work.T.get(), pl.work.ampL.get(), pl.work.ampR.get(),
               lmax, 1, dv.get(), st, tail ? 0 : 1);
  dv.download(output, npoint, st);
  NAPMO_HIP(hipStreamSynchronize(st));
}

}  // namespace napmo
