/* C/OpenMP restatement of the heavy loops of the nAPMO grid hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for parity tests and the CPU baseline of
 * bench.py. The product never links it. Built by ``make -C oracle`` into
 * oracle/liboracle.so and bound in oracle/cext.py.
 *
 * All arithmetic is float64 and follows the reference loop structure; each
 * function cites the reference file:line it restates.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define INV_SQRT_4_PI 0.28209479177387814347

/* ------------------------------------------------------------------------------
 * Hermite finite-element assembly (src/horton/ode2.cpp:232-321).
 * tab2[64] / tab3[512]: per-segment integrals of products of Hermite shape
 * functions, indexed by 3 bits per factor (relative node 0/1, kind, derivative),
 * computed exactly by oracle/core.py and rounded once.
 * ---------------------------------------------------------------------------- */
static int fbits(long i, int deriv, long seg) {
  long c = i / 2 - seg;
  if (c < 0 || c > 1) return -1;
  return (int)((c << 2) | ((i % 2) << 1) | (deriv ? 1 : 0));
}

static double seg2(const double *tab2, long seg, long i0, int d0, long i1, int d1) {
  int a = fbits(i0, d0, seg), b = fbits(i1, d1, seg);
  if (a < 0 || b < 0) return 0.0;
  return tab2[(a << 3) | b];
}

static double seg3(const double *tab3, long seg, long i0, int d0, long i1, int d1, long i2,
                   int d2) {
  int a = fbits(i0, d0, seg), b = fbits(i1, d1, seg), c = fbits(i2, d2, seg);
  if (a < 0 || b < 0 || c < 0) return 0.0;
  return tab3[(a << 6) | (b << 3) | c];
}

static double ov2(const double *tab2, long xmax, long i0, int d0, long i1, int d1) {
  long c0 = i0 / 2, c1 = i1 / 2;
  if (labs(c1 - c0) >= 2) return 0.0;
  double r = 0.0;
  if (c0 >= 1 && c0 <= xmax) r += seg2(tab2, c0 - 1, i0, d0, i1, d1);
  if (c0 >= 0 && c0 <= xmax - 1) r += seg2(tab2, c0, i0, d0, i1, d1);
  return r;
}

static double ov3(const double *tab3, long xmax, long i0, int d0, long i1, int d1, long i2,
                  int d2) {
  long c0 = i0 / 2;
  if (labs(i1 / 2 - c0) >= 2 || labs(i2 / 2 - c0) >= 2) return 0.0;
  double r = 0.0;
  if (c0 >= 1 && c0 <= xmax) r += seg3(tab3, c0 - 1, i0, d0, i1, d1, i2, d2);
  if (c0 >= 0 && c0 <= xmax - 1) r += seg3(tab3, c0, i0, d0, i1, d1, i2, d2);
  return r;
}

static double node(long x, long i, int deriv) {
  return ((i % 2) == deriv) ? (double)(x == i / 2) : 0.0;
}

void oracle_build_ode2(const double *b, const double *a, const double *f, const double *bcs,
                       double *coeffs, double *rhs, long npoint, const double *tab2,
                       const double *tab3) {
  long nfn = 2 * npoint, xmax = npoint - 1;
  long skip[2] = {-1, -1};
  int cnt = 0;
  const long rows[4] = {0, 1, nfn - 2, nfn - 1};
  for (int k = 0; k < 4; ++k) {
    if (bcs[k] == bcs[k]) {
      coeffs[rows[k] * nfn + rows[k]] = 1.0;
      rhs[rows[k]] = bcs[k];
      if (cnt < 2) skip[cnt] = rows[k];
      cnt++;
    }
  }
  for (long irow = 0; irow < nfn; ++irow) {
    if (irow == skip[0] || irow == skip[1]) continue;
    long begin = 2 * (irow / 2 - 1), end = 2 * (irow / 2 + 2);
    if (begin < 0) begin = 0;
    if (end > nfn) end = nfn;
    for (long icol = begin; icol < end; ++icol) {
      double c = coeffs[irow * nfn + icol];
      c -= ov2(tab2, xmax, irow, 1, icol, 1);
      for (long k = begin; k < end; ++k) {
        if (b[k] != 0) c += b[k] * ov3(tab3, xmax, irow, 0, icol, 1, k, 0);
        if (a[k] != 0) c += a[k] * ov3(tab3, xmax, irow, 0, icol, 0, k, 0);
      }
      c += node(xmax, irow, 0) * node(xmax, icol, 1);
      c -= node(0, irow, 0) * node(0, icol, 1);
      coeffs[irow * nfn + icol] = c;
    }
    double acc = rhs[irow];
    for (long k = begin; k < end; ++k)
      if (f[k] != 0) acc += f[k] * ov2(tab2, xmax, irow, 0, k, 0);
    rhs[irow] = acc;
  }
}

/* ------------------------------------------------------------------------------
 * Becke weights (src/grid/becke.cpp:79-176), generalised to uneven atomic grids
 * through offsets[M+1] (equal to the reference when all grids have one size).
 * ---------------------------------------------------------------------------- */
void oracle_becke_weights(const double *points, long npoint, const double *origin,
                          const double *radii, int ncenter, const long *offsets, double *out) {
  int m = ncenter;
  double R[m * m], A[m * m];
  for (int i = 0; i < m; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) {
        double d = origin[3 * i + k] - origin[3 * j + k];
        s += d * d;
      }
      R[i * m + j] = R[j * m + i] = sqrt(s);
      double u = (radii[i] - radii[j]) / (radii[i] + radii[j]);
      u = u / (u * u - 1.0);
      if (u > 0.45) u = 0.45;
      else if (u < -0.45) u = -0.45;
      A[i * m + j] = A[j * m + i] = u;
    }
  for (int atom = 0; atom < m; ++atom) {
#pragma omp parallel for schedule(static)
    for (long p = offsets[atom]; p < offsets[atom + 1]; ++p) {
      const double *pt = points + 3 * p;
      double sum = 0.0, mine = 0.0;
      for (int i = 0; i < m; ++i) {
        double prod = 1.0;
        for (int j = 0; j < m; ++j) {
          if (i == j) continue;
          double ri = 0.0, rj = 0.0;
          for (int k = 0; k < 3; ++k) {
            double d = pt[k] - origin[3 * i + k];
            ri += d * d;
            double e = pt[k] - origin[3 * j + k];
            rj += e * e;
          }
          ri = sqrt(ri);
          rj = sqrt(rj);
          double mu = (ri - rj) / R[i * m + j];
          double s = mu + A[i * m + j] * ((1.0 - 2 * (i < j)) * (1.0 - mu * mu));
          for (int k = 1; k <= 3; ++k) s = 0.5 * s * (3 - s * s);
          s = 0.5 * (1.0 - s);
          prod *= s;
        }
        if (i == atom) mine = prod;
        sum += prod;
      }
      out[p] = mine / sum;
    }
  }
}

/* ------------------------------------------------------------------------------
 * Collocation (src/system/basis_set.cpp:69-82, gto.cpp:41-51, 182-192).
 * fn_origin[3n], fn_l[3n], fn_nprim[n], fn_first[n] index into prim_zeta/coef
 * (coef already times the primitive norm), fn_norm[n] contraction norms.
 * out [N x n] point-major.
 * ---------------------------------------------------------------------------- */
void oracle_basis_compute(const double *points, long npoint, int nbasis, const double *fn_origin,
                          const int *fn_l, const int *fn_nprim, const int *fn_first,
                          const double *fn_norm, const double *prim_zeta, const double *prim_coef,
                          double *out) {
#pragma omp parallel for schedule(static)
  for (long p = 0; p < npoint; ++p) {
    for (int mu = 0; mu < nbasis; ++mu) {
      double d[3], r2 = 0.0, ang = 1.0;
      for (int k = 0; k < 3; ++k) {
        d[k] = points[3 * p + k] - fn_origin[3 * mu + k];
        ang *= pow(d[k], fn_l[3 * mu + k]);
        r2 += d[k] * d[k];
      }
      double acc = 0.0;
      for (int k = 0; k < fn_nprim[mu]; ++k) {
        int q = fn_first[mu] + k;
        acc += prim_coef[q] * ang * exp(-prim_zeta[q] * r2);
      }
      out[p * nbasis + mu] = acc * fn_norm[mu];
    }
  }
}

/* ------------------------------------------------------------------------------
 * Racah solid harmonics of (z, x, y) for l = 1..lmax (src/horton/moments.cpp:59-125).
 * ---------------------------------------------------------------------------- */
static void pure_polys(double *work, int lmax) {
  if (lmax <= 1) return;
  double z = work[0], x = work[1], y = work[2];
  double r2 = x * x + y * y + z * z;
  double pi_old[lmax + 1], pi_new[lmax + 1], a[lmax + 1], b[lmax + 1];
  pi_old[0] = 1;
  pi_new[0] = z;
  pi_new[1] = 1;
  a[1] = x;
  b[1] = y;
  int off = 0, np = 3;
  for (int l = 2; l <= lmax; ++l) {
    int nnp = np + 2, noff = off + np;
    double factor = 2 * l - 1;
    for (int m = 0; m <= l - 2; ++m) {
      double tmp = pi_old[m];
      pi_old[m] = pi_new[m];
      pi_new[m] = (z * factor * pi_old[m] - r2 * (l + m - 1) * tmp) / (l - m);
    }
    pi_old[l - 1] = pi_new[l - 1];
    pi_new[l] = factor * pi_old[l - 1];
    pi_new[l - 1] = z * pi_new[l];
    a[l] = x * a[l - 1] - y * b[l - 1];
    b[l] = x * b[l - 1] + y * a[l - 1];
    work[noff] = pi_new[0];
    factor = M_SQRT2;
    for (int m = 1; m <= l; ++m) {
      factor /= sqrt((l + m) * (l - m + 1));
      work[noff + 2 * m - 1] = factor * a[m] * pi_new[m];
      work[noff + 2 * m] = factor * b[m] * pi_new[m];
    }
    np = nnp;
    off = noff;
  }
}

/* Radial transform evaluated as src/horton/rtransform.cpp: kind 0 Chebyshev
 * (p0 = radii), kind 1 Power (p0 = rmin, p1 = power), kind 2 identity. */
static double rtf_inv(int kind, double p0, double p1, int n, double r) {
  if (kind == 0) return -(acos((r - p0) / (r + p0)) / (M_PI / (n + 1))) + n;
  if (kind == 1) return pow(r / p0, 1.0 / p1) - 1;
  return r;
}

/* ------------------------------------------------------------------------------
 * Synthesis eval_decomposition_grid (src/horton/evaluate.cpp:129-223, no periodic
 * images) with the potential splines of one centre: y, dt [nspline x n],
 * extrapolation amplitudes amp_left/right [nspline] (PotentialExtrapolation,
 * src/horton/extrapolation.cpp:156-167). out[N] accumulates.
 * ---------------------------------------------------------------------------- */
void oracle_eval_decomposition(const double *y, const double *dt, const double *amp_left,
                               const double *amp_right, int n, int nspline, int rkind, double rp0,
                               double rp1, double first_x, double last_x, const double *center,
                               const double *points, long npoint, double *out) {
  int lmax = (int)sqrt((double)nspline) - 1;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < npoint; ++i) {
    double work[nspline > 1 ? nspline - 1 : 1];
    double x = points[3 * i + 0] - center[0];
    double zz = points[3 * i + 2] - center[2];
    double yy = points[3 * i + 1] - center[1];
    double d = sqrt(x * x + yy * yy + zz * zz);
    int reg = (d < first_x) ? 0 : ((d <= last_x) ? 1 : 2);
    int j = 0;
    double u = 0.0;
    if (reg == 1) {
      double t = rtf_inv(rkind, rp0, rp1, n, d);
      j = (int)floor(t);
      if (j == n - 1) j = n - 2;
      if (j < 0) j = 0;
      u = t - j;
    }
    double acc = out[i];
    for (int s = 0; s < nspline; ++s) {
      int l = (int)sqrt((double)s);
      double v;
      if (reg == 0) v = amp_left[s] * pow(d, l);
      else if (reg == 2) v = amp_right[s] / pow(d, l + 1);
      else {
        const double *ys = y + (long)s * n, *ds = dt + (long)s * n;
        double z = ys[j + 1] - ys[j];
        v = ys[j] + u * (ds[j] + u * (3 * z - 2 * ds[j] - ds[j + 1] + u * (-2 * z + ds[j] + ds[j + 1])));
      }
      if (s == 0) {
        acc += v * INV_SQRT_4_PI;
        if (lmax > 0) {
          work[0] = zz;
          work[1] = x;
          work[2] = yy;
          if (lmax > 1) pure_polys(work, lmax);
        }
      } else {
        double dpowl = 1.0;
        for (int k = 0; k < l; ++k) dpowl = dpowl / d;
        acc += v * sqrt(2 * l + 1) * INV_SQRT_4_PI * dpowl * work[s - 1];
      }
    }
    out[i] = acc;
  }
}

/* ------------------------------------------------------------------------------
 * Grid contraction nwavefunction_compute_coupling (src/wavefunction/nwavefunction.cpp
 * :77-107): M_ij = sum_p phi_i phi_j v bw w for i <= j, mirrored. The reference
 * materialises all pair rows first; here each pair row is streamed (same sums).
 * phi is function-major [n x N] as the reference passes it.
 * ---------------------------------------------------------------------------- */
void oracle_coupling(const double *phi, const double *v, const double *bw, const double *w,
                     long npoint, int n, double *out) {
  long npair = (long)n * (n + 1) / 2;
#pragma omp parallel for schedule(dynamic, 4)
  for (long q = 0; q < npair; ++q) {
    int i = (int)((sqrt(8.0 * q + 1) - 1) / 2);
    while ((long)i * (i + 1) / 2 > q) --i;
    while ((long)(i + 1) * (i + 2) / 2 <= q) ++i;
    int j = (int)(q - (long)i * (i + 1) / 2);
    const double *a = phi + (long)i * npoint, *b = phi + (long)j * npoint;
    double s = 0.0;
    for (long p = 0; p < npoint; ++p) s += a[p] * b[p] * v[p] * bw[p] * w[p];
    out[(long)j * n + i] = s;
    out[(long)i * n + j] = s;
  }
}

/* Density rho_p = phi_p^T P phi_p per point (napmo/wavefunction/psi_hybrid.py:144-154,
 * a per-point matvec). phi point-major [N x n]. */
void oracle_density(const double *phi, const double *P, long npoint, int n, double *out) {
#pragma omp parallel for schedule(static)
  for (long p = 0; p < npoint; ++p) {
    const double *f = phi + p * n;
    double rho = 0.0;
    for (int i = 0; i < n; ++i) {
      double t = 0.0;
      for (int j = 0; j < n; ++j) t += P[(long)i * n + j] * f[j];
      rho += f[i] * t;
    }
    out[p] = rho;
  }
}

/* Thread count of the OpenMP loops above (the CPU baseline reports 1 and all cores). */
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_max_threads(void) { return omp_get_max_threads(); }
