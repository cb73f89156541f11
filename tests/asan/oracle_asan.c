/* Address/undefined-behaviour sanitizer driver for the oracle's C restatement
 * (oracle/napmo_oracle.c): every entry point on small inputs. Built with
 * gcc -fsanitize=address,undefined by tests/test_asan_host.py. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

void oracle_build_ode2(const double *b, const double *a, const double *f, const double *bcs,
                       double *coeffs, double *rhs, long npoint, const double *tab2,
                       const double *tab3);
void oracle_becke_weights(const double *points, long npoint, const double *origin,
                          const double *radii, int ncenter, const long *offsets, double *out);
void oracle_basis_compute(const double *points, long npoint, int nbasis, const double *fn_origin,
                          const int *fn_l, const int *fn_nprim, const int *fn_first,
                          const double *fn_norm, const double *prim_zeta, const double *prim_coef,
                          double *out);
void oracle_eval_decomposition(const double *y, const double *dt, const double *amp_left,
                               const double *amp_right, int n, int nspline, int rkind, double rp0,
                               double rp1, double first_x, double last_x, const double *center,
                               const double *points, long npoint, double *out);
void oracle_coupling(const double *phi, const double *v, const double *bw, const double *w,
                     long npoint, int n, double *out);
void oracle_density(const double *phi, const double *P, long npoint, int n, double *out);
void oracle_set_threads(int n);

static double urand(unsigned *s) {
  *s = *s * 1664525u + 1013904223u;
  return (double)(*s >> 8) / 16777216.0;
}

int main(void) {
  unsigned seed = 7;
  oracle_set_threads(2);
  /* Hermite finite-element system of 12 nodes */
  long np = 12, nfn = 2 * np;
  double b[24], a[24], f[24], bcs[4] = {0.0, NAN, 0.0, NAN}, t2[64], t3[512];
  for (int i = 0; i < 24; ++i) b[i] = urand(&seed), a[i] = urand(&seed), f[i] = urand(&seed);
  for (int i = 0; i < 64; ++i) t2[i] = urand(&seed);
  for (int i = 0; i < 512; ++i) t3[i] = urand(&seed);
  double *coeffs = calloc((size_t)nfn * nfn, sizeof(double)), rhs[24] = {0};
  oracle_build_ode2(b, a, f, bcs, coeffs, rhs, np, t2, t3);
  /* Becke weights: 3 centres with uneven point counts */
  long offsets[4] = {0, 40, 70, 130};
  long N = offsets[3];
  double *pts = malloc(sizeof(double) * 3 * N), *bw = malloc(sizeof(double) * N);
  double origin[9] = {0, 0, 0, 1.4, 0, 0, 0, 1.7, 0.3}, radii[3] = {0.6, 1.2, 0.9};
  for (long p = 0; p < 3 * N; ++p) pts[p] = 4.0 * urand(&seed) - 2.0;
  oracle_becke_weights(pts, N, origin, radii, 3, offsets, bw);
  for (long p = 0; p < N; ++p)
    if (!(bw[p] >= 0.0 && bw[p] <= 1.0)) { fprintf(stderr, "becke weight out of [0, 1]\n"); return 1; }
  /* basis: 4 contractions of 2 primitives */
  int nb = 4, fl[12] = {0, 0, 0, 1, 0, 0, 0, 1, 1, 2, 0, 0}, fnp[4] = {2, 2, 2, 2}, ff[4] = {0, 2, 4, 6};
  double fo[12], fnorm[4] = {1, 1, 1, 1}, pz[8], pc[8];
  for (int i = 0; i < 12; ++i) fo[i] = urand(&seed);
  for (int i = 0; i < 8; ++i) pz[i] = 0.2 + 3.0 * urand(&seed), pc[i] = urand(&seed);
  double *phi = malloc(sizeof(double) * N * nb);
  oracle_basis_compute(pts, N, nb, fo, fl, fnp, ff, fnorm, pz, pc, phi);
  /* density and coupling */
  double P[16], rho[130], v[130], w[130], F[16], *phiT = malloc(sizeof(double) * N * nb);
  for (int i = 0; i < 16; ++i) P[i] = urand(&seed);
  for (long p = 0; p < N; ++p) v[p] = urand(&seed), w[p] = urand(&seed);
  for (long p = 0; p < N; ++p)
    for (int m = 0; m < nb; ++m) phiT[(long)m * N + p] = phi[p * nb + m];
  oracle_density(phi, P, N, nb, rho);
  oracle_coupling(phiT, v, bw, w, N, nb, F);
  /* synthesis of one centre's splines (lmax 3: 16 splines on a 20-point Chebyshev grid) */
  int n = 20, ns = 16;
  double *y = malloc(sizeof(double) * ns * n), *dt = malloc(sizeof(double) * ns * n), aL[16], aR[16];
  for (int i = 0; i < ns * n; ++i) y[i] = urand(&seed), dt[i] = urand(&seed);
  for (int i = 0; i < ns; ++i) aL[i] = urand(&seed), aR[i] = urand(&seed);
  double out[130] = {0}, c[3] = {0.1, 0.2, 0.3};
  double rp1 = M_PI / (n + 1), first = 0.6 * (1 + cos(n * rp1)) / (1 - cos(n * rp1));
  double last = 0.6 * (1 + cos(rp1)) / (1 - cos(rp1));
  oracle_eval_decomposition(y, dt, aL, aR, n, ns, 0, 0.6, rp1, first, last, c, pts, N, out);
  for (long p = 0; p < N; ++p)
    if (!isfinite(out[p])) { fprintf(stderr, "non-finite synthesis\n"); return 1; }
  free(coeffs), free(pts), free(bw), free(phi), free(phiT), free(y), free(dt);
  printf("oracle asan driver: ok\n");
  return 0;
}
