// Address/undefined-behaviour sanitizer driver for the library's host code (host.cpp):
// every host routine of the grid API on small inputs, with light checks of the results.
// Built with g++ -fsanitize=address,undefined by tests/test_asan_host.py; no device code.
#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

#include "host.hpp"

using namespace napmo;

static int failures = 0;
static void check(bool ok, const char *what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++failures;
  }
}

int main() {
  // radial transforms: radius / inverse round trip, derivative by central difference
  std::unique_ptr<RTransform> tfs[3] = {std::unique_ptr<RTransform>(make_chebyshev(1.1, 40)),
                                        std::unique_ptr<RTransform>(make_power(1e-4, 20.0, 40)),
                                        std::unique_ptr<RTransform>(make_identity(40))};
  for (auto &t : tfs)
    for (int i = 1; i < 39; ++i) {
      const double r = t->radius(i);
      check(std::fabs(t->inv(r) - i) < 1e-8 * (1 + i), "rtransform inv(radius(t)) = t");
      const double d = (t->radius(i + 1e-5) - t->radius(i - 1e-5)) / 2e-5;
      check(std::fabs(d - t->deriv(i)) < 1e-5 * (1 + std::fabs(d)), "rtransform deriv");
      (void)t->deriv2(i), (void)t->deriv3(i);
    }
  // Lebedev rules: weights sum to one, points on the unit sphere
  for (int npt : {6, 14, 26, 110, 194, 302, 590, 1202, 5810}) {
    check(lebedev_known(npt), "lebedev size known");
    std::unique_ptr<AngularGrid> a(make_angular(npt));
    double s = 0.0;
    for (int k = 0; k < a->lorder; ++k) {
      s += a->weights[k];
      const double *p = &a->points[3 * k];
      check(std::fabs(p[0] * p[0] + p[1] * p[1] + p[2] * p[2] - 1.0) < 1e-12, "lebedev point norm");
    }
    check(std::fabs(s - 1.0) < 1e-12, "lebedev weights sum");
  }
  // radial and atomic grids: integral of exp(-r^2) = pi^(3/2)
  std::unique_ptr<RadialGrid> rad(make_radial(tfs[0].get(), 1.1));
  std::unique_ptr<AngularGrid> ang(make_angular(110));
  const double origin[3] = {0.1, -0.2, 0.3};
  std::unique_ptr<AtomicGrid> at(make_atomic(ang.get(), rad.get(), origin));
  double integral = 0.0;
  for (int p = 0; p < at->size(); ++p) {
    const double *x = &at->points[3 * p];
    const double r2 = (x[0] - origin[0]) * (x[0] - origin[0]) + (x[1] - origin[1]) * (x[1] - origin[1]) +
                      (x[2] - origin[2]) * (x[2] - origin[2]);
    integral += at->weights[p] * std::exp(-r2);
  }
  check(std::fabs(integral - std::pow(M_PI, 1.5)) < 1e-6, "atomic grid integral");
  // deriv2 of exp(-r^2) on the radial grid
  const int n = rad->points.size();
  std::vector<double> f(n), d2(n);
  for (int i = 0; i < n; ++i) f[i] = std::exp(-rad->points[i] * rad->points[i]);
  radial_deriv2(1.1, rad->points.data(), f.data(), d2.data(), n);
  // splines with every extrapolation kind, evaluated inside, left and right of the grid
  std::vector<double> y(n), dt(n), xs = {1e-6, 1e-3, 0.5, 1.0, 3.0, 50.0, 1e3}, out(xs.size()),
                                        der(xs.size());
  for (int i = 0; i < n; ++i) y[i] = std::exp(-0.5 * rad->points[i]);
  spline_derivatives(y.data(), dt.data(), n);
  for (int kind = 0; kind < 4; ++kind) {
    Extrapolation ex{(ExtrapKind)kind};
    ex.power = -2.0;
    ex.l = 2;
    CubicSpline cs{&ex, tfs[0].get(), tfs[0]->radius(0), tfs[0]->radius(n - 1), y.data(), dt.data(), n};
    ex.prepare(cs);
    cs.eval(xs.data(), out.data(), (int)xs.size());
    cs.eval_deriv(xs.data(), der.data(), (int)xs.size());
    for (double v : out) check(std::isfinite(v), "spline value finite");
    (void)ex.has_tail();
  }
  std::vector<double> wts(n);
  cubic_spline_int_weights(wts.data(), n);
  // Gaussian basis: values, overlaps, flattening
  BasisSet basis;
  const int ls[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 1}, {2, 0, 0}};
  for (int c = 0; c < 4; ++c) {
    std::vector<Primitive *> prims;
    for (double z : {0.3, 2.0, 11.0}) prims.push_back(make_primitive(ls[c], origin, z, 0.5));
    std::unique_ptr<Contraction> con(make_contraction(prims.data(), (int)prims.size()));
    for (auto *p : prims) {
      check(std::fabs(primitive_overlap(*p, *p) - p->coeff * p->coeff) < 1e-10, "primitive norm");
      const double r[3] = {0.3, 0.1, -0.4};
      check(std::isfinite(p->value(r)), "primitive value");
      delete p;
    }
    const double r[3] = {0.2, 0.0, 0.1};
    check(std::isfinite(con->value(r)), "contraction value");
    basis.cont.push_back(*con);
  }
  check(basis.max_l() == 2 && basis.max_nprim() == 3, "basis maxima");
  FlatBasis fb = flatten(basis);
  check(fb.nbasis == 4, "flattened basis size");
  // harmonics
  std::vector<double> sh(13 * 13), ylm(13 * 13);
  solid_harmonics(0.36, 0.48, 0.8, 12, sh.data());
  double th, ph;
  unit_to_angles(0.36, 0.48, 0.8, &th, &ph);
  real_ylm(th, ph, 12, ylm.data());
  check(std::fabs(ylm[0] - 0.5 / std::sqrt(M_PI)) < 1e-14, "Y00");
  if (failures) return 1;
  std::printf("host asan driver: ok\n");
  return 0;
}
