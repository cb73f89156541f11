// Host-side objects of the drop-in grid API: radial transforms, Lebedev rules,
// radial / atomic grids, splines and extrapolations, Gaussian basis descriptions.
// These mirror the reference's object model (src/horton, src/grid, src/system);
// everything on the hot path that touches grid points runs on the GPU.
#pragma once

#include <cmath>
#include <memory>
#include <vector>

#include "common.hpp"

namespace napmo {

// ------------------------------------------------------------- radial transforms
// Parametrised by kind so kernels can evaluate them (src/horton/rtransform.cpp).
enum RKind { RK_CHEBYSHEV = 0, RK_POWER = 1, RK_IDENTITY = 2 };

struct RTransform {
  RKind kind;
  int npoint;
  double p0 = 0.0;  // Chebyshev: radii;  Power: rmin
  double p1 = 0.0;  // Chebyshev: z;      Power: power
  double p2 = 0.0;  // Power: rmax
  double radius(double t) const;
  double deriv(double t) const;
  double deriv2(double t) const;
  double deriv3(double t) const;
  double inv(double r) const;
  double radii() const { return kind == RK_POWER ? p2 : p0; }
};

RTransform *make_chebyshev(double radii, int npoint);
RTransform *make_power(double rmin, double rmax, int npoint);
RTransform *make_identity(int npoint);

// -------------------------------------------------------------------- Lebedev
struct AngularGrid {
  int lorder;                   // number of points
  std::vector<double> points;   // [lorder x 3]
  std::vector<double> weights;  // [lorder], sum 1
};
AngularGrid *make_angular(int npoint);
bool lebedev_known(int npoint);

// --------------------------------------------------------------- radial, atomic
struct RadialGrid {
  RTransform *rtf;  // borrowed (owned by the caller, as in the reference)
  double radii;
  std::vector<double> points, weights;  // r_i, 4 pi r_i^2 dr/dt
};
RadialGrid *make_radial(RTransform *rtf, double radii);
// RadialGrid::deriv2 (src/grid/radial.cpp:53-192): (2/r) f' + f'' on the Chebyshev grid of
// radius parameter radii, by 7-point (ends) and 9-point finite differences in
// omega = (i + 1) pi / (n + 1); points r [n]; n >= 8.
void radial_deriv2(double radii, const double *points, const double *f, double *out, int n);

struct AtomicGrid {
  AngularGrid *ang;  // borrowed
  RadialGrid *rad;   // borrowed
  double origin[3];
  std::vector<double> points, weights;  // radial-major [nrad*nang x 3], aw*rw
  int size() const { return (int)weights.size(); }
};
AtomicGrid *make_atomic(AngularGrid *ang, RadialGrid *rad, const double *origin);

// ------------------------------------------------------- splines and extrapolation
enum ExtrapKind { EX_ZERO = 0, EX_CUSP = 1, EX_POWER = 2, EX_POTENTIAL = 3 };
struct CubicSpline;
struct Extrapolation {
  ExtrapKind kind;
  double power = 0.0;  // PowerExtrapolation
  long l = 0;          // PotentialExtrapolation
  double a0 = 0, b0 = 0, x0 = 0, amp = 0, amp_left = 0, amp_right = 0;
  void prepare(const CubicSpline &cs);
  double eval_left(double x) const;
  double eval_right(double x) const;
  double deriv_left(double x) const;
  double deriv_right(double x) const;
  // power-law tails (src/horton/extrapolation.h: Power and Potential kinds)
  bool has_tail() const { return kind == EX_POWER || kind == EX_POTENTIAL; }
};

struct CubicSpline {
  Extrapolation *ext;
  RTransform *rtf;
  double first_x, last_x;
  double *y, *dt;  // aliased caller buffers (src/horton/cubic_spline.cpp:138-145)
  int n;
  void eval(const double *x, double *out, int m) const;
  // dy/dr: the interpolant's t-derivative divided by dr/dt (src/horton/cubic_spline.cpp:183-214)
  void eval_deriv(const double *x, double *out, int m) const;
};

void spline_derivatives(const double *y, double *dt, int n);
// Quadrature weights of the natural cubic spline through (t, y_t), t = 0..n-1: the integral
// of the spline of the unit vector e_i (src/horton/cubic_spline.cpp:105-131 semantics).
void cubic_spline_int_weights(double *weights, int n);

// ---------------------------------------------------------------- Gaussian basis
struct Primitive {
  int l[3];
  double origin[3];
  double zeta, coeff, norma;
  // coeff * norma * dx^lx dy^ly dz^lz exp(-zeta r^2) at r (src/system/gto.cpp:41-51)
  double value(const double *r) const;
};
struct Contraction {
  double origin[3];
  int l[3];
  std::vector<Primitive> prim;
  double norma;
  double value(const double *r) const;  // norma * sum of the primitives
};
struct BasisSet {
  std::vector<Contraction> cont;
  int max_l() const;      // largest lx + ly + lz over the contractions (0 when empty)
  int max_nprim() const;  // largest primitive count over the contractions (0 when empty)
};
Primitive *make_primitive(const int *l, const double *origin, double zeta, double coeff);
Contraction *make_contraction(Primitive **prims, int n);
double primitive_overlap(const Primitive &a, const Primitive &b);

// Flattened basis (device-friendly): one entry per contracted function.
struct FlatBasis {
  int nbasis = 0;
  std::vector<double> origin;   // [3n]
  std::vector<int> lxyz;        // [3n]
  std::vector<int> first, nprim;
  std::vector<double> norm;     // contraction norm
  std::vector<double> zeta, coef;  // per primitive, coef = c * primitive norm
};
FlatBasis flatten(const BasisSet &b);

// Real solid harmonics of a unit vector in the reference order
// [C_l0, C_l1, S_l1, ..., C_ll, S_ll] for l = 0..lmax (src/horton/moments.cpp:59-125).
void solid_harmonics(double x, double y, double z, int lmax, double *out);

// Polar and azimuthal angles of a unit vector: theta = acos(z), phi from x / sqrt(x^2 + y^2)
// with the sign of y (src/grid/sphere_lebedev_rule.cpp:6718-6760).
void unit_to_angles(double x, double y, double z, double *theta, double *phi);
// Real spherical harmonics Y_lm(theta, phi) for l = 0..lmax in the order m = -l..l per l:
// m = 0: N_l0 P_l(cos theta); m > 0: sqrt2 N_lm P_l^m cos(m phi); m < 0: sqrt2 N_l|m| P_l^|m|
// sin(|m| phi), with N_lm P_l^m the orthonormalised associated Legendre function including
// the Condon-Shortley phase (-1)^m (src/utils/spherical_harmonics.cpp:11-27, GSL sphPlm).
void real_ylm(double theta, double phi, int lmax, double *out);

}  // namespace napmo
