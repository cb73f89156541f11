// extern "C" surface of libnapmo_gpu.so (declared in include/napmo_gpu.h).
#include "napmo_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "common.hpp"
#include "apmo_plan.hpp"
#include "batch_plan.hpp"
#include "fock_plan.hpp"
#include "host.hpp"
#include "kernels.hpp"
#include "legacy_kernels.hpp"
#include "poisson_plan.hpp"

namespace napmo {

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
void clear_error() { g_error.clear(); }

size_t max_lds_per_block() {
  static thread_local long cached = -1;
  if (cached < 0) {
    int dev = 0, v = 0;
    NAPMO_HIP(hipGetDevice(&dev));
    NAPMO_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    cached = v;
  }
  return (size_t)cached;
}

void require_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    throw Error("napmo_gpu: no HIP device available (this library has no CPU path)");
}

// Legacy entry points return values, not status codes: report and abort on failure
// like the reference (whose C++ exceptions terminate), but with a readable message.
template <class F> auto legacy(const char *name, F &&f) -> decltype(f()) {
  try {
    clear_error();
    return f();
  } catch (const std::exception &e) {
    set_error(e.what());
    std::fprintf(stderr, "napmo_gpu: %s: %s\n", name, e.what());
    std::fflush(stderr);
    std::abort();
  }
}


static AtomInfo atom_info(const AtomicGrid &g) {
  AtomInfo a;
  std::memcpy(a.origin, g.origin, sizeof(a.origin));
  a.radii = g.rad->radii;
  a.nrad = (int)g.rad->points.size();
  a.nang = g.ang->lorder;
  a.rtf = *g.rad->rtf;
  a.r = g.rad->points;
  a.rw = g.rad->weights;
  a.u = g.ang->points;
  a.aw = g.ang->weights;
  return a;
}

// Host points (AoS [n x 3]) to device SoA.
struct DevicePoints {
  DeviceBuffer<double> x, y, z;
  void upload(const double *pts, long n, hipStream_t s) {
    std::vector<double> hx(n), hy(n), hz(n);
    for (long i = 0; i < n; ++i) hx[i] = pts[3 * i], hy[i] = pts[3 * i + 1], hz[i] = pts[3 * i + 2];
    x.upload(hx, s), y.upload(hy, s), z.upload(hz, s);
    NAPMO_HIP(hipStreamSynchronize(s));
  }
};

// A stream for one-shot legacy calls that are not tied to a grid handle.
static hipStream_t scratch_stream() {
  static thread_local hipStream_t s = nullptr;
  if (!s) {
    require_device();
    NAPMO_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  }
  return s;
}

// F = Phi^T diag(s) Y for host function-major Phi / Y [n x N] (legacy contractions).
static void host_contract(const DeviceGrid &g, const double *phi, const double *yfm,
                          const double *vec, int n, bool upper_mirror, double scale, double *out) {
  hipStream_t st = g.stream;
  const long N = g.N, Npad = round_up(N, TILE);
  const int ld = (int)round_up(n, TILE);
  auto to_point_major = [&](const double *fm, DeviceBuffer<double> &dst) {
    std::vector<double> h((size_t)Npad * ld, 0.0);
    for (int i = 0; i < n; ++i)
      for (long p = 0; p < N; ++p) h[(size_t)p * ld + i] = fm[(size_t)i * N + p];
    dst.upload(h, st);
  };
  DeviceBuffer<double> X, Y, v, s, part, F;
  to_point_major(phi, X);
  if (yfm) to_point_major(yfm, Y);
  s.resize(Npad);
  if (vec) {
    v.upload(vec, N, st);
    launch_scale3(v.get(), g.wbw.get(), nullptr, N, Npad, s.get(), st);
  } else {
    launch_scale3(g.wbw.get(), g.wbw.get(), nullptr, 0, Npad, s.get(), st);  // zeros
    launch_scale3(g.w.get(), g.bw.get(), nullptr, N, Npad, s.get(), st);
  }
  FockShape sh = fock_shape(ld, Npad, yfm == nullptr);
  part.resize((size_t)sh.nchunk * sh.cstride());
  F.resize((size_t)n * n);
  launch_fock_f64(sh, X.get(), yfm ? Y.get() : X.get(), s.get(), part.get(), st);
  launch_fock_reduce(sh, part.get(), n, upper_mirror ? 1 : 0, scale, 0, F.get(), st);
  F.download(out, (size_t)n * n, st);
  NAPMO_HIP(hipStreamSynchronize(st));
}

}  // namespace napmo

using namespace napmo;

extern "C" {

const char *napmo_last_error(void) { return g_error.c_str(); }
void napmo_free(void *ptr) { std::free(ptr); }
int napmo_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// ------------------------------------------------------------- radial transforms
void *ChebyshevRTransform_new(double radii, int npoint) {
  return legacy(__func__, [&] { return (void *)make_chebyshev(radii, npoint); });
}
double ChebyshevRTransform_get_radii(void *r) { return ((RTransform *)r)->p0; }
double ChebyshevRTransform_get_z(void *r) { return ((RTransform *)r)->p1; }
void *PowerRTransform_new(double rmin, double rmax, int npoint) {
  return legacy(__func__, [&] { return (void *)make_power(rmin, rmax, npoint); });
}
double PowerRTransform_get_rmin(void *r) { return ((RTransform *)r)->p0; }
double PowerRTransform_get_rmax(void *r) { return ((RTransform *)r)->p2; }
double PowerRTransform_get_power(void *r) { return ((RTransform *)r)->p1; }
void *IdentityRTransform_new(int npoint) {
  return legacy(__func__, [&] { return (void *)make_identity(npoint); });
}
int RTransform_get_npoint(void *r) { return ((RTransform *)r)->npoint; }
double RTransform_radius(void *r, double t) { return ((RTransform *)r)->radius(t); }
double RTransform_deriv(void *r, double t) { return ((RTransform *)r)->deriv(t); }
double RTransform_deriv2(void *r, double t) { return ((RTransform *)r)->deriv2(t); }
double RTransform_deriv3(void *r, double t) { return ((RTransform *)r)->deriv3(t); }
double RTransform_inv(void *r, double x) { return ((RTransform *)r)->inv(x); }
#define NAPMO_RTF_ARRAY(name, method)                                          \
  void name(void *r, double *in, double *out, int n) {                         \
    for (int i = 0; i < n; ++i) out[i] = ((RTransform *)r)->method(in[i]);     \
  }
NAPMO_RTF_ARRAY(RTransform_radius_array, radius)
NAPMO_RTF_ARRAY(RTransform_deriv_array, deriv)
NAPMO_RTF_ARRAY(RTransform_deriv2_array, deriv2)
NAPMO_RTF_ARRAY(RTransform_deriv3_array, deriv3)
NAPMO_RTF_ARRAY(RTransform_inv_array, inv)
void RTransform_del(void *r) { delete (RTransform *)r; }

// ------------------------------------------------------------------------ grids
void *AngularGrid_new(int lorder) {
  return legacy(__func__, [&] { return (void *)make_angular(lorder); });
}
int AngularGrid_get_lorder(void *g) { return ((AngularGrid *)g)->lorder; }
double *AngularGrid_get_points(void *g) { return ((AngularGrid *)g)->points.data(); }
double *AngularGrid_get_weights(void *g) { return ((AngularGrid *)g)->weights.data(); }
double AngularGrid_integrate(void *g, int nfunc, double *f) {
  // src/grid/angular.cpp: 4 pi sum_j prod_k f_k(j) aw_j
  AngularGrid *a = (AngularGrid *)g;
  double s = 0.0;
  for (int j = 0; j < a->lorder; ++j) {
    double t = f[j];
    for (int k = 1; k < nfunc; ++k) t *= f[(size_t)k * a->lorder + j];
    s += t * a->weights[j];
  }
  return s * 4.0 * M_PI;
}
void AngularGrid_del(void *g) { delete (AngularGrid *)g; }
void AngularGrid_spherical(void *g) {
  // points become (1, theta, phi) in place (src/grid/angular.cpp:39-48)
  AngularGrid *a = (AngularGrid *)g;
  for (int j = 0; j < a->lorder; ++j) {
    double *u = &a->points[3 * (size_t)j];
    double t, p;
    unit_to_angles(u[0], u[1], u[2], &t, &p);
    u[0] = 1.0, u[1] = t, u[2] = p;
  }
}
void AngularGrid_spherical_expansion(void *g, int lmax, int size_f, double *f, double *out) {
  // out[k][lm] = 4 pi sum_j aw_j f[k][j] Y_lm(u_j), lm in the order m = -l..l per l
  // (src/grid/angular.cpp:50-93)
  legacy(__func__, [&] {
    NAPMO_REQUIRE(lmax >= 0, "lmax can not be negative.");
    AngularGrid *a = (AngularGrid *)g;
    const int L = (lmax + 1) * (lmax + 1), na = a->lorder;
    std::vector<double> Y((size_t)na * L);
    for (int j = 0; j < na; ++j) {
      const double *u = &a->points[3 * (size_t)j];
      double t, p;
      unit_to_angles(u[0], u[1], u[2], &t, &p);
      real_ylm(t, p, lmax, &Y[(size_t)j * L]);
    }
    for (int k = 0; k < size_f; ++k)
      for (int lm = 0; lm < L; ++lm) {
        double acc = 0.0;
        for (int j = 0; j < na; ++j) acc += f[(size_t)k * na + j] * a->weights[j] * Y[(size_t)j * L + lm];
        out[(size_t)k * L + lm] = acc * 4.0 * M_PI;
      }
    return 0;
  });
}
void AngularGrid_eval_expansion(void *g, int lmax, int size_f, double *decomposition, double *out) {
  // out[k][j] = sum_lm decomposition[k][lm] Y_lm(u_j) (src/grid/angular.cpp:95-139)
  legacy(__func__, [&] {
    NAPMO_REQUIRE(lmax >= 0, "lmax can not be negative.");
    AngularGrid *a = (AngularGrid *)g;
    const int L = (lmax + 1) * (lmax + 1), na = a->lorder;
    std::vector<double> Y(L);
    for (int j = 0; j < na; ++j) {
      const double *u = &a->points[3 * (size_t)j];
      double t, p;
      unit_to_angles(u[0], u[1], u[2], &t, &p);
      real_ylm(t, p, lmax, Y.data());
      for (int k = 0; k < size_f; ++k) {
        double acc = 0.0;
        for (int lm = 0; lm < L; ++lm) acc += decomposition[(size_t)k * L + lm] * Y[lm];
        out[(size_t)k * na + j] = acc;
      }
    }
    return 0;
  });
}

void *RadialGrid_new(void *rtf, double radii) {
  return legacy(__func__, [&] { return (void *)make_radial((RTransform *)rtf, radii); });
}
int RadialGrid_get_size(void *g) { return (int)((RadialGrid *)g)->points.size(); }
double RadialGrid_get_radii(void *g) { return ((RadialGrid *)g)->radii; }
double *RadialGrid_get_points(void *g) { return ((RadialGrid *)g)->points.data(); }
double *RadialGrid_get_weights(void *g) { return ((RadialGrid *)g)->weights.data(); }
double *RadialGrid_deriv2(void *g, double *f) {
  return legacy(__func__, [&] {
    RadialGrid *rg = (RadialGrid *)g;
    const int n = (int)rg->points.size();
    double *out = (double *)std::malloc(sizeof(double) * std::max(n, 1));
    NAPMO_REQUIRE(out != nullptr, "RadialGrid_deriv2: out of host memory");
    radial_deriv2(rg->radii, rg->points.data(), f, out, n);
    return out;
  });
}
double RadialGrid_integrate(void *g, int segments, double *f) {
  // src/grid/radial.cpp:31-51
  RadialGrid *r = (RadialGrid *)g;
  const size_t n = r->points.size();
  double out = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double t = f[i];
    for (int k = 1; k < segments; ++k) t *= f[k * n + i];
    out += t * r->weights[i];
  }
  return out;
}
void RadialGrid_del(void *g) { delete (RadialGrid *)g; }

void *AtomicGrid_new(void *ang, void *rad, double *origin) {
  return legacy(__func__,
                [&] { return (void *)make_atomic((AngularGrid *)ang, (RadialGrid *)rad, origin); });
}
int AtomicGrid_get_size(void *g) { return ((AtomicGrid *)g)->size(); }
double AtomicGrid_get_radii(void *g) { return ((AtomicGrid *)g)->rad->radii; }
double *AtomicGrid_get_origin(void *g) { return ((AtomicGrid *)g)->origin; }
double *AtomicGrid_get_points(void *g) { return ((AtomicGrid *)g)->points.data(); }
double *AtomicGrid_get_weights(void *g) { return ((AtomicGrid *)g)->weights.data(); }
void AtomicGrid_del(void *g) { delete (AtomicGrid *)g; }

double *AtomicGrid_integrate(void *g, int nfunc, int segments, int size, double *f) {
  // src/grid/atomic_grid.cpp:53-80: per segment sum of the product of nfunc arrays
  const long total = (long)size * segments;
  double *out = (double *)std::calloc(segments, sizeof(double));
  for (int s = 0; s < segments; ++s)
    for (int j = 0; j < size; ++j) {
      const long i = (long)s * size + j;
      double t = f[i];
      for (int k = 1; k < nfunc; ++k) t *= f[k * total + i];
      out[s] += t;
    }
  return out;
}

double *AtomicGrid_spherical_expansion(void *g, double *f, int lmax) {
  // src/grid/atomic_grid.cpp:82-136 on the GPU (projection kernel on a one-centre grid)
  return legacy(__func__, [&] {
    NAPMO_REQUIRE(lmax >= 0, "lmax can not be negative.");
    AtomicGrid *ag = (AtomicGrid *)g;
    DeviceGrid dg;
    dg.build({atom_info(*ag)});
    DeviceBuffer<double> dens, ones;
    dens.upload(f, dg.N, dg.stream);
    // spherical_expansion projects f itself (no Becke weight): use bw == 1
    std::vector<double> one(dg.N, 1.0);
    dg.bw.upload(one, dg.stream);
    PoissonPlan &pl = dg.plan(lmax);
    AtomTables at = dg.tables(pl);
    launch_project(at, dens.get(), dg.bw.get(), pl.yhat.get(), lmax, pl.work.rho_lm.get(), dg.stream);
    const int L = (lmax + 1) * (lmax + 1);
    double *out = (double *)std::malloc(sizeof(double) * dg.atoms[0].nrad * L);
    pl.work.rho_lm.download(out, (size_t)dg.atoms[0].nrad * L, dg.stream);
    NAPMO_HIP(hipStreamSynchronize(dg.stream));
    return out;
  });
}

void *BeckeGrid_new(void **atgrids, int ncenter, int ablmax, double abldep) {
  return legacy(__func__, [&] {
    std::vector<AtomInfo> atoms;
    for (int i = 0; i < ncenter; ++i) atoms.push_back(atom_info(*(AtomicGrid *)atgrids[i]));
    auto *h = new BeckeHandle;
    h->ablmax = ablmax;
    h->abldep = abldep;
    h->grid.build(atoms);
    return (void *)h;
  });
}
void *BeckeGrid_from_points(double *pw, int size, int ncenter) {
  // src/grid/becke.cpp:57-77 (Lowdin2 grids: rows x, y, z, Becke weight; weights 1)
  return legacy(__func__, [&] {
    auto *h = new BeckeHandle;
    try {
      h->grid.build_points(pw, size, ncenter);
    } catch (...) {
      delete h;
      throw;
    }
    return (void *)h;
  });
}

int BeckeGrid_get_ncenter(void *g) { return ((BeckeHandle *)g)->grid.M; }
int BeckeGrid_get_size(void *g) { return (int)((BeckeHandle *)g)->grid.N; }
double *BeckeGrid_get_radii(void *g) { return ((BeckeHandle *)g)->grid.radii.data(); }
double *BeckeGrid_get_points(void *g) { return ((BeckeHandle *)g)->grid.points.data(); }
double *BeckeGrid_get_weights(void *g) { return ((BeckeHandle *)g)->grid.weights.data(); }
double *BeckeGrid_get_origin(void *g) { return ((BeckeHandle *)g)->grid.origin.data(); }
double *BeckeGrid_get_becke_weights(void *g) { return ((BeckeHandle *)g)->grid.becke.data(); }
double BeckeGrid_integrate(void *g, double *f) {
  // src/grid/becke.cpp:178-184
  DeviceGrid &d = ((BeckeHandle *)g)->grid;
  double s = 0.0;
  for (long p = 0; p < d.N; ++p) s += f[p] * d.becke[p] * d.weights[p];
  return s;
}
void BeckeGrid_del(void *g) { delete (BeckeHandle *)g; }

void *AuxiliaryBasis_new(void *grid) {
  return legacy(__func__, [&] {
    BeckeHandle *h = (BeckeHandle *)grid;
    return (void *)make_aux_basis(h->grid, h->ablmax, h->abldep);
  });
}
void AuxiliaryBasis_del(void *ab) { delete (AuxBasis *)ab; }
void AuxiliaryBasis_compute_aobasis(void *ab) { (void)ab; }  // computed by AuxiliaryBasis_new
double *AuxiliaryBasis_get_aobasis(void *ab) { return ((AuxBasis *)ab)->aobasis.data(); }
int AuxiliaryBasis_get_nao(void *ab) { return ((AuxBasis *)ab)->nao; }
int AuxiliaryBasis_get_ncenter(void *ab) { return ((AuxBasis *)ab)->ncenter; }

// ----------------------------------------------------- splines, extrapolation, ODE
void solve_cubic_spline_system(double *y, double *dt, int npoint) {
  legacy(__func__, [&] {
    spline_derivatives(y, dt, npoint);
    return 0;
  });
}
void *CubicSpline_new(double *y, double *dt, void *ext, void *rtf, int n) {
  return legacy(__func__, [&] {
    RTransform *r = (RTransform *)rtf;
    NAPMO_REQUIRE(r && r->npoint == n, "CubicSpline: transform size mismatch");
    auto *cs = new CubicSpline{(Extrapolation *)ext, r, r->radius(0), r->radius(n - 1), y, dt, n};
    cs->ext->prepare(*cs);
    return (void *)cs;
  });
}
void CubicSpline_eval(void *s, double *x, double *out, int n) { ((CubicSpline *)s)->eval(x, out, n); }
void CubicSpline_eval_deriv(void *s, double *x, double *out, int n) {
  ((CubicSpline *)s)->eval_deriv(x, out, n);
}
void *CubicSpline_get_rtransform(void *s) { return ((CubicSpline *)s)->rtf; }
void *CubicSpline_get_extrapolation(void *s) { return ((CubicSpline *)s)->ext; }
void compute_cubic_spline_int_weights(double *weights, int npoint) {
  legacy(__func__, [&] {
    cubic_spline_int_weights(weights, npoint);
    return 0;
  });
}
double CubicSpline_get_first_x(void *s) { return ((CubicSpline *)s)->first_x; }
double CubicSpline_get_last_x(void *s) { return ((CubicSpline *)s)->last_x; }
void CubicSpline_del(void *s) { delete (CubicSpline *)s; }

void *PotentialExtrapolation_new(int l) {
  return legacy(__func__, [&] {
    NAPMO_REQUIRE(l >= 0, "The argument l cannot be negative.");
    auto *e = new Extrapolation{EX_POTENTIAL};
    e->l = l;
    return (void *)e;
  });
}
void *ZeroExtrapolation_new(void) { return new Extrapolation{EX_ZERO}; }
void *CuspExtrapolation_new(void) { return new Extrapolation{EX_CUSP}; }
void *PowerExtrapolation_new(double power) {
  auto *e = new Extrapolation{EX_POWER};
  e->power = power;
  return e;
}
void Extrapolation_del(void *e) { delete (Extrapolation *)e; }
void Extrapolation_prepare(void *e, void *cs) { ((Extrapolation *)e)->prepare(*(CubicSpline *)cs); }
double Extrapolation_eval_left(void *e, double x) { return ((Extrapolation *)e)->eval_left(x); }
double Extrapolation_eval_right(void *e, double x) { return ((Extrapolation *)e)->eval_right(x); }
double Extrapolation_deriv_left(void *e, double x) { return ((Extrapolation *)e)->deriv_left(x); }
double Extrapolation_deriv_right(void *e, double x) { return ((Extrapolation *)e)->deriv_right(x); }
bool Extrapolation_has_tail(void *e) { return ((Extrapolation *)e)->has_tail(); }
double PowerExtrapolation_get_power(void *e) { return ((Extrapolation *)e)->power; }

void *Cell_new(double *rvecs, int nvec) {
  (void)rvecs;
  return legacy(__func__, [&] {
    NAPMO_REQUIRE(nvec == 0, "only the non-periodic cell (nvec = 0) is supported");
    return (void *)new int(0);
  });
}
void Cell_del(void *c) { delete (int *)c; }

void build_ode2(double *b, double *a, double *f, double *bcs, double *coeffs, double *rhs,
                long npoint) {
  legacy(__func__, [&] {
    fe_assemble(b, a, f, bcs, coeffs, rhs, npoint);
    return 0;
  });
}

void eval_decomposition_grid(void **splines, double *center, double *output, double *points,
                             void *cell, long nspline, long npoint) {
  (void)cell;
  legacy(__func__, [&] {
    const long lmax = (long)std::sqrt((double)nspline) - 1;
    NAPMO_REQUIRE((lmax + 1) * (lmax + 1) == nspline,
                  "The number of splines does not match a well-defined lmax.");
    CubicSpline **sp = (CubicSpline **)splines;
    decomposition_on_points(sp, (int)nspline, (int)lmax, center, points, npoint, output);
    return 0;
  });
}

void dot_multi_moments(long npoint, long nvector, double *data, double *points, double *center,
                       long lmax, long mtype, long *segments, double *output, long nmoment) {
  legacy(__func__, [&] {
    // the reference's checks (src/horton/dot_multi.cpp:152-157)
    NAPMO_REQUIRE(lmax >= 0, "lmax can not be negative.");
    NAPMO_REQUIRE(mtype >= 1 && mtype <= 4, "mtype should be 1, 2, 3 or 4.");
    NAPMO_REQUIRE(lmax <= MOMENT_LMAX, "dot_multi_moments: lmax above 16 is not supported");
    NAPMO_REQUIRE(nvector >= 1, "dot_multi_moments: nvector must be positive");
    if (npoint <= 0) return 0;
    // Segment ranges as the reference's shift_segment walk assigns them (dot_multi.cpp:46-52):
    // one shift at most per point, so after an empty segment every later point stays in it.
    std::vector<long> st, sp, ix;
    long k = 0, end = segments[0], pos = 0;
    while (pos < npoint) {
      if (pos == end) end += segments[++k];
      const long stop = end > pos ? std::min(end, npoint) : npoint;
      st.push_back(pos), sp.push_back(stop), ix.push_back(k);
      pos = stop;
    }
    const long nout = (k + 1) * nmoment;
    hipStream_t s = scratch_stream();
    DeviceBuffer<long> dst, dsp, dix;
    DeviceBuffer<double> dd, dp, dout;
    dst.upload(st, s), dsp.upload(sp, s), dix.upload(ix, s);
    dd.upload(data, (size_t)(nvector * npoint), s);
    dp.upload(points, (size_t)(3 * npoint), s);
    dout.upload(output, (size_t)nout, s);
    launch_dot_multi_moments(dst.get(), dsp.get(), dix.get(), (int)st.size(), npoint, nvector,
                             dd.get(), dp.get(), center, (int)lmax, (int)mtype, (int)nmoment,
                             dout.get(), s);
    dout.download(output, (size_t)nout, s);
    NAPMO_HIP(hipStreamSynchronize(s));
    return 0;
  });
}

// ------------------------------------------------------------------- Gaussians
void *PrimitiveGaussian_new(int *l, double *origin, double zeta, double coeff) {
  return legacy(__func__, [&] { return (void *)make_primitive(l, origin, zeta, coeff); });
}
double PrimitiveGaussian_get_norma(void *p) { return ((Primitive *)p)->norma; }
double PrimitiveGaussian_get_zeta(void *p) { return ((Primitive *)p)->zeta; }
double PrimitiveGaussian_get_coeff(void *p) { return ((Primitive *)p)->coeff; }
void PrimitiveGaussian_get_l(void *p, int *l) {
  for (int i = 0; i < 3; ++i) l[i] = ((Primitive *)p)->l[i];
}
void PrimitiveGaussian_get_origin(void *p, double *o) {
  for (int i = 0; i < 3; ++i) o[i] = ((Primitive *)p)->origin[i];
}
double PrimitiveGaussian_compute(void *p, double *r, double *out, int size) {
  // values at size points (src/system/gto.cpp:97-104); the reference declares a double
  // result (napmo/__init__.py:206) but returns nothing: 0 here
  const Primitive *g = (const Primitive *)p;
  for (int i = 0; i < size; ++i) out[i] = g->value(&r[3 * (size_t)i]);
  return 0.0;
}
double PrimitiveGaussian_overlap(void *p, void *q) {
  return primitive_overlap(*(Primitive *)p, *(Primitive *)q);
}
void *ContractedGaussian_new(void **prims, int n) {
  return legacy(__func__, [&] { return (void *)make_contraction((Primitive **)prims, n); });
}
double ContractedGaussian_get_norma(void *c) { return ((Contraction *)c)->norma; }
int ContractedGaussian_get_nprim(void *c) { return (int)((Contraction *)c)->prim.size(); }
void ContractedGaussian_get_l(void *c, int *l) {
  for (int i = 0; i < 3; ++i) l[i] = ((Contraction *)c)->l[i];
}
void ContractedGaussian_get_origin(void *c, double *o) {
  for (int i = 0; i < 3; ++i) o[i] = ((Contraction *)c)->origin[i];
}
void ContractedGaussian_compute(void *c, double *r, double *out, int size) {
  const Contraction *g = (const Contraction *)c;
  for (int i = 0; i < size; ++i) out[i] = g->value(&r[3 * (size_t)i]);
}
double ContractedGaussian_overlap(void *c, void *d) {
  // norma_c norma_d sum over primitive pairs (src/system/gto.cpp:196-206)
  const Contraction *a = (const Contraction *)c, *b = (const Contraction *)d;
  double s = 0.0;
  for (const auto &pa : a->prim)
    for (const auto &pb : b->prim) s += primitive_overlap(pa, pb);
  return s * a->norma * b->norma;
}
void *BasisSet_new(void **conts, int n) {
  auto *b = new BasisSet;
  for (int i = 0; i < n; ++i) b->cont.push_back(*(Contraction *)conts[i]);
  return b;
}
void *BasisSet_new_empty(void) { return new BasisSet; }
void BasisSet_update(void *basis, void *other) {
  auto *b = (BasisSet *)basis;
  for (auto &c : ((BasisSet *)other)->cont) b->cont.push_back(c);
}
int BasisSet_get_nbasis(void *basis) { return (int)((BasisSet *)basis)->cont.size(); }
int BasisSet_get_max_l(void *basis) { return ((BasisSet *)basis)->max_l(); }
int BasisSet_get_max_nprim(void *basis) { return ((BasisSet *)basis)->max_nprim(); }

void BasisSet_deriv(void *basis, double *r, double *out, int npoint) {
  // out [npoint x nbasis x 3]: the Cartesian gradient of every function, on the GPU
  legacy(__func__, [&] {
    hipStream_t st = scratch_stream();
    FlatBasis fb = flatten(*(BasisSet *)basis);
    if (npoint <= 0 || fb.nbasis == 0) return 0;
    DevicePoints pts;
    pts.upload(r, npoint, st);
    DeviceBasis db;
    db.upload(fb, st);
    DeviceBuffer<double> dout((size_t)npoint * fb.nbasis * 3);
    launch_basis_gradient(pts.x.get(), pts.y.get(), pts.z.get(), npoint, fb.nbasis,
                          db.origin.get(), db.lxyz.get(), db.first.get(), db.nprim.get(),
                          db.norm.get(), db.zeta.get(), db.coef.get(), dout.get(), st);
    dout.download(out, (size_t)npoint * fb.nbasis * 3, st);
    NAPMO_HIP(hipStreamSynchronize(st));
    return 0;
  });
}

void BasisSet_compute(void *basis, double *r, double *out, int npoint) {
  legacy(__func__, [&] {
    hipStream_t st = scratch_stream();
    FlatBasis fb = flatten(*(BasisSet *)basis);
    DevicePoints pts;
    pts.upload(r, npoint, st);
    DeviceBasis db;
    db.upload(fb, st);
    DeviceBuffer<double> dout((size_t)npoint * fb.nbasis);
    db.collocate(pts.x.get(), pts.y.get(), pts.z.get(), npoint, npoint, fb.nbasis, dout.get(),
                 nullptr, st);
    dout.download(out, (size_t)npoint * fb.nbasis, st);
    NAPMO_HIP(hipStreamSynchronize(st));
    return 0;
  });
}

// --------------------------------------------------------- grid contractions
void nwavefunction_compute_2body_matrix(WaveFunction *psi, void *grid, double *phi, double *J,
                                        double *K) {
  legacy(__func__, [&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    const int n = psi->ndim;
    std::vector<double> G((size_t)n * n), E((size_t)n * n);
    host_contract(g, phi, nullptr, J, n, false, 1.0, G.data());
    if (psi->x_factor != 0.0) {
      // exchange(i, j) = sum_p phi_i K_j bw w for i <= j, mirrored (nwavefunction.cpp:50-62)
      host_contract(g, phi, K, nullptr, n, true, psi->x_factor, E.data());
      for (size_t e = 0; e < G.size(); ++e) G[e] += E[e];
    }
    std::memcpy(psi->G, G.data(), sizeof(double) * n * n);
    return 0;
  });
}

void nwavefunction_compute_coupling(WaveFunction *psi, void *grid, double *phi, double *other_J,
                                    double *res) {
  legacy(__func__, [&] {
    host_contract(((BeckeHandle *)grid)->grid, phi, nullptr, other_J, psi->ndim, false, 1.0, res);
    return 0;
  });
}

void nwavefunction_compute_xc_matrix(WaveFunction *psi, void *grid, double *phi, double *vrho) {
  legacy(__func__, [&] {
    host_contract(((BeckeHandle *)grid)->grid, phi, nullptr, vrho, psi->ndim, false, 1.0, psi->XC);
    return 0;
  });
}

void nwavefunction_compute_density_from_dm(void *basis, void *grid, double *dm, double *output,
                                           double epsilon, double *dmmaxrow) {
  legacy(__func__, [&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    hipStream_t s = g.stream;
    const FlatBasis fb = flatten(*(BasisSet *)basis);
    const int n = fb.nbasis;
    const long N = g.N;
    if (N == 0 || n == 0) return 0;
    NAPMO_REQUIRE(epsilon <= 0 || dmmaxrow, "density_from_dm: dmmaxrow is required when epsilon > 0");
    DeviceBasis db;
    db.upload(fb, s);
    DeviceBuffer<double> phi((size_t)N * n), ddm, dmax, absmax, upper, thr, out(N);
    DeviceBuffer<long> stop(1);
    db.collocate(g.x.get(), g.y.get(), g.z.get(), N, N, n, phi.get(), nullptr, s);
    ddm.upload(dm, (size_t)n * n, s);
    if (epsilon > 0) {
      dmax.upload(dmmaxrow, (size_t)n, s);
      absmax.resize(N), upper.resize(N), thr.resize(N);
    }
    launch_dm_density(phi.get(), n, N, n, ddm.get(), dmax.get(), epsilon, absmax.get(),
                      upper.get(), thr.get(), stop.get(), out.get(), s);
    long nwritten = N;
    if (epsilon > 0) stop.download(&nwritten, 1, s);
    NAPMO_HIP(hipStreamSynchronize(s));
    out.download(output, (size_t)nwritten, s);
    NAPMO_HIP(hipStreamSynchronize(s));
    return 0;
  });
}

// ----------------------------------------------------------- device-resident path
int napmo_compute_coulomb(void *grid, const double *dens, int lmax, double *v_out) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    DeviceBuffer<double> d, v(g.N);
    d.upload(dens, g.N, g.stream);
    g.poisson_expand(d.get(), lmax);
    g.poisson_synthesize(lmax, v.get());
    v.download(v_out, g.N, g.stream);
    NAPMO_HIP(hipStreamSynchronize(g.stream));
  });
}

int napmo_compute_kinetic(void *grid, const double *psi, int lmax, double *out) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    DeviceBuffer<double> d, o(g.N);
    d.upload(psi, g.N, g.stream);
    g.kinetic(d.get(), lmax, o.get());
    o.download(out, g.N, g.stream);
    NAPMO_HIP(hipStreamSynchronize(g.stream));
  });
}

int napmo_kinetic_expand(void *grid, const double *psi, int lmax, double *U_y, double *U_dt,
                         double *rho_lm) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    DeviceBuffer<double> d;
    d.upload(psi, g.N, g.stream);
    g.kinetic_expand(d.get(), lmax);
    PoissonPlan &pl = g.kinetic_plan(lmax);
    pl.work.Uy.download(U_y, pl.work.Uy.size(), g.stream);
    pl.work.Udt.download(U_dt, pl.work.Udt.size(), g.stream);
    if (rho_lm) pl.work.rho_lm.download(rho_lm, pl.work.rho_lm.size(), g.stream);
    NAPMO_HIP(hipStreamSynchronize(g.stream));
  });
}

int napmo_compute_exchange_potential(void *grid, int ndim, const double *psi, const double *D,
                                     int lmax, double *K_out) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    NAPMO_REQUIRE(ndim >= 0, "napmo_compute_exchange_potential: negative ndim");
    const long N = g.N;
    hipStream_t st = g.stream;
    DeviceBuffer<double> dpsi, dD, K((size_t)ndim * N);
    dpsi.upload(psi, (size_t)ndim * N, st);
    dD.upload(D, (size_t)ndim * ndim, st);
    g.exchange_potential(ndim, dpsi.get(), dD.get(), lmax, K.get());
    K.download(K_out, (size_t)ndim * N, st);
    NAPMO_HIP(hipStreamSynchronize(st));
  });
}

int napmo_compute_exchange_potential_device(void *grid, int ndim, const double *psi,
                                            const double *D, int lmax, double *K_out) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    NAPMO_REQUIRE(ndim >= 0, "napmo_compute_exchange_potential_device: negative ndim");
    g.exchange_potential(ndim, psi, D, lmax, K_out);
    NAPMO_HIP(hipStreamSynchronize(g.stream));
  });
}

int napmo_compute_kinetic_device(void *grid, const double *psi, int lmax, double *out) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    g.kinetic(psi, lmax, out);
    NAPMO_HIP(hipStreamSynchronize(g.stream));
  });
}

int napmo_compute_dpsi(void *grid, int lmax, int norb, const double *phi, const double *doi,
                       const double *oi, const double *ri, const double *V_tot, double mass_inv,
                       double *dpsi_out) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    NAPMO_REQUIRE(norb >= 0, "napmo_compute_dpsi: negative orbital count");
    if (norb == 0) return;
    const size_t n = (size_t)norb * g.N;
    DeviceBuffer<double> dphi, dri, dV, ddoi, doi_, out(n);
    dphi.upload(phi, n, g.stream), dri.upload(ri, n, g.stream), dV.upload(V_tot, n, g.stream);
    ddoi.upload(doi, norb, g.stream), doi_.upload(oi, norb, g.stream);
    g.orbital_correction(lmax, norb, dphi.get(), ddoi.get(), doi_.get(), dri.get(), dV.get(),
                         mass_inv, out.get());
    out.download(dpsi_out, n, g.stream);
    NAPMO_HIP(hipStreamSynchronize(g.stream));
  });
}

int napmo_becke_dickson_step(void *grid, int lmax, int norb, const double *psi, const double *T,
                             const double *V_tot, const double *e_prev, double mass_inv,
                             double *dpsi_out, double *de_out, double *R_out) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    NAPMO_REQUIRE(norb >= 0, "napmo_becke_dickson_step: negative orbital count");
    if (norb == 0) return;
    const size_t n = (size_t)norb * g.N;
    DeviceBuffer<double> dpsi, dT, dV, de, dout(n), dde(norb), dR(n);
    dpsi.upload(psi, n, g.stream), dT.upload(T, n, g.stream), dV.upload(V_tot, n, g.stream);
    de.upload(e_prev, norb, g.stream);
    g.becke_dickson_step(lmax, norb, dpsi.get(), dT.get(), dV.get(), de.get(), mass_inv,
                         dout.get(), dde.get(), dR.get());
    dout.download(dpsi_out, n, g.stream);
    dde.download(de_out, norb, g.stream);
    if (R_out) dR.download(R_out, n, g.stream);
    NAPMO_HIP(hipStreamSynchronize(g.stream));
  });
}

int napmo_becke_dickson_step_device(void *grid, int lmax, int norb, const double *psi,
                                    const double *T, const double *V_tot, const double *e_prev,
                                    double mass_inv, double *dpsi_out, double *de_out,
                                    double *R_out) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    NAPMO_REQUIRE(R_out != nullptr, "napmo_becke_dickson_step_device: R_out is required");
    g.becke_dickson_step(lmax, norb, psi, T, V_tot, e_prev, mass_inv, dpsi_out, de_out, R_out);
  });
}

int napmo_epc17_2(long n, const double *rhoE, const double *rhoN, double *c_ene, double *c_pot_a,
                  double *c_pot_b) {
  return guarded([&] {
    NAPMO_REQUIRE(n >= 0, "napmo_epc17_2: negative length");
    if (n == 0) return;
    hipStream_t st = scratch_stream();
    DeviceBuffer<double> e, p, ene(n), pa(n), pb(n);
    DeviceBuffer<int> flag(1);
    e.upload(rhoE, n, st), p.upload(rhoN, n, st);
    NAPMO_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), st));
    launch_epc17_2(e.get(), p.get(), n, flag.get(), ene.get(), pa.get(), pb.get(), st);
    ene.download(c_ene, n, st), pa.download(c_pot_a, n, st), pb.download(c_pot_b, n, st);
    NAPMO_HIP(hipStreamSynchronize(st));
  });
}

int napmo_epc17_2_grid(void *grid_a, void *grid_b, const int *index, int npair,
                       const double *rho_a, const double *rho_b, double *vrho_a, double *vrho_b,
                       double *energy) {
  return guarded([&] {
    DeviceGrid &ga = ((BeckeHandle *)grid_a)->grid, &gb = ((BeckeHandle *)grid_b)->grid;
    NAPMO_REQUIRE(npair >= 0, "napmo_epc17_2_grid: negative pair count");
    for (long k = 0; k < npair; ++k)
      NAPMO_REQUIRE(index[2 * k] >= 0 && index[2 * k] < ga.N && index[2 * k + 1] >= 0 &&
                        index[2 * k + 1] < gb.N,
                    "napmo_epc17_2_grid: common-point index out of range of the grids");
    hipStream_t st = gb.stream;
    if (ga.stream != gb.stream) NAPMO_HIP(hipStreamSynchronize(ga.stream));
    DeviceBuffer<double> ra, rb, va, vb, part(EPC_PART), e(1);
    DeviceBuffer<int> idx, flag(1);
    ra.upload(rho_a, ga.N, st), rb.upload(rho_b, gb.N, st);
    va.upload(vrho_a, ga.N, st), vb.upload(vrho_b, gb.N, st);
    idx.upload(index, 2 * (size_t)npair, st);
    NAPMO_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), st));
    launch_epc17_2_pairs(ra.get(), rb.get(), idx.get(), npair, ga.N, gb.N, gb.wbw.get(), flag.get(),
                         va.get(), vb.get(), part.get(), e.get(), st);
    va.download(vrho_a, ga.N, st), vb.download(vrho_b, gb.N, st);
    e.download(energy, 1, st);
    NAPMO_HIP(hipStreamSynchronize(st));
  });
}

int napmo_epc17_2_grid_noalloc(void *grid_a, void *grid_b, const int *d_index, int npair,
                               const double *d_rho_a, const double *d_rho_b, double *d_vrho_a,
                               double *d_vrho_b, double *d_energy) {
  return guarded([&] {
    NAPMO_REQUIRE(grid_a && grid_b && npair >= 0, "napmo_epc17_2_grid_noalloc: bad arguments");
    DeviceGrid &ga = ((BeckeHandle *)grid_a)->grid, &gb = ((BeckeHandle *)grid_b)->grid;
    launch_epc17_2_pairs_noalloc(d_rho_a, d_rho_b, d_index, npair, ga.N, gb.N, gb.wbw.get(),
                                 d_vrho_a, d_vrho_b, d_energy, gb.stream);
  });
}

int napmo_grid_sync(void *grid) {
  return guarded([&] { NAPMO_HIP(hipStreamSynchronize(((BeckeHandle *)grid)->grid.stream)); });
}

int napmo_grid_profile(void *grid, int enable) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    if (enable)
      g.profile_start();
    else
      g.prof_on = false;
  });
}

int napmo_grid_profile_read(void *grid, double *avg_ms, int nphase, int *ncalls) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    const int n = g.profile_read(avg_ms, nphase);
    if (ncalls) *ncalls = n;
  });
}

int napmo_compute_nuclear(void *grid, int ncharge, const double *charges, const double *origins,
                          double *V_out) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    NAPMO_REQUIRE(ncharge >= 0, "napmo_compute_nuclear: negative charge count");
    DeviceBuffer<double> q, o, V(g.N);
    q.upload(charges, ncharge, g.stream);
    o.upload(origins, 3 * (size_t)ncharge, g.stream);
    launch_nuclear(g.x.get(), g.y.get(), g.z.get(), g.N, q.get(), o.get(), ncharge, V.get(), g.stream);
    V.download(V_out, g.N, g.stream);
    NAPMO_HIP(hipStreamSynchronize(g.stream));
  });
}

int napmo_poisson_solve(void *grid, const double *dens, int lmax, double *U_y, double *U_dt) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    DeviceBuffer<double> d;
    d.upload(dens, g.N, g.stream);
    g.poisson_expand(d.get(), lmax);
    PoissonPlan &pl = g.plan(lmax);
    pl.work.Uy.download(U_y, pl.work.Uy.size(), g.stream);
    pl.work.Udt.download(U_dt, pl.work.Udt.size(), g.stream);
    NAPMO_HIP(hipStreamSynchronize(g.stream));
  });
}

int napmo_poisson_solve_expansion(void *grid, const double *rho_lm, int lmax, double *U_y,
                                  double *U_dt) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    PoissonPlan &pl = g.plan(lmax);
    NAPMO_REQUIRE(rho_lm != nullptr, "napmo_poisson_solve_expansion: null expansion");
    pl.work.rho_lm.upload(rho_lm, pl.work.rho_lm.size(), g.stream);
    AtomTables at = g.tables(pl);
    launch_radial(at, pl.work.rho_lm.get(), pl.green.get(), lmax, pl.work.Uy.get(), pl.work.Udt.get(),
                  g.stream);
    pl.work.Uy.download(U_y, pl.work.Uy.size(), g.stream);
    pl.work.Udt.download(U_dt, pl.work.Udt.size(), g.stream);
    NAPMO_HIP(hipStreamSynchronize(g.stream));
  });
}

int napmo_fock_plan_new(void *grid, void *basis, int lmax, int prec_mode, void **plan_out) {
  return guarded([&] {
    NAPMO_REQUIRE(grid && basis && plan_out, "napmo_fock_plan_new: null argument");
    auto *p = new FockPlan(((BeckeHandle *)grid)->grid, *(BasisSet *)basis, lmax, prec_mode);
    *plan_out = p;
  });
}
void napmo_fock_plan_del(void *plan) { delete (FockPlan *)plan; }
int napmo_fock_plan_info(void *plan, long *npoint, int *nbasis, int *ncenter, int *lmax) {
  return guarded([&] {
    FockPlan *p = (FockPlan *)plan;
    if (npoint) *npoint = p->grid.N;
    if (nbasis) *nbasis = p->n;
    if (ncenter) *ncenter = p->grid.M;
    if (lmax) *lmax = p->lmax;
  });
}
int napmo_grid_fock_build(void *plan, const double *P, double *F_out, double *E_out,
                          double *rho_out, double *v_out) {
  return guarded([&] { ((FockPlan *)plan)->build_host(P, F_out, E_out, rho_out, v_out); });
}
int napmo_grid_fock_build_device(void *plan, const double *dP, double *dF, double *dE) {
  return guarded([&] { ((FockPlan *)plan)->build_device(dP, dF, dE); });
}
int napmo_grid_fock_build_orbitals(void *plan, const double *C, int nocc, double eta,
                                   double *F_out, double *E_out, double *rho_out, double *v_out) {
  return guarded(
      [&] { ((FockPlan *)plan)->build_orbitals_host(C, nocc, eta, F_out, E_out, rho_out, v_out); });
}
int napmo_grid_fock_build_orbitals_device(void *plan, const double *dC, int nocc, double eta,
                                          double *dF, double *dE) {
  return guarded([&] { ((FockPlan *)plan)->build_orbitals_device(dC, nocc, eta, dF, dE); });
}
int napmo_device_copy(double *dst, const double *src, long n) {
  return guarded([&] {
    launch_copy(src, dst, n, nullptr);
    NAPMO_HIP(hipDeviceSynchronize());
  });
}
int napmo_selftest_mfma(int which, double *out) {
  return guarded([&] {
    require_device();
    DeviceBuffer<double> d(256);
    launch_selftest_mfma(which, d.get(), nullptr);
    NAPMO_HIP(hipMemcpy(out, d.get(), 256 * sizeof(double), hipMemcpyDeviceToHost));
  });
}
int napmo_grid_overlap(void *plan, double *S_out) {
  return guarded([&] { ((FockPlan *)plan)->overlap(S_out); });
}

int napmo_fock_plan_contract(void *plan, const double *v, double *F_out) {
  return guarded([&] { ((FockPlan *)plan)->contract_potential(v, F_out); });
}
int napmo_fock_plan_profile(void *plan, int max_builds) {
  return guarded([&] { ((FockPlan *)plan)->profile_start(max_builds); });
}
int napmo_fock_plan_profile_read(void *plan, double *avg_ms, int nphase, int *nbuild) {
  return guarded([&] {
    int n = ((FockPlan *)plan)->profile_read(avg_ms, nphase);
    if (nbuild) *nbuild = n;
  });
}
const char *napmo_fock_phase_name(int k) { return FockPlan::phase_name(k); }

int napmo_fock_plan_centres(void *plan, int *ncentre, double *xyz) {
  return guarded([&] {
    const FockPlan *p = (const FockPlan *)plan;
    const int nc = (int)p->centres.size() / 3;
    if (ncentre) *ncentre = nc;
    if (xyz) std::memcpy(xyz, p->centres.data(), sizeof(double) * 3 * nc);
  });
}
int napmo_grid_fock_gradient(void *plan, const double *P, double *grad_out) {
  return guarded([&] {
    FockPlan *p = (FockPlan *)plan;
    hipStream_t st = p->stream;
    const int nc = (int)p->centres.size() / 3;
    DeviceBuffer<double> dP, dG(3 * nc);
    dP.upload(P, (size_t)p->n * p->n, st);
    p->gradient_device(dP.get(), dG.get());
    dG.download(grad_out, 3 * nc, st);
    NAPMO_HIP(hipStreamSynchronize(st));
  });
}
int napmo_grid_fock_gradient_device(void *plan, const double *dP, double *dG) {
  return guarded([&] { ((FockPlan *)plan)->gradient_device(dP, dG); });
}
int napmo_grid_fock_energy_dP(void *plan, double *dEdP_out) {
  return guarded([&] {
    FockPlan *p = (FockPlan *)plan;
    DeviceBuffer<double> d((size_t)p->n * p->n);
    p->energy_dP(d.get());
    d.download(dEdP_out, (size_t)p->n * p->n, p->stream);
    NAPMO_HIP(hipStreamSynchronize(p->stream));
  });
}
int napmo_grid_fock_energy_dP_device(void *plan, double *d_dEdP) {
  return guarded([&] { ((FockPlan *)plan)->energy_dP(d_dEdP); });
}
int napmo_grid_poisson_adjoint(void *plan, const double *y, double *z_out) {
  return guarded([&] {
    FockPlan *p = (FockPlan *)plan;
    hipStream_t st = p->stream;
    DeviceBuffer<double> dy, dz(std::max<long>(p->grid.N, 1));
    dy.upload(y, p->grid.N, st);
    p->adjoint_potential(dy.get(), dz.get());
    dz.download(z_out, p->grid.N, st);
    NAPMO_HIP(hipStreamSynchronize(st));
  });
}

int napmo_fock_plan_set_tuning(void *plan, int key, long value) {
  return guarded([&] { ((FockPlan *)plan)->set_tuning(key, value); });
}
int napmo_fock_plan_get_tuning(void *plan, int key, long *value) {
  return guarded([&] { *value = ((FockPlan *)plan)->get_tuning(key); });
}
int napmo_fock_plan_int8_screen_stats(void *plan, long *kept, long *total) {
  return guarded([&] {
    NAPMO_REQUIRE(kept && total, "int8 screening statistics: null output");
    const auto st = ((FockPlan *)plan)->int8_screen_stats();
    *kept = st.first, *total = st.second;
  });
}

int napmo_grid_set_tuning(void *grid, int key, long value) {
  return guarded([&] {
    DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    if (key == NAPMO_TUNE_SYNTH_NP) {
      NAPMO_REQUIRE(value == 0 || value == 1 || value == 2, "synthesis pairs per lane must be 1 or 2");
      g.synth_np = value ? (int)value : 2;
    } else if (key == NAPMO_TUNE_SYNTH_MFMA_MIN) {
      NAPMO_REQUIRE(value >= 0 && value <= 17, "MFMA synthesis threshold must be 0 to 17");
      g.synth_mfma_min = value ? (int)value : SYNTH_MFMA_MIN;
    } else {
      throw Error("unknown grid tuning key " + std::to_string(key));
    }
  });
}
int napmo_grid_get_tuning(void *grid, int key, long *value) {
  return guarded([&] {
    const DeviceGrid &g = ((BeckeHandle *)grid)->grid;
    if (key == NAPMO_TUNE_SYNTH_NP)
      *value = g.synth_np;
    else if (key == NAPMO_TUNE_SYNTH_MFMA_MIN)
      *value = g.synth_mfma_min;
    else
      throw Error("unknown grid tuning key " + std::to_string(key));
  });
}

int napmo_batch_plan_new(int nplan, void **plans, int nlanes, void **batch_out) {
  return guarded([&] {
    NAPMO_REQUIRE(nplan >= 1 && plans && batch_out, "napmo_batch_plan_new: bad arguments");
    std::vector<FockPlan *> v(nplan);
    for (int m = 0; m < nplan; ++m) v[m] = (FockPlan *)plans[m];
    *batch_out = new BatchPlan(v, nlanes);
  });
}
void napmo_batch_plan_del(void *batch) { delete (BatchPlan *)batch; }
int napmo_batch_build(void *batch, const double **dP, double **dF, double **dE) {
  return guarded([&] { ((BatchPlan *)batch)->build(dP, dF, dE); });
}
int napmo_batch_capture(void *batch, const double **dP, double **dF, double **dE) {
  return guarded([&] { ((BatchPlan *)batch)->capture(dP, dF, dE); });
}
int napmo_batch_build_orbitals(void *batch, const double **dC, const int *nocc, double eta,
                               double **dF, double **dE) {
  return guarded([&] { ((BatchPlan *)batch)->build_orbitals(dC, nocc, eta, dF, dE); });
}
int napmo_batch_capture_orbitals(void *batch, const double **dC, const int *nocc, double eta,
                                 double **dF, double **dE) {
  return guarded([&] { ((BatchPlan *)batch)->capture_orbitals(dC, nocc, eta, dF, dE); });
}
int napmo_batch_replay(void *batch) {
  return guarded([&] { ((BatchPlan *)batch)->replay(); });
}
int napmo_batch_sync(void *batch) {
  return guarded([&] { ((BatchPlan *)batch)->synchronize(); });
}

int napmo_apmo_plan_new(int nspecies, void **grids, void **bases, const double *charges,
                        const int *lmax, int prec_mode, void **plan_out) {
  return guarded([&] {
    require_device();
    NAPMO_REQUIRE(nspecies >= 1 && grids && bases && charges && lmax && plan_out,
                  "napmo_apmo_plan_new: bad arguments");
    std::vector<DeviceGrid *> g(nspecies);
    std::vector<const BasisSet *> b(nspecies);
    for (int s = 0; s < nspecies; ++s) {
      g[s] = &((BeckeHandle *)grids[s])->grid;
      b[s] = (const BasisSet *)bases[s];
    }
    *plan_out = new ApmoPlan(g, b, std::vector<double>(charges, charges + nspecies),
                             std::vector<int>(lmax, lmax + nspecies), prec_mode);
  });
}
void napmo_apmo_plan_del(void *plan) { delete (ApmoPlan *)plan; }
int napmo_apmo_build(void *plan, const double **P, double **F_out, double *E_out) {
  return guarded([&] { ((ApmoPlan *)plan)->build_host(P, F_out, E_out); });
}
int napmo_apmo_build_device(void *plan, const double **dP, double **dF, double *dE) {
  return guarded([&] { ((ApmoPlan *)plan)->build_device(dP, dF, dE); });
}
int napmo_apmo_sync(void *plan) {
  return guarded([&] { ((ApmoPlan *)plan)->synchronize(); });
}
int napmo_apmo_plan_set_one_launch(void *plan, int on) {
  return guarded([&] { ((ApmoPlan *)plan)->set_one_launch(on != 0); });
}
int napmo_apmo_plan_set_tuning(void *plan, int species, int key, long value) {
  return guarded([&] {
    ApmoPlan *a = (ApmoPlan *)plan;
    NAPMO_REQUIRE(species >= 0 && species < a->nspecies(), "apmo tuning: species index out of range");
    NAPMO_REQUIRE(!a->group, "apmo tuning: turn the one-launch mode off first");
    a->synchronize();
    a->sp[species].plan->set_tuning(key, value);
  });
}
int napmo_apmo_capture(void *plan, const double **dP, double **dF, double *dE) {
  return guarded([&] { ((ApmoPlan *)plan)->capture(dP, dF, dE); });
}
int napmo_apmo_replay(void *plan) {
  return guarded([&] { ((ApmoPlan *)plan)->replay(); });
}
void *napmo_device_malloc(size_t bytes) {
  void *p = nullptr;
  if (guarded([&] {
        require_device();
        NAPMO_HIP(hipMalloc(&p, bytes));
      }) != 0)
    return nullptr;
  return p;
}
void napmo_device_free(void *p) { (void)hipFree(p); }
int napmo_memcpy(void *dst, const void *src, size_t bytes, int kind) {
  return guarded([&] {
    NAPMO_REQUIRE(kind >= 1 && kind <= 3, "napmo_memcpy: kind 1 = H2D, 2 = D2H, 3 = D2D");
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                                      : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    NAPMO_HIP(hipMemcpy(dst, src, bytes, k));
  });
}

void *napmo_fock_plan_stream(void *plan) { return (void *)((FockPlan *)plan)->stream; }
int napmo_fock_plan_sync(void *plan) {
  return guarded([&] { NAPMO_HIP(hipStreamSynchronize(((FockPlan *)plan)->stream)); });
}

}  // extern "C"
