// Device mirror of a Becke grid and its Poisson plans.
#pragma once

#include <map>
#include <memory>
#include <vector>

#include "common.hpp"
#include "host.hpp"
#include "kernels.hpp"

namespace napmo {

// Hermite finite-element system of the radial Poisson equation
// V'' + (2/r) V' - l(l+1)/r^2 V = -4 pi rho on the t-axis (napmo/grids/poisson_solver.py
// :59-116, napmo/utilities/ode2.py:35-118, src/horton/ode2.cpp:232-321). The matrix depends
// only on (radial transform, l); rho enters linearly through the right-hand side and the
// two boundary values, so G = A^{-1} B maps rho_lm(r_j) to the potential spline (y, dt) at
// the nodes. Returns Gy, Gd [n x n] row-major (output node, input node).
void radial_green(const RTransform &rtf, const std::vector<double> &rw, int l,
                  std::vector<double> &Gy, std::vector<double> &Gd);

// build_ode2 of src/horton/ode2.cpp:232-321 (same row/column ranges and boundary rows).
void fe_assemble(const double *b, const double *a, const double *f, const double *bcs,
                 double *coeffs, double *rhs, long npoint);

// eval_decomposition_grid on the GPU for one centre (src/horton/evaluate.cpp:129-223):
// output[npoint] += synthesis of the given splines at the given host points.
void decomposition_on_points(CubicSpline **sp, int nspline, int lmax, const double *center,
                             const double *points, long npoint, double *output);

// Kinetic radial matrices for one l: Gy = deriv2 - l(l+1)/r^2 (on unit vectors), Gd = the
// cubic-spline slopes of those columns; same layout as radial_green.
void kinetic_green(const RTransform &rtf, double radii, const std::vector<double> &r, int l,
                   std::vector<double> &Gy, std::vector<double> &Gd);

// Finite-element tables of the orbital-correction equation u'' + (2/r) u' + a u = f on one
// radial grid (napmo/wavefunction/ndpsi.py:52-81 through napmo/utilities/ode2.py:35-118 and
// src/horton/ode2.cpp:232-321), with u = 0 at both ends: c0 [2n x 7] the Galerkin band
// without the a term (slot = col - row + 3), oa [2n x 7 x 6] the integrals multiplying the
// merged a coefficients begin(row) + k, o2 [2n x 6] the right-hand-side integrals of the
// merged source f_k, k = 2 (row / 2 - 1) + kk (zero rows for the two value conditions),
// and the t-derivatives j1, j2 of r(t).
void helmholtz_tables(const RTransform &rtf, std::vector<double> &c0, std::vector<double> &oa,
                      std::vector<double> &o2, std::vector<double> &j1, std::vector<double> &j2);

// Per-Lebedev-rule projection table Yhat[j][lm] = aw_j S_lm(u_j) sqrt(4 pi (2l+1)) / sum aw.
std::vector<double> projection_table(const AngularGrid &ang, int lmax);

struct AtomInfo {
  double origin[3];
  double radii;
  int nrad, nang;
  RTransform rtf;               // copy
  std::vector<double> r, rw;    // radial nodes and weights
  std::vector<double> u, aw;    // Lebedev unit vectors and weights
};

// Per-build Poisson workspace: rho_lm and U (y, dt) [sum_A nrad_A x L], synthesis tables
// T [M x (max_nrad - 1) x L x 4] and extrapolation amplitudes [M x L].
struct PoissonWork {
  DeviceBuffer<double> rho_lm, Uy, Udt, T, ampL, ampR;
};

// Per-(grid, lmax) tables shared by every build on the grid, and a default workspace.
struct PoissonPlan {
  int lmax = -1;
  DeviceBuffer<double> yhat, green;
  DeviceBuffer<long> yhat_off, green_off;
  PoissonWork work;
};

// Densities (orbitals, orbital pairs) whose projections, radial solves and synthesis tables
// go into one launch each in the orbital update and the exchange build.
constexpr int SET_BATCH = 64;
// MFMA synthesis: at most this many bytes of per-centre tile sums [M][N][16] at once (at
// least one tile)
constexpr long SYNTH_MFMA_BUDGET = 4L << 30;
// A tile of 16 densities on MFMA costs about as much as ten one-density syntheses (2.7 ms
// against 0.28 ms each on 16 centres x 472,000 points at lmax 12), so a last tile takes the
// MFMA path from ten densities on.
constexpr int SYNTH_MFMA_MIN = 10;

// Per-geometry pair buckets of the bucketed synthesis (see launch_synth2).
struct SynthPairs {
  long nwaves = 0;
  DeviceBuffer<int> list, wave_bucket;
  DeviceBuffer<long> slot0;  // [M (max_nrad + 1) + 1] first slot of every bucket
  DeviceBuffer<double> geo;  // [nwaves][4][64]: u or d, unit vector (launch_pair_geometry)
  DeviceBuffer<double> contrib;  // [M x N]
};

// Orbital-correction tables of every atom (one copy per distinct radial grid).
struct HelmholtzPlan {
  DeviceBuffer<double> c0, oa, o2, j1, j2;
  DeviceBuffer<long> c0_off, oa_off, o2_off;
  HelmholtzTables tables() const;
};

// Workspace of the orbital corrections, kept between calls (resized when norb or lmax
// change): sources, projections mu, solutions (y, dt), shell averages, synthesis tables.
struct OrbitalWork {
  DeviceBuffer<double> src, mu, Uy, Udt, Vsph, resid;
  DeviceBuffer<int> status;
  PoissonWork pw;
  int lmax = -1;
};

struct DeviceGrid {
  hipStream_t stream = nullptr;
  long N = 0;
  int M = 0;
  std::vector<AtomInfo> atoms;
  std::vector<long> offsets;  // [M+1]
  // host mirrors (borrowed views returned by the getters)
  std::vector<double> points, weights, becke, origin, radii;
  // device arrays
  DeviceBuffer<double> x, y, z, w, bw, wbw;
  DeviceBuffer<int> owner;
  DeviceBuffer<long> d_point_off, d_rad_off;
  DeviceBuffer<int> d_nrad, d_nang, d_rkind;
  DeviceBuffer<double> d_rp0, d_rp1, d_first, d_last, d_center, d_r;
  int max_nrad = 0;
  int synth_np = 2;  // (point, centre) pairs per lane of the synthesis (1 or 2)
  std::map<int, std::unique_ptr<PoissonPlan>> plans, kplans;  // Poisson / kinetic, per lmax
  std::unique_ptr<SynthPairs> pairs;
  std::unique_ptr<HelmholtzPlan> hplan;
  std::unique_ptr<OrbitalWork> owork;
  // Per-phase HIP-event timing of the orbital corrections and exchange builds (phases:
  // residual, sources and projection, shell averages, radial solves, synthesis tables,
  // one-density synthesis, reduction, MFMA synthesis).
  static constexpr int DPSI_NPHASE = 8;
  bool prof_on = false;
  int prof_calls = 0;
  size_t prof_next = 0;
  std::vector<hipEvent_t> prof_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_marks;
  hipEvent_t prof_start_ev = nullptr;
  int prof_phase = -1;
  void tic(int phase);
  void toc();
  void profile_start();
  int profile_read(double *avg_ms, int nphase);

  ~DeviceGrid();
  void build(const std::vector<AtomInfo> &atoms);
  // build = build_geometry (points, weights, atom tables) + becke_weights (bw, wbw).
  void build_geometry(const std::vector<AtomInfo> &atoms);
  void becke_weights();
  // Per-atom harmonic cut of the Poisson plans (empty: none): projections of l > lcut[a]
  // are zero (the grid of several molecules, each keeping its own lmax).
  std::vector<int> lcut;
  // A grid given only as points and Becke weights (the Lowdin2 compatibility grid of
  // src/grid/becke.cpp:57-77): pw [N x 4] rows (x, y, z, bw), weights 1, ncenter centres
  // without atomic grids. Integration, collocation and contractions work on it; the
  // Poisson, kinetic and orbital-correction paths need atomic grids and refuse it.
  void build_points(const double *pw, long n, int ncenter);
  void require_atoms(const char *what) const;
  AtomTables tables(const PoissonPlan &plan) const;
  PoissonPlan &plan(int lmax);
  // Same projection tables; radial matrices of the kinetic operator: U_lm = the spline
  // (y, dt) of deriv2(f_lm) - l(l+1)/r^2 f_lm (napmo/wavefunction/nkinetic.py:15-73).
  PoissonPlan &kinetic_plan(int lmax);
  PoissonPlan &build_plan(int lmax, bool kinetic);
  // d_out [N] = -1/2 laplacian of d_psi [N] (grid order), on the grid's stream.
  void kinetic(const double *d_psi, int lmax, double *d_out);
  // Its first half: the (l, m) splines (y, dt) into kinetic_plan(lmax).work.Uy / Udt.
  void kinetic_expand(const double *d_psi, int lmax);
  // A private workspace sized for plan(lmax) (for plans that must not share one).
  void alloc_work(int lmax, PoissonWork &w) const;
  // Poisson of dens (device [N]) into the plan's U / tables; then synthesis to v (device).
  void poisson_expand(const double *d_dens, int lmax);
  void poisson_synthesize(int lmax, double *d_v);
  SynthPairs &synth_pairs();
  HelmholtzPlan &helmholtz();
  // Becke-Dickson orbital corrections (napmo/wavefunction/ndpsi.py:15-89) of norb orbitals,
  // all device arrays [norb x N] (doi, oi: device [norb]): out[o] = compute_dpsi(phi[o],
  // doi[o], oi[o], ri[o], V[o]). Synchronises; throws on a singular radial system.
  void orbital_correction(int lmax, int norb, const double *d_phi, const double *d_doi,
                          const double *d_oi, const double *d_ri, const double *d_V,
                          double mass_inv, double *d_out);
  // Grid exchange potential (napmo/wavefunction/psi_numeric.py:186-215), all device
  // arrays: K [ndim x N] = sum_k sum_l psi_k V[psi_l psi_j] D[k][l] by ndim(ndim+1)/2
  // Poisson solves of orbital products.
  void exchange_potential(int ndim, const double *d_psi, const double *d_D, int lmax, double *d_K);
  DeviceBuffer<double> xchi, xdens, xaux;  // its workspace, kept between calls
  DeviceBuffer<double> xrho, xUy, xUdt;    // per batch of orbital pairs
  // Synthesis of nset densities on this grid from their radial solutions (set s at
  // Uy, Udt + s ustride, plan(lmax) order): out[s ostride + p] = scale v_s[p]. Whole tiles
  // of 16 densities, and a last tile of at least synth_mfma_min, go through the FP64 MFMA
  // synthesis (launch_synth_mfma); the others one at a time through the bucketed synthesis.
  // Profiling phases 4 (tables), 5 (one-density synthesis), 7 (MFMA synthesis), 6 (centre
  // sums).
  void synth_sets(int lmax, int nset, const double *Uy, const double *Udt, long ustride,
                  double *out, long ostride, double scale = 1.0);
  int synth_mfma_min = SYNTH_MFMA_MIN;
  DeviceBuffer<double> mT, maL, maR, mcontrib;  // its workspace (interleaved tables, tiles)
  DeviceBuffer<double> sT, saL, saR;            // tables of the one-at-a-time densities
  // Residual R = T + V - e psi and de = integrate(R psi) of norb orbitals, then their
  // corrections (napmo/wavefunction/psi_optimization.py:175-203).
  void becke_dickson_step(int lmax, int norb, const double *d_psi, const double *d_T,
                          const double *d_V, const double *d_e, double mass_inv, double *d_dpsi,
                          double *d_de, double *d_R);
  // Atom arrays without a Poisson plan (enough for bucketing pairs).
  AtomTables geometry_tables() const;
  // Synthesis of this grid's expansion (plan(lmax), after poisson_expand / expand) at the
  // target points of sp: v (+)= scale * sum_A contrib[A], on stream st.
  void synthesize_at(int lmax, const PoissonWork &w, const SynthPairs &sp, const double *tx,
                     const double *ty, const double *tz, long tN, double scale, int accumulate,
                     double *d_v, hipStream_t st);
};

// What a BeckeGrid_* handle points to.
struct BeckeHandle {
  DeviceGrid grid;
  int ablmax = 0;        // auxiliary-basis settings kept from BeckeGrid_new
  double abldep = 1e-6;
};

// Orthonormal auxiliary basis on a Becke grid (AuxiliaryBasis_*): aobasis [N x nao]
// point-major.
struct AuxBasis {
  int ncenter = 0, nao = 0;
  std::vector<double> aobasis;
};
AuxBasis *make_aux_basis(const DeviceGrid &g, int ablmax, double abldep);

// (point, centre) pair buckets of src's centres and radial grids at the target points
// x, y, z [N] (device): the synthesis lists of src's expansion evaluated at those points.
std::unique_ptr<SynthPairs> build_synth_pairs(const DeviceGrid &src, const double *x,
                                              const double *y, const double *z, long N,
                                              hipStream_t stream);

}  // namespace napmo
