/* napmo_gpu.h -- C ABI of libnapmo_gpu.so, the MI355X (gfx950) grid hot path of nAPMO.
 *
 * Two layers:
 *
 *  1. Drop-in symbols. Same names and C signatures as the reference library that
 *     napmo/__init__.py:95-585 binds with ctypes, restricted to the grid hot path
 *     (grids, radial transforms, splines, ODE assembly, Poisson synthesis, GTO
 *     collocation, grid contractions). Each cites the reference declaration it
 *     replaces. Host arrays in and out, as before; the arithmetic runs on the GPU.
 *     Pointer-returning getters return borrowed views owned by the handle.
 *     Callee-allocated results (AtomicGrid_spherical_expansion, AtomicGrid_integrate)
 *     are released with napmo_free().
 *
 *  2. Device-resident entry points (napmo_*). Return int status (0 = ok) and never
 *     throw or exit; napmo_last_error() returns the thread's last message.
 *
 * Every array is C-contiguous float64 unless stated. No torch types cross this ABI.
 */
#ifndef NAPMO_GPU_H
#define NAPMO_GPU_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors, memory */

/* Thread-local message of the last failed call ("" when none). */
const char *napmo_last_error(void);
/* Releases a callee-allocated result (the reference returns new[] buffers that
 * Python never frees: src/grid/atomic_grid.cpp:68, 116). */
void napmo_free(void *ptr);
/* Number of visible GPUs (0 when none); never initialises a context on failure. */
int napmo_device_count(void);

/* ------------------------------------------------- radial transforms (rtransform.h) */
/* napmo/__init__.py RTransform block; src/horton/rtransform.cpp:308-341, :257-283 */
void *ChebyshevRTransform_new(double radii, int npoint);
double ChebyshevRTransform_get_radii(void *rtf);
double ChebyshevRTransform_get_z(void *rtf);
void *PowerRTransform_new(double rmin, double rmax, int npoint);
double PowerRTransform_get_rmin(void *rtf);
double PowerRTransform_get_rmax(void *rtf);
double PowerRTransform_get_power(void *rtf);
void *IdentityRTransform_new(int npoint);
int RTransform_get_npoint(void *rtf);
double RTransform_radius(void *rtf, double t);
double RTransform_deriv(void *rtf, double t);
double RTransform_deriv2(void *rtf, double t);
double RTransform_deriv3(void *rtf, double t);
double RTransform_inv(void *rtf, double r);
void RTransform_radius_array(void *rtf, double *t, double *r, int n);
void RTransform_deriv_array(void *rtf, double *t, double *d, int n);
void RTransform_deriv2_array(void *rtf, double *t, double *d, int n);
void RTransform_deriv3_array(void *rtf, double *t, double *d, int n);
void RTransform_inv_array(void *rtf, double *r, double *t, int n);
void RTransform_del(void *rtf);

/* ----------------------------------------------------------- grids (src/grid) */
/* src/grid/angular.cpp:12-37 (Lebedev rule with lorder points) */
void *AngularGrid_new(int lorder);
int AngularGrid_get_lorder(void *grid);
double *AngularGrid_get_points(void *grid);
double *AngularGrid_get_weights(void *grid);
double AngularGrid_integrate(void *grid, int nfunc, double *f);
void AngularGrid_del(void *grid);
/* src/grid/angular.cpp:39-48 (napmo/__init__.py:284-285): points -> (1, theta, phi) in place */
void AngularGrid_spherical(void *grid);
/* src/grid/angular.cpp:50-93 (napmo/__init__.py:287-289): output[k][lm] = 4 pi sum_j aw_j
 * f[k][j] Y_lm(u_j) for size_f rows of lorder values; real Y_lm of GSL's convention
 * (Condon-Shortley phase), lm ordered m = -l..l per l */
void AngularGrid_spherical_expansion(void *grid, int lmax, int size_f, double *f, double *output);
/* src/grid/angular.cpp:95-139 (napmo/__init__.py:291-293): output[k][j] = sum_lm
 * decomposition[k][lm] Y_lm(u_j) */
void AngularGrid_eval_expansion(void *grid, int lmax, int size_f, double *decomposition,
                                double *output);
/* src/grid/radial.cpp:14-51 */
void *RadialGrid_new(void *rtf, double radii);
int RadialGrid_get_size(void *grid);
double RadialGrid_get_radii(void *grid);
double *RadialGrid_get_points(void *grid);
double *RadialGrid_get_weights(void *grid);
double RadialGrid_integrate(void *grid, int segments, double *f);
/* src/grid/radial.cpp:53-192: (2/r) f' + f'' by finite differences (Chebyshev grids,
 * size >= 8); returns [size], release with napmo_free. */
double *RadialGrid_deriv2(void *grid, double *f);
void RadialGrid_del(void *grid);
/* src/grid/atomic_grid.cpp:12-136 */
void *AtomicGrid_new(void *angular, void *radial, double *origin);
int AtomicGrid_get_size(void *grid);
double AtomicGrid_get_radii(void *grid);
double *AtomicGrid_get_origin(void *grid);
double *AtomicGrid_get_points(void *grid);
double *AtomicGrid_get_weights(void *grid);
/* Returns [nrad x (lmax+1)^2], release with napmo_free. GPU projection. */
double *AtomicGrid_spherical_expansion(void *grid, double *f, int lmax);
/* Returns [segments], release with napmo_free. */
double *AtomicGrid_integrate(void *grid, int nfunc, int segments, int size, double *f);
void AtomicGrid_del(void *grid);
/* src/grid/becke.cpp:12-54, 79-190. Becke weights computed on the GPU. Atomic grids
 * may have different sizes (the reference assumes equal sizes, becke.cpp:121). */
void *BeckeGrid_new(void **atgrids, int ncenter, int ablmax, double abldep);
int BeckeGrid_get_ncenter(void *grid);
int BeckeGrid_get_size(void *grid);
double *BeckeGrid_get_radii(void *grid);
double *BeckeGrid_get_points(void *grid);
double *BeckeGrid_get_weights(void *grid);
double *BeckeGrid_get_origin(void *grid);
double *BeckeGrid_get_becke_weights(void *grid);
double BeckeGrid_integrate(void *grid, double *f);
void BeckeGrid_del(void *grid);
/* Lowdin2 compatibility grid (src/grid/becke.cpp:57-77, bound at napmo/__init__.py:368-369,
 * read by napmo/grids/becke.py:42-50): pw [size x 4] rows (x, y, z, Becke weight), weights
 * 1, ncenter centres without atomic grids (get_origin / get_radii return NULL). Poisson,
 * kinetic and orbital-correction entry points report an error on such a grid. */
void *BeckeGrid_from_points(double *pw, int size, int ncenter);

/* ------------------------------------------- auxiliary basis (auxiliary_basis.cpp) */
/* src/grid/auxiliary_basis.cpp:10-280 (napmo/__init__.py:422-432): orthonormal auxiliary
 * functions of the grid (ablmax, abldep from BeckeGrid_new); computed on creation. */
void *AuxiliaryBasis_new(void *grid);
void AuxiliaryBasis_del(void *auxiliary_basis);
/* no-op: the basis is computed by AuxiliaryBasis_new (auxiliary_basis.h:59) */
void AuxiliaryBasis_compute_aobasis(void *auxiliary_basis);
/* borrowed [npoint x nao] point-major */
double *AuxiliaryBasis_get_aobasis(void *auxiliary_basis);
int AuxiliaryBasis_get_nao(void *auxiliary_basis);
int AuxiliaryBasis_get_ncenter(void *auxiliary_basis);

/* ------------------------------------------------------------ MultiGrid (f4 glue) */
/* src/grid/multi_grid.cpp:14-132 (bound at napmo/__init__.py:400-419): common points of two
 * species' grids by exact coordinate equality. The reference compares every pair of
 * points (O(Na Nb)); here the smaller grid's points are hashed (O(Na + Nb)) and the result
 * is the same list: rows (i, j), i in the larger grid (grid a on a tie), j in the
 * other, ordered by i then j. Pairs are cached per unordered (a, b); with three or more
 * grids every pair is kept separately (the reference's pair slot formula collides there). */
void *MultiGrid_new(int nspecies);
void MultiGrid_del(void *multi_grid);
int MultiGrid_add_grid(void *multi_grid, void *grid);
/* Borrowed [size x 2] int array, valid until MultiGrid_del. a == b: rows (i, i). */
int *MultiGrid_get_common_index(void *multi_grid, int id_grid_a, int id_grid_b);
/* Row count of a computed pair, -1 before MultiGrid_get_common_index was called for it. */
int MultiGrid_get_common_index_size(void *multi_grid, int id_grid_a, int id_grid_b);
int MultiGrid_get_ngrids(void *multi_grid);
int MultiGrid_get_nspecies(void *multi_grid);

/* ------------------------------------------- splines, extrapolation, ODE (horton) */
/* src/horton/cubic_spline.cpp:81-103 */
void solve_cubic_spline_system(double *y, double *dt, int npoint);
/* src/horton/cubic_spline.cpp:138-181; aliases y and dt (caller keeps them alive) */
void *CubicSpline_new(double *y, double *dt, void *extrapolation, void *rtf, int n);
void CubicSpline_eval(void *spline, double *new_x, double *new_y, int new_n);
/* src/horton/cubic_spline.cpp:183-214 (napmo/__init__.py:446-448): dy/dr */
void CubicSpline_eval_deriv(void *spline, double *new_x, double *new_dx, int new_n);
/* napmo/__init__.py:456-460: the transform / extrapolation the spline was built with */
void *CubicSpline_get_rtransform(void *spline);
void *CubicSpline_get_extrapolation(void *spline);
/* src/horton/cubic_spline.cpp:105-131 (napmo/__init__.py:574-575): integration weights of
 * the cubic spline on t = 0..npoint-1 */
void compute_cubic_spline_int_weights(double *weights, int npoint);
double CubicSpline_get_first_x(void *spline);
double CubicSpline_get_last_x(void *spline);
void CubicSpline_del(void *spline);
/* src/horton/extrapolation.cpp:60-183 */
void *PotentialExtrapolation_new(int l);
void *ZeroExtrapolation_new(void);
void *CuspExtrapolation_new(void);
void *PowerExtrapolation_new(double power);
void Extrapolation_del(void *extrapolation);
/* src/horton/extrapolation.cpp:11-37 (napmo/__init__.py:468-488) */
void Extrapolation_prepare(void *extrapolation, void *spline);
double Extrapolation_eval_left(void *extrapolation, double x);
double Extrapolation_eval_right(void *extrapolation, double x);
double Extrapolation_deriv_left(void *extrapolation, double x);
double Extrapolation_deriv_right(void *extrapolation, double x);
/* true for the power-law kinds (Power, Potential), src/horton/extrapolation.h:64-149 */
bool Extrapolation_has_tail(void *extrapolation);
/* napmo/__init__.py:496-497 */
double PowerExtrapolation_get_power(void *extrapolation);
/* src/horton/ode2.cpp:232-321: coeffs [2n x 2n] (+=, pre-zeroed) and rhs [2n] */
void build_ode2(double *b, double *a, double *f, double *bcs, double *coeffs, double *rhs,
                long npoint);
/* src/horton/cell.cpp:407; only the non-periodic cell (nvec = 0) is supported */
void *Cell_new(double *rvecs, int nvec);
void Cell_del(void *cell);
/* src/horton/evaluate.cpp:129-223: output[npoint] += sum_lm U_lm(d) Y_lm, on the GPU */
void eval_decomposition_grid(void **splines, double *center, double *output, double *points,
                             void *cell, long nspline, long npoint);
/* src/horton/dot_multi.cpp:148-186 (bound at napmo/__init__.py:583-585): output[seg][m] +=
 * sum over the segment's points of prod_v data[v][p] * poly_m(points[p] - center), mtype
 * 1 Cartesian, 2 pure, 3 radial, 4 pure of the unit vector; nmoment must equal the count for
 * (lmax, mtype) and lmax <= 16. Same accumulation order as the reference, on the GPU. */
void dot_multi_moments(long npoint, long nvector, double *data, double *points, double *center,
                       long lmax, long mtype, long *segments, double *output, long nmoment);

/* ------------------------------------------------------------ Gaussian basis (gto.h) */
/* src/system/gto.cpp:17-51, 161-227; src/system/basis_set.cpp:12-82 */
void *PrimitiveGaussian_new(int *l, double *origin, double zeta, double coeff);
double PrimitiveGaussian_get_norma(void *p);
/* src/system/gto.cpp:85-129 (napmo/__init__.py:203-223). _compute writes out[size] at the
 * size points r [size x 3]; its double result (declared by the reference's ctypes table,
 * void in C) is 0. */
double PrimitiveGaussian_compute(void *p, double *r, double *output, int size);
double PrimitiveGaussian_overlap(void *p, void *other);
void PrimitiveGaussian_get_l(void *p, int *l);
void PrimitiveGaussian_get_origin(void *p, double *origin);
double PrimitiveGaussian_get_zeta(void *p);
double PrimitiveGaussian_get_coeff(void *p);
void *ContractedGaussian_new(void **primitives, int nprim);
double ContractedGaussian_get_norma(void *c);
/* src/system/gto.cpp:208-260 (napmo/__init__.py:229-243) */
int ContractedGaussian_get_nprim(void *c);
void ContractedGaussian_get_l(void *c, int *l);
void ContractedGaussian_get_origin(void *c, double *origin);
void ContractedGaussian_compute(void *c, double *r, double *output, int size);
double ContractedGaussian_overlap(void *c, void *other);
void *BasisSet_new(void **contractions, int n);
void *BasisSet_new_empty(void);
void BasisSet_update(void *basis, void *other);
int BasisSet_get_nbasis(void *basis);
/* src/system/basis_set.cpp:10-19, 115-119 (napmo/__init__.py:268-274): largest x power
 * (the shell l for complete Cartesian shells) and largest primitive count */
int BasisSet_get_max_l(void *basis);
int BasisSet_get_max_nprim(void *basis);
/* napmo/__init__.py:259-261, napmo/gto/basis_set.py:160-173: out [npoint x nbasis x 3],
 * the Cartesian gradient of every function at every point, on the GPU. (The reference's
 * src/system/basis_set.cpp:84-100 copies function values with indexes past its buffer;
 * the derivative is what the Python API documents.) */
void BasisSet_deriv(void *basis, double *r, double *out, int npoint);
/* out [npoint x nbasis] point-major; GPU collocation */
void BasisSet_compute(void *basis, double *r, double *out, int npoint);

/* --------------------------------------------- grid contractions (nwavefunction.cpp) */
/* WaveFunction: the by-reference struct of src/wavefunction/wavefunction.h:20-44
 * (== PSIB, napmo/wavefunction/psi_base.py:22-45). */
typedef struct WaveFunction {
  double *S, *X, *T, *V, *H, *C, *D, *L, *G, *J, *XC, *F, *O;
  int nbasis, ndim, occupation;
  double eta, kappa, x_factor, xc_energy, energy, rmsd;
} WaveFunction;
/* src/wavefunction/nwavefunction.cpp:15-75: psi->G = J-part + x_factor * K-part.
 * phi and K are function-major [ndim x N], J [N]. */
void nwavefunction_compute_2body_matrix(WaveFunction *psi, void *grid, double *phi, double *J,
                                        double *K);
/* nwavefunction.cpp:77-107: res[ndim x ndim] = sum_p phi_i phi_j otherJ bw w */
void nwavefunction_compute_coupling(WaveFunction *psi, void *grid, double *phi, double *other_J,
                                    double *res);
/* nwavefunction.cpp:161-197: psi->XC = sum_p phi_i phi_j vrho bw w */
void nwavefunction_compute_xc_matrix(WaveFunction *psi, void *grid, double *phi, double *xc_vrho);
/* nwavefunction.cpp:109-159 (bound at napmo/__init__.py:182-184): output[p] = rho from the
 * lower triangle of dm [nbasis x nbasis] with the basis collocated on the grid's points.
 * epsilon > 0 screens as the reference does, including its running threshold and its stop
 * at the first point whose bound falls below it (later points are left untouched);
 * dmmaxrow [nbasis] is read only then. */
void nwavefunction_compute_density_from_dm(void *basis, void *grid, double *dm, double *output,
                                           double epsilon, double *dmmaxrow);

/* ------------------------------------------------------- device-resident hot path */
/* Precision of the grid Fock contractions. */
enum {
  NAPMO_PREC_FP64 = 0,   /* everything float64 (FP64 MFMA) */
  NAPMO_PREC_FP32 = 1,   /* Phi and contractions float32 (FP32 MFMA), FP64 Poisson + traces */
  NAPMO_PREC_BF16X3 = 2, /* 3-term bf16 split on BF16 MFMA, FP32 accumulate, FP64 Poisson */
  NAPMO_PREC_INT8 = 3    /* F on int8 MFMA: 28-bit digit slices of sqrt|w bw| Phi (per geometry)
                            and of v times them (per build), exact int32 accumulation per chunk,
                            FP64 across chunks; rho, Poisson and traces FP64 */
};

/* Poisson: dens [N] (grid order) -> potential v [N] (napmo/wavefunction/ntwobody.py:15-46). */
int napmo_compute_coulomb(void *grid, const double *dens, int lmax, double *v_out);
/* Poisson, expansion only (napmo/grids/poisson_solver.py:18-118): for every centre A and
 * lm, the potential spline on A's radial grid: U_y, U_dt [sum_A nrad_A x L] (atom-major,
 * then radial, then lm). */
int napmo_poisson_solve(void *grid, const double *dens, int lmax, double *U_y, double *U_dt);
/* napmo/grids/poisson_solver.py:47-50 with sph_exp given: the radial solves of a
 * precomputed spherical expansion rho_lm [sum_A nrad_A x (lmax+1)^2] (atoms in grid order,
 * each atom's [nrad x L] block as AtomicGrid_spherical_expansion returns it, i.e. already
 * including the Becke weight), into U_y / U_dt as napmo_poisson_solve. */
int napmo_poisson_solve_expansion(void *grid, const double *rho_lm, int lmax, double *U_y,
                                  double *U_dt);
/* Kinetic operator on the grid (napmo/wavefunction/nkinetic.py:15-73): out [N] =
 * -1/2 laplacian of psi [N] from the Becke partition of psi, per (atom, l, m)
 * deriv2 - l(l+1)/r^2 of its spherical expansion, splined with PotentialExtrapolation(l)
 * and synthesised at every point. Chebyshev radial grids of >= 8 points. */
int napmo_compute_kinetic(void *grid, const double *psi, int lmax, double *out);
/* Its radial half: the spline values and slopes (y, dt) of deriv2 - l(l+1)/r^2 of every
 * (atom, l, m) component, U_y, U_dt [sum_A nrad_A x L] (as napmo_poisson_solve), and the
 * projected components themselves into rho_lm (same shape; may be NULL). */
int napmo_kinetic_expand(void *grid, const double *psi, int lmax, double *U_y, double *U_dt,
                         double *rho_lm);
/* Grid exchange potential (napmo/wavefunction/psi_numeric.py:186-215): for psi [ndim x N]
 * (function-major, grid order) and D [ndim x ndim], aux_sv = V[psi_s psi_v] for s >= v
 * (ndim(ndim+1)/2 Poisson solves) and K_out[j] = sum_k sum_l psi_k aux_lj D[k][l]
 * [ndim x N], the Kpot handed to nwavefunction_compute_2body_matrix. */
int napmo_compute_exchange_potential(void *grid, int ndim, const double *psi, const double *D,
                                     int lmax, double *K_out);
/* napmo_compute_exchange_potential and napmo_compute_kinetic on device pointers (resident
 * on the grid's device); they return when the result is complete. */
int napmo_compute_exchange_potential_device(void *grid, int ndim, const double *psi,
                                            const double *D, int lmax, double *K_out);
int napmo_compute_kinetic_device(void *grid, const double *psi, int lmax, double *out);
/* Point-charge potential on the grid (napmo/wavefunction/nnuclear.py:15-37): V_out [N] =
 * sum_A charges[A] / |r - origins[A]|, origins [ncharge x 3]. */
int napmo_compute_nuclear(void *grid, int ncharge, const double *charges, const double *origins,
                          double *V_out);
/* Becke-Dickson orbital correction (napmo/wavefunction/ndpsi.py:15-89, compute_dpsi), for
 * norb orbitals at once: phi, ri, V_tot [norb x N] (grid order), doi, oi [norb]; dpsi_out
 * [norb x N]. Per atom and (l, m) it solves u'' + (2/r) u' + a u = -2/m_inv mu_lm with
 * a = -2/m_inv (l(l+1)/r^2 + V_sph - oi - doi), u = 0 at both radial ends, mu the spherical
 * expansion of (phi doi - ri) bw and V_sph the spherical average of V_tot bw; the solutions
 * are synthesised at every point. Radial grids of 3 to 240 points. */
int napmo_compute_dpsi(void *grid, int lmax, int norb, const double *phi, const double *doi,
                       const double *oi, const double *ri, const double *V_tot, double mass_inv,
                       double *dpsi_out);
/* The per-orbital steps of PSIO_BECKE.optimize (napmo/wavefunction/psi_optimization.py
 * :175-203): R = T + V_tot - e_prev psi (Rgrid), de = integrate(R psi) (delta_e), then
 * dpsi = compute_dpsi(psi, de, e_prev, R, V_tot). psi, T, V_tot [norb x N]; outputs dpsi
 * [norb x N], de [norb], R [norb x N] (R_out may be NULL). */
int napmo_becke_dickson_step(void *grid, int lmax, int norb, const double *psi, const double *T,
                             const double *V_tot, const double *e_prev, double mass_inv,
                             double *dpsi_out, double *de_out, double *R_out);
/* Same on device pointers (all arrays resident on the grid's device, R_out required);
 * returns after the step has finished. */
int napmo_becke_dickson_step_device(void *grid, int lmax, int norb, const double *psi,
                                    const double *T, const double *V_tot, const double *e_prev,
                                    double mass_inv, double *dpsi_out, double *de_out,
                                    double *R_out);
/* epc17-2 electron-proton correlation (napmo/functional/interspecies_correlation.py:36-66),
 * elementwise on n densities: c_ene = -rhoN / (a - b sqrt(rhoE rhoN) + c rhoE rhoN) and the
 * potentials c_pot_a = d(rhoE c_ene)/d rhoE, c_pot_b = d(rhoE c_ene)/d rhoN, with the
 * reference's choice of formula when any density is below 1e-200. */
int napmo_epc17_2(long n, const double *rhoE, const double *rhoN, double *c_ene, double *c_pot_a,
                  double *c_pot_b);
/* The functional part of compute_c_2species_grid (napmo/wavefunction/psi_analytic.py
 * :214-262): index [npair x 2] common points (MultiGrid_get_common_index) of grid_a
 * (electrons, column 0) and grid_b (column 1), rho_a [Na], rho_b [Nb] the densities on
 * the grids; vrho_a[i] += c_pot_a, vrho_b[j] += c_pot_b and energy = sum over the pairs of
 * bw_b w_b c_ene rho_a. Every index is checked against the grid sizes. */
int napmo_epc17_2_grid(void *grid_a, void *grid_b, const int *index, int npair,
                       const double *rho_a, const double *rho_b, double *vrho_a, double *vrho_b,
                       double *energy);
/* No counterpart in the reference: napmo_epc17_2_grid for callers that cannot allocate
 * (HIP graph capture, tight serving loops). All arrays are device pointers: d_index
 * [npair x 2] int, d_rho_a [grid_a size], d_rho_b [grid_b size], d_vrho_a / d_vrho_b (+=),
 * d_energy [1]. Runs on grid_b's stream without synchronising; calls no allocator. Its
 * scratch (the small-density flag and 256 partial energies, 2.3 KB) is a static device array
 * per GPU shared by every call: calls that overlap on one GPU race. Index rows outside the
 * grids are skipped (napmo_epc17_2_grid reports them instead). */
int napmo_epc17_2_grid_noalloc(void *grid_a, void *grid_b, const int *d_index, int npair,
                               const double *d_rho_a, const double *d_rho_b, double *d_vrho_a,
                               double *d_vrho_b, double *d_energy);
/* Waits for the work queued on a grid's stream (device-pointer grid entry points do not). */
int napmo_grid_sync(void *grid);
/* HIP-event timing of the orbital corrections on this grid: start (enable = 1) or stop;
 * read gives the mean milliseconds per call of each phase (residual, sources and projection,
 * shell averages, radial solves, synthesis tables, synthesis, reduction) since the start. */
int napmo_grid_profile(void *grid, int enable);
int napmo_grid_profile_read(void *grid, double *avg_ms, int nphase, int *ncalls);

/* Grid Fock plan: device-resident Phi (collocated once), Poisson plan, workspace. */
int napmo_fock_plan_new(void *grid, void *basis, int lmax, int prec_mode, void **plan_out);
void napmo_fock_plan_del(void *plan);
int napmo_fock_plan_info(void *plan, long *npoint, int *nbasis, int *ncenter, int *lmax);
/* One grid Fock build from host P [n x n]: rho = diag(Phi P Phi^T), v = Poisson(rho),
 * F = Phi^T diag(w bw v) Phi [n x n], E = 1/2 sum P o F (FP64 trace). Any output may be
 * NULL. rho_out / v_out [N] on request. */
int napmo_grid_fock_build(void *plan, const double *P, double *F_out, double *E_out,
                          double *rho_out, double *v_out);
/* Same on device pointers, asynchronous on the plan's stream (no host sync). */
int napmo_grid_fock_build_device(void *plan, const double *dP, double *dF, double *dE);
/* Orbital form of the density (napmo/wavefunction/psi_numeric.py:361-369: Dgrid = eta psi^2
 * with psi = Phi C): C [n x nocc] row-major, rho = eta sum_i (Phi C)_i^2 in FP64,
 * P = eta C C^T for the trace. E_out[0] = 1/2 sum P o F, E_out[1] = 1/2 sum_p w bw rho v
 * (FP64 in every mode; equal in exact arithmetic). */
int napmo_grid_fock_build_orbitals(void *plan, const double *C, int nocc, double eta,
                                   double *F_out, double *E_out, double *rho_out, double *v_out);
int napmo_grid_fock_build_orbitals_device(void *plan, const double *dC, int nocc, double eta,
                                          double *dF, double *dE);
/* No reference counterpart: runs one wavefront through an MFMA builtin (which = 0:
 * f64 16x16x4, 1: bf16 16x16x32) and writes D [16 x 16] with the lane layout the
 * kernels assume, for the layout test. */
int napmo_selftest_mfma(int which, double *out);
/* No reference counterpart: plain device copy of n doubles (n even) with a streaming
 * kernel, used to calibrate the HBM byte counters (FETCH_SIZE / WRITE_SIZE). */
int napmo_device_copy(double *dst, const double *src, long n);
/* Grid overlap S = Phi^T diag(w bw) Phi [n x n] from the plan's resident Phi (FP64).
 * Same contraction as nwavefunction_compute_coupling with other_J = 1. */
int napmo_grid_overlap(void *plan, double *S_out);
/* F_out [n x n] = Phi^T diag(w bw v) Phi for a host potential v [N] of any sign, with the
 * plan's precision and contraction kernel: the coupling matrix of
 * nwavefunction_compute_coupling (src/wavefunction/nwavefunction.cpp:77-107) on the
 * plan's resident Phi. */
int napmo_fock_plan_contract(void *plan, const double *v, double *F_out);
/* Per-phase HIP-event timing of the next max_builds builds on the plan's stream; read
 * back the average milliseconds of each phase (napmo_fock_phase_name(k) names them). */
int napmo_fock_plan_profile(void *plan, int max_builds);
int napmo_fock_plan_profile_read(void *plan, double *avg_ms, int nphase, int *nbuild);
const char *napmo_fock_phase_name(int k);
/* Device memory for callers without a GPU framework (kind 1 H2D, 2 D2H, 3 D2D). */
void *napmo_device_malloc(size_t bytes);
void napmo_device_free(void *ptr);
int napmo_memcpy(void *dst, const void *src, size_t bytes, int kind);
/* Plan's HIP stream (hipStream_t) and a full device synchronisation of it. */
void *napmo_fock_plan_stream(void *plan);
int napmo_fock_plan_sync(void *plan);

/* Gradients of the build's energy E = 1/2 sum_p w bw rho v (no reference counterpart: the
 * reference has no grid gradients). Basis centres are the distinct function origins in
 * order of first appearance; xyz [ncentre x 3] may be NULL. */
int napmo_fock_plan_centres(void *plan, int *ncentre, double *xyz);
/* Derivatives of the last build's E_grid = 1/2 sum_p w bw rho v = 1/2 rho^T W V rho, exact for
 * the discrete Becke-Poisson operator V (no counterpart in the reference; the backward of the
 * torch glue). With ssym = (W v + V^T (W rho)) / 2:
 *  - napmo_grid_fock_gradient: dE/dR_c [ncentre x 3], the grid fixed and every basis function
 *    moving with its centre; the last build's density matrix P must be passed again; grad phi
 *    is recomputed on the fly;
 *  - napmo_grid_fock_energy_dP: dE/dP [n x n] = Phi^T diag(ssym) Phi (FP64); the build's own F
 *    uses W v alone, so F differs from dE/dP by the asymmetry of V;
 *  - napmo_grid_poisson_adjoint: z [N] = V^T y [N] on the plan's grid and lmax.
 * The _device forms take device pointers and return without synchronising. */
int napmo_grid_fock_gradient(void *plan, const double *P, double *grad_out);
int napmo_grid_fock_gradient_device(void *plan, const double *dP, double *dG);
int napmo_grid_fock_energy_dP(void *plan, double *dEdP_out);
int napmo_grid_fock_energy_dP_device(void *plan, double *d_dEdP);
int napmo_grid_poisson_adjoint(void *plan, const double *y, double *z_out);

/* Launch-shape knobs of a plan, for the autotuner (no reference counterpart). Set them
 * before a batch is captured; not while a build is running. */
enum {
  NAPMO_TUNE_FOCK_WG = 0,  /* workgroups of the per-tile Fock kernels (chunks x tiles); 0 = default */
  NAPMO_TUNE_PASS_WG = 1,  /* workgroups of the one-pass BF16x3 Fock kernel; 0 = default */
  NAPMO_TUNE_USE_PASS = 2, /* 0: per-tile BF16x3 kernel, 1: one-pass BF16x3 kernel,
                              2: one-pass kernel on G = sqrt|s| Phi (F = G^T diag(sign s) G) */
  NAPMO_TUNE_DENSITY_KS = 3, /* K per LDS stage of the density kernels: 16 or 32; 0 = 16 */
  NAPMO_TUNE_DENSITY_PF = 4, /* stages in flight in the density kernels: 1 or 2; 0 = 1 */
  NAPMO_TUNE_SYNTH_NP = 5,   /* (point, centre) pairs per lane of the synthesis, 1 or 2; 0 = 2
                                (a property of the grid: shared by every plan on it) */
  NAPMO_TUNE_DENSITY_STREAM = 6, /* orbital-form density: 1..4 keep C in LDS and stream Phi when
                                   C fits (nocc <= 64), in one of three launch shapes (1: default);
                                   0 uses the LDS-staged kernels */
  NAPMO_TUNE_INT8_CHUNK = 7,  /* int8 mode: points per chunk (a multiple of 64, at most 4032;
                                 0 = about N / 256). Setting it re-slices the per-geometry digits */
  NAPMO_TUNE_INT8_WAVE_BLOCKS = 8, /* int8 mode: 16 x 16 blocks of F per wave in the row-group
                                     packing, 1..5; 0 = 5 (fewer: more, smaller row groups) */
  NAPMO_TUNE_SYNTH_MFMA_MIN = 9, /* many-density synthesis (exchange potential, orbital
                                   corrections): whole tiles of 16 densities, and a last tile
                                   of at least this many, run on FP64 MFMA, the others one at a
                                   time; 1..16, 17 = one at a time only; 0 = 10 (a property of
                                   the grid) */
  NAPMO_TUNE_BACKWARD_POINTS = 10, /* energy derivatives (dE/dP, dE/dR): 0 = from the stored
                                     Phi with Phi P for all points at once; k > 0 = Phi
                                     recollocated in chunks of about k points (rounded up to
                                     a multiple of the contraction chunks), neither the stored
                                     Phi nor a full Phi P read: the same results bit for bit */
  NAPMO_TUNE_INT8_SCREEN = 11, /* int8 mode: S > 0 skips, per chunk of points, the 16 x 16 blocks
                                 of F whose chunk sum is bounded by 2^-S in every element (from
                                 the chunk maxima of |sqrt(w bw) Phi| and |v|), so each element
                                 moves by at most (chunks) 2^-S; 0 = off (default). Not a
                                 launch shape: results change within that bound */
  NAPMO_TUNE_DENSITY_SCREEN = 12, /* orbital-form density on the streamed kernel (nocc <= 64):
                                    S > 0 skips, per 16-point tile, the 16-function groups whose
                                    terms of X = Phi C are bounded by 2^-S (from the tile maxima
                                    of |Phi| and the group's sums of max |C|), so each X_pc moves
                                    by at most (ld / 16) 2^-S; keeps a copy of Phi with its
                                    columns reordered (compact functions first); 0 = off
                                    (default). Not a launch shape: rho changes within that bound */
  NAPMO_TUNE_INT8_PACK = 13 /* int8 mode: 1 = the blocks each chunk keeps (all, or those
                               NAPMO_TUNE_INT8_SCREEN keeps) are packed into as few workgroup
                               tasks as the LDS and wave slots allow, taken in turn by one
                               workgroup per compute unit; a chunk's tasks are formed again
                               only when its |v| exponent grows; 0 = fixed row groups
                               (default). A launch shape: the same results bit for bit */
};
int napmo_fock_plan_set_tuning(void *plan, int key, long value);
int napmo_fock_plan_get_tuning(void *plan, int key, long *value);
/* int8 plan: the upper 16 x 16 blocks x chunks of points that the last contraction computed
 * under NAPMO_TUNE_INT8_SCREEN (*kept) and all of them (*total); equal when screening is off.
 * Synchronises the plan's stream. No reference counterpart (measurement). */
int napmo_fock_plan_int8_screen_stats(void *plan, long *kept, long *total);
/* The grid-level keys (NAPMO_TUNE_SYNTH_NP, NAPMO_TUNE_SYNTH_MFMA_MIN) on a BeckeGrid handle. */
int napmo_grid_set_tuning(void *grid, int key, long value);
int napmo_grid_get_tuning(void *grid, int key, long *value);

/* Batched builds of independent molecules (no reference counterpart; the reference
 * runs molecules one at a time). Each plan keeps its own grid, basis and buffers.
 * nlanes 1..16: the batch spreads the plans' own launches over nlanes HIP streams forked
 * from and joined to one origin stream. nlanes 0: one launch per phase for all molecules
 * (FP64 plans with the per-tile contraction): the molecules' atoms and points side by
 * side in one grid for the Poisson steps (each molecule keeps its own lmax and its own
 * (point, centre) pairs), grouped launches for the per-molecule padding, density,
 * centre sums, Fock partials, reduction and traces; the results equal the plans' own
 * builds bit for bit. Either batch can capture one build of every plan into a HIP graph:
 * after napmo_batch_capture, napmo_batch_replay re-runs all builds on the same device
 * pointers (write the next P into them between replays). The plans must outlive the
 * batch. */
int napmo_batch_plan_new(int nplan, void **plans, int nlanes, void **batch_out);
void napmo_batch_plan_del(void *batch);
int napmo_batch_build(void *batch, const double **dP, double **dF, double **dE);
int napmo_batch_capture(void *batch, const double **dP, double **dF, double **dE);
/* Orbital form (nlanes >= 1): dC[m] [n_m x nocc[m]] row-major, P_m = eta C C^T; the same
 * results as the plans' own orbital-form builds. */
int napmo_batch_build_orbitals(void *batch, const double **dC, const int *nocc, double eta,
                               double **dF, double **dE);
int napmo_batch_capture_orbitals(void *batch, const double **dC, const int *nocc, double eta,
                                 double **dF, double **dE);
int napmo_batch_replay(void *batch);
int napmo_batch_sync(void *batch);

/* Multi-species (APMO) Coulomb build: species s (grid s, basis s, charge q_s, lmax s), one
 * Becke grid per species (two species may pass the same grid). Per build, with
 * rho_s from P_s:
 *   v_eff_s = q_s^2 V[rho_s] + q_s sum_{t != s} q_t V[rho_t]  at s's points
 *   F_s = Phi_s^T diag(w bw v_eff_s) Phi_s  = q_s^2 G_s(Coulomb) + J_s(coupling)
 * replacing compute_coulomb + nwavefunction_compute_2body_matrix (charge^2) and the
 * per-pair nwavefunction_compute_coupling loop of napmo/wavefunction/psi_numeric.py:152-241
 * and psi_hybrid.py:81-110 (the reference shares the electronic grid,
 * napmo/solver/hf_solver.py:95-98; V[rho_t] here is t's expansion synthesised at s's
 * points, which equals the reference's when the grids are shared).
 * E_out [4 nspecies]: 1/2 sum P_s o F_s, 1/2 sum w bw rho_s v_eff_s, the own-Coulomb part
 * 1/2 q_s^2 sum w bw rho_s V[rho_s], and the coupling part (second minus third). */
int napmo_apmo_plan_new(int nspecies, void **grids, void **bases, const double *charges,
                        const int *lmax, int prec_mode, void **plan_out);
void napmo_apmo_plan_del(void *plan);
/* Host pointers: P[s] [n_s x n_s]; F_out[s] [n_s x n_s] (entries may be NULL). */
int napmo_apmo_build(void *plan, const double **P, double **F_out, double *E_out);
/* Device pointers; asynchronous on the species' streams (napmo_apmo_sync waits). */
int napmo_apmo_build_device(void *plan, const double **dP, double **dF, double *dE);
int napmo_apmo_sync(void *plan);
/* on = 1: later builds run every phase for all species as one launch (FP64 plans, no
 * reference counterpart): the species' atoms and points side by side, the synthesis of
 * own and cross-species potentials over one pair list per lmax, grouped density, Fock,
 * reduction and trace kernels. Equal to the per-species build to rounding. on = 0: back
 * to one stream per species. */
int napmo_apmo_plan_set_one_launch(void *plan, int on);
/* napmo_fock_plan_set_tuning on species s's Fock plan (no reference counterpart), e.g.
 * NAPMO_TUNE_INT8_SCREEN; not in one-launch mode. */
int napmo_apmo_plan_set_tuning(void *plan, int species, int key, long value);
/* One-launch mode only: record one build on fixed device pointers into a HIP graph;
 * napmo_apmo_replay re-runs it (write the next P into the same buffers between replays);
 * napmo_apmo_sync waits. */
int napmo_apmo_capture(void *plan, const double **dP, double **dF, double *dE);
int napmo_apmo_replay(void *plan);

#ifdef __cplusplus
}
#endif
#endif /* NAPMO_GPU_H */
