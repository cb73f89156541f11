// Device kernels behind the legacy exports dot_multi_moments and
// nwavefunction_compute_density_from_dm.
#pragma once

#include "common.hpp"

namespace napmo {

constexpr int MOMENT_LMAX = 16;        // largest lmax dot_multi_moments accepts
constexpr int MOMENT_PER_THREAD = 4;   // nmoment <= 1024 (Cartesian lmax 16: 969)

// Moments per segment for (lmax, mtype): 1 + the polynomials the reference fills.
int moment_count(int lmax, int mtype);

// output[seg_index[b] * nmoment + m] += sum over points of segment b (in point order) of
// term_p * poly_m(p - center); points [npoint x 3], data [nvector x npoint] (device).
void launch_dot_multi_moments(const long *seg_start, const long *seg_stop, const long *seg_index,
                              int nseg, long npoint, long nvector, const double *data,
                              const double *points, const double *center, int lmax, int mtype,
                              int nmoment, double *output, hipStream_t s);

// rho from a lower-triangular DM with the reference's screening (eps > 0) on Phi
// point-major [N x ld]; work: absmax, upper, thr [N], stop [1]. With screening only points
// before *stop are written.
void launch_dm_density(const double *phi, int ld, long N, int n, const double *dm,
                       const double *dmmaxrow, double eps, double *absmax, double *upper,
                       double *thr, long *stop, double *out, hipStream_t s);

// Cartesian gradient of every contracted function at every point (BasisSet_deriv):
// out[(p n + mu) 3 + k] = d phi_mu / d r_k at point p; points x, y, z [N] (device), the
// flattened basis arrays as for the collocation.
void launch_basis_gradient(const double *x, const double *y, const double *z, long N, int n,
                           const double *fo, const int *fl, const int *ff, const int *fnp,
                           const double *fnorm, const double *pz, const double *pc, double *out,
                           hipStream_t s);

}  // namespace napmo
