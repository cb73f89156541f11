// Nuclear-centre gradient of the grid Coulomb energy E = 1/2 sum_p w_p bw_p rho_p v_p,
// rho_p = sum_{mu nu} phi_mu(p) P_mu nu phi_nu(p), on a fixed grid with every basis function
// moving with its centre c:
//
//   dE/dR_c = sum_p s_p d rho_p / dR_c,   s = w bw v,
//   d rho_p / dR_c = -2 sum_{mu on c} (grad phi_mu)(p) X_p mu,   X = Phi P,
//
// with v the derivative of E with respect to rho_p / (w bw) (exact when the discrete
// Coulomb operator is self-adjoint in the w bw inner product; the same approximation makes
// dE/dP the grid Fock matrix). grad phi is recomputed analytically here, not stored:
// for phi = N_c x^a y^b z^c sum_k c_k e^{-zeta_k r^2},
//   d phi / dx = N_c y^b z^c (a x^{a-1} g + x^{a+1} h),  g = sum_k c_k e_k,  h = -2 sum_k c_k zeta_k e_k.
#include "common.hpp"
#include "kernels.hpp"

namespace napmo {

namespace {

__device__ inline double ipow_g(double v, int e) {
  double r = 1.0;
  for (int k = 0; k < e; ++k) r *= v;
  return r;
}

// Partial sums over one chunk of points for one function: partial[(mu cstride + c0 + c) 3 + k]
// (cstride = nchunk, c0 = 0 for one call over all points; a call over the points from
// 4096 c0 on writes its chunks at c0 ...).
__global__ void __launch_bounds__(256) k_grad_basis(
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    long N, long chunk, int cstride, int c0, const double *__restrict__ s,
    const double *__restrict__ X, int ldx, const double *__restrict__ fo,
    const int *__restrict__ fl, const int *__restrict__ ff, const int *__restrict__ fnp,
    const double *__restrict__ fnorm, const double *__restrict__ pz, const double *__restrict__ pc,
    double *__restrict__ partial) {
  __shared__ double red[4][3];
  const int mu = blockIdx.y, c = blockIdx.x;
  const double ox = fo[3 * mu], oy = fo[3 * mu + 1], oz = fo[3 * mu + 2];
  const int la = fl[3 * mu], lb = fl[3 * mu + 1], lc = fl[3 * mu + 2];
  const int f0 = ff[mu], np = fnp[mu];
  const long pbeg = (long)c * chunk, pend = min(N, pbeg + chunk);
  double gx = 0.0, gy = 0.0, gz = 0.0;
  for (long p = pbeg + threadIdx.x; p < pend; p += 256) {
    const double dx = x[p] - ox, dy = y[p] - oy, dz = z[p] - oz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    double g = 0.0, h = 0.0;
    for (int k = 0; k < np; ++k) {
      const double e = pc[f0 + k] * exp(-pz[f0 + k] * r2);
      g += e;
      h -= 2.0 * pz[f0 + k] * e;
    }
    const double px = ipow_g(dx, la), py = ipow_g(dy, lb), pzv = ipow_g(dz, lc);
    const double ax = (la ? la * ipow_g(dx, la - 1) : 0.0) * g + px * dx * h;
    const double ay = (lb ? lb * ipow_g(dy, lb - 1) : 0.0) * g + py * dy * h;
    const double az = (lc ? lc * ipow_g(dz, lc - 1) : 0.0) * g + pzv * dz * h;
    const double wgt = s[p] * X[p * ldx + mu] * fnorm[mu];
    gx += wgt * ax * py * pzv;
    gy += wgt * px * ay * pzv;
    gz += wgt * px * py * az;
  }
  for (int o = 32; o > 0; o >>= 1) {
    gx += __shfl_xor(gx, o);
    gy += __shfl_xor(gy, o);
    gz += __shfl_xor(gz, o);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave][0] = gx, red[wave][1] = gy, red[wave][2] = gz;
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    partial[((long)mu * cstride + c0 + c) * 3 + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
  }
}

// grad[centre][k] = -2 sum over the centre's functions and the chunks (fixed order).
__global__ void k_grad_reduce(const double *__restrict__ partial, int n, int nchunk,
                              const int *__restrict__ centre_of, int ncentre,
                              double *__restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ncentre * 3) return;
  const int c = e / 3, k = e % 3;
  double sum = 0.0;
  for (int mu = 0; mu < n; ++mu) {
    if (centre_of[mu] != c) continue;
    for (int q = 0; q < nchunk; ++q) sum += partial[((long)mu * nchunk + q) * 3 + k];
  }
  grad[e] = -2.0 * sum;
}

}  // namespace

constexpr long GRAD_CHUNK = 4096;
long grad_chunk_points() { return GRAD_CHUNK; }

int grad_chunks(long N) { return (int)std::max<long>(1, div_up(N, GRAD_CHUNK)); }

void launch_grad_centres(const double *x, const double *y, const double *z, long N, const double *s,
                         const double *X, int ldx, int n, const double *fo, const int *fl,
                         const int *ff, const int *fnp, const double *fnorm, const double *pz,
                         const double *pc, const int *centre_of, int ncentre, double *partial,
                         double *grad, hipStream_t st) {
  if (n <= 0 || ncentre <= 0) return;
  const int nchunk = grad_chunks(N);
  if (N > 0)
    k_grad_basis<<<dim3(nchunk, n), 256, 0, st>>>(x, y, z, N, GRAD_CHUNK, nchunk, 0, s, X, ldx, fo,
                                                   fl, ff, fnp, fnorm, pz, pc, partial);
  k_grad_reduce<<<div_up(ncentre * 3, 64), 64, 0, st>>>(partial, n, N > 0 ? nchunk : 0, centre_of,
                                                        ncentre, grad);
  NAPMO_HIP(hipGetLastError());
}

void launch_grad_basis_part(const double *x, const double *y, const double *z, long N, long p0,
                            long Ntotal, const double *s, const double *X, int ldx, int n,
                            const double *fo, const int *fl, const int *ff, const int *fnp,
                            const double *fnorm, const double *pz, const double *pc,
                            double *partial, hipStream_t st) {
  NAPMO_REQUIRE(p0 % GRAD_CHUNK == 0, "gradient: point ranges must start on a 4096-point chunk");
  if (n <= 0 || N <= 0) return;
  k_grad_basis<<<dim3(grad_chunks(N), n), 256, 0, st>>>(
      x + p0, y + p0, z + p0, N, GRAD_CHUNK, grad_chunks(Ntotal), (int)(p0 / GRAD_CHUNK), s + p0,
      X, ldx, fo, fl, ff, fnp, fnorm, pz, pc, partial);
  NAPMO_HIP(hipGetLastError());
}

void launch_grad_reduce(const double *partial, int n, long Ntotal, const int *centre_of,
                        int ncentre, double *grad, hipStream_t st) {
  if (ncentre <= 0) return;
  k_grad_reduce<<<div_up(ncentre * 3, 64), 64, 0, st>>>(partial, n, Ntotal > 0 ? grad_chunks(Ntotal) : 0,
                                                        centre_of, ncentre, grad);
  NAPMO_HIP(hipGetLastError());
}

}  // namespace napmo
