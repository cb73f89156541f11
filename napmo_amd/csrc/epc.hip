// epc17-2 electron-proton correlation functional (napmo/functional/interspecies_correlation.py
// :36-66) and its use on the common points of two species' grids
// (napmo/wavefunction/psi_analytic.py:214-262, compute_c_2species_grid):
//   den = a - b sqrt(rE rN) + c rE rN,  eps = -rN / den,  a, b, c = 2.35, 2.4, 6.6,
//   vE = d(rE eps)/d rE, vN = d(rE eps)/d rN,
// written in the reference's two algebraically equal forms; the form without 1/sqrt(rE rN)
// is taken for the whole array when any density on it is below 1e-200, as the reference
// decides. Contraction is off so every product and sum is rounded as the reference's numpy
// expression is.
#include <map>
#include <mutex>

#include "common.hpp"
#include "kernels.hpp"

namespace napmo {

namespace {

constexpr double EPC_A = 2.35, EPC_B = 2.4, EPC_C = 6.6, EPC_TINY = 1.0e-200;

// index rows outside [0, Na) x [0, Nb) are skipped (the host entry point rejects them first)
__device__ inline bool epc_pair_ok(const int *index, long k, long Na, long Nb) {
  const int i = index[2 * k], j = index[2 * k + 1];
  return i >= 0 && i < Na && j >= 0 && j < Nb;
}

__global__ void k_epc_small(const double *__restrict__ rE, const double *__restrict__ rN,
                            const int *__restrict__ index, long n, long Na, long Nb,
                            int *__restrict__ flag) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (index && !epc_pair_ok(index, k, Na, Nb)) return;
  const double e = index ? rE[index[2 * k]] : rE[k];
  const double p = index ? rN[index[2 * k + 1]] : rN[k];
  if (fabs(e) < EPC_TINY || fabs(p) < EPC_TINY) atomicOr(flag, 1);
}

__global__ void k_epc_clear(int *flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *flag = 0;
}

struct EpcOut {
  double ene, potE, potN;
};

__device__ inline EpcOut epc17_2(double rhoE, double rhoN, bool small) {
#pragma clang fp contract(off)
  const double s = sqrt(rhoE * rhoN);
  const double den = EPC_A - EPC_B * s + EPC_C * rhoE * rhoN;
  const double den2 = den * den;
  EpcOut o;
  o.ene = -rhoN / den;
  if (small) {
    o.potE = (EPC_B * sqrt(rhoE) * pow(rhoN, 1.5) - 2.0 * EPC_A * rhoN) / den2 / 2.0;
    o.potN = (EPC_B * sqrt(rhoN) * pow(rhoE, 1.5) - 2.0 * EPC_A * rhoE) / den2 / 2.0;
  } else {
    o.potE = (rhoE * rhoN * (EPC_C * rhoN - EPC_B * rhoN / (2.0 * s)) - rhoN * den) / den2;
    o.potN = (rhoN * rhoE * (EPC_C * rhoE - EPC_B * rhoE / (2.0 * s)) - rhoE * den) / den2;
  }
  return o;
}

__global__ void k_epc_arrays(const double *__restrict__ rE, const double *__restrict__ rN, long n,
                             const int *__restrict__ flag, double *__restrict__ ene,
                             double *__restrict__ potE, double *__restrict__ potN) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const EpcOut o = epc17_2(rE[k], rN[k], *flag != 0);
  ene[k] = o.ene, potE[k] = o.potE, potN[k] = o.potN;
}

// Common points (i, j) = index rows: vrhoE[i] += vE, vrhoN[j] += vN (each grid's points are
// distinct, so every i and every j occurs once), partial energies (wbwN[j] eps) rhoE[i].
constexpr int EPC_BLOCKS = 256;

__global__ void __launch_bounds__(256) k_epc_pairs(const double *__restrict__ rE,
                                                   const double *__restrict__ rN,
                                                   const int *__restrict__ index, long n,
                                                   long Na, long Nb, const int *__restrict__ flag,
                                                   const double *__restrict__ wbwN,
                                                   double *__restrict__ vE, double *__restrict__ vN,
                                                   double *__restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double red[4];
  const bool small = *flag != 0;
  double acc = 0.0;
  for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long)gridDim.x * 256) {
    if (!epc_pair_ok(index, k, Na, Nb)) continue;
    const int i = index[2 * k], j = index[2 * k + 1];
    const double e = rE[i];
    const EpcOut o = epc17_2(e, rN[j], small);
    vE[i] += o.potE;
    vN[j] += o.potN;
    acc += wbwN[j] * o.ene * e;
  }
  for (int s = 32; s > 0; s >>= 1) acc += __shfl_xor(acc, s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ void k_epc_final(const double *__restrict__ part, int nb, double *__restrict__ energy) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[b];
  energy[0] = s;
}

}  // namespace

void launch_epc17_2(const double *rE, const double *rN, long n, int *flag, double *ene,
                    double *potE, double *potN, hipStream_t st) {
  if (n <= 0) return;
  k_epc_small<<<div_up(n, 256), 256, 0, st>>>(rE, rN, nullptr, n, 0, 0, flag);
  k_epc_arrays<<<div_up(n, 256), 256, 0, st>>>(rE, rN, n, flag, ene, potE, potN);
  NAPMO_HIP(hipGetLastError());
}

int epc_pair_blocks(long n) { return (int)std::max<long>(1, std::min<long>(EPC_BLOCKS, div_up(n, 256))); }

void launch_epc17_2_pairs(const double *rE, const double *rN, const int *index, long n, long Na,
                          long Nb, const double *wbwN, int *flag, double *vE, double *vN,
                          double *part, double *energy, hipStream_t st) {
  const int nb = epc_pair_blocks(n);
  if (n > 0) k_epc_small<<<div_up(n, 256), 256, 0, st>>>(rE, rN, index, n, Na, Nb, flag);
  k_epc_pairs<<<nb, 256, 0, st>>>(rE, rN, index, n, Na, Nb, flag, wbwN, vE, vN, part);
  k_epc_final<<<1, 64, 0, st>>>(part, nb, energy);
  NAPMO_HIP(hipGetLastError());
}

// ---------------------------------------------------------- allocation-free variant
// Scratch of napmo_epc17_2_grid_noalloc: the partial energies (EPC_PART doubles) and the
// small-density flag, in one file-scope device array; each GPU holds its own copy, found once
// per device by hipGetSymbolAddress. Every call on a GPU shares it (calls that overlap on one
// GPU race); the flag is cleared by a kernel at the start of each call.
constexpr int EPC_SCRATCH_DOUBLES = EPC_PART + 32;
static __device__ __align__(256) double scratch[EPC_SCRATCH_DOUBLES];

static double *epc_scratch_ptr() {
  static std::mutex mu;
  static std::map<int, double *> ptrs;
  int dev = 0;
  NAPMO_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto it = ptrs.find(dev);
  if (it != ptrs.end()) return it->second;
  void *p = nullptr;
  NAPMO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  NAPMO_REQUIRE(p != nullptr, "epc17-2 scratch: no device address");
  ptrs[dev] = (double *)p;
  return (double *)p;
}

void launch_epc17_2_pairs_noalloc(const double *rE, const double *rN, const int *index, long n,
                                  long Na, long Nb, const double *wbwN, double *vE, double *vN,
                                  double *energy, hipStream_t st) {
  const int nb = epc_pair_blocks(n);
  NAPMO_REQUIRE(nb + 1 <= EPC_SCRATCH_DOUBLES,
                "epc17-2 (no allocation): the partial energies exceed the static scratch");
  double *s = epc_scratch_ptr();
  int *flag = (int *)(s + EPC_PART);
  k_epc_clear<<<1, 64, 0, st>>>(flag);
  launch_epc17_2_pairs(rE, rN, index, n, Na, Nb, wbwN, flag, vE, vN, s, energy, st);
}

}  // namespace napmo
