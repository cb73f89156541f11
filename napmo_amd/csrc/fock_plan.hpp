// Device-resident grid Fock build: Phi collocated once per geometry, then per build
// rho = diag(Phi P Phi^T), Poisson, F = Phi^T diag(w bw v) Phi, FP64 energy trace.
#pragma once

#include "napmo_gpu.h"
#include "common.hpp"
#include "host.hpp"
#include "kernels.hpp"
#include "ozaki.hpp"
#include "poisson_plan.hpp"

namespace napmo {

struct DeviceBasis {
  int nbasis = 0;
  DeviceBuffer<double> origin, norm, zeta, coef;
  DeviceBuffer<int> lxyz, first, nprim;
  void upload(const FlatBasis &fb, hipStream_t s);
  // Phi into out64 / out32 [Npad x ld] (either may be null).
  void collocate(const double *x, const double *y, const double *z, long N, long Npad, int ld,
                 double *out64, float *out32, hipStream_t s) const;
};

struct FockPlan {
  DeviceGrid &grid;  // borrowed: the BeckeGrid handle must outlive the plan
  int n, ld, lmax, prec;
  long Npad;
  DeviceBasis basis;
  PoissonWork work;  // this plan's own Poisson workspace (plans may share a grid)
  DeviceBuffer<double> phi64;          // [Npad x ld]
  DeviceBuffer<unsigned short> phiT_hi, phiT_lo;  // bf16 bits, function-major [ld x Npad]
  DeviceBuffer<float> phiT32;          // function-major [ld x Npad]
  DeviceBuffer<double> Cpad, Pfull;    // orbital-form inputs
  DeviceBuffer<double> P, rho_part, rho, v, s, F_part, F, E, trace_work;
  FockShape fsh, psh;  // per-tile shape; one-pass shape (mixed mode)
  // int8 mode: digit planes of sqrt|w bw| Phi, chunk exponents, per-stage max |d1|, the
  // per-build signed potential vs = sign(w bw) v, partials, wave tasks (+ group rows)
  OzakiLayout oz;
  DeviceBuffer<signed char> oz_planes;
  DeviceBuffer<int> oz_eA, oz_tasks;
  DeviceBuffer<unsigned char> oz_maxd1;
  DeviceBuffer<double> oz_vs, oz_part;
  long oz_chunk = 0;       // NAPMO_TUNE_INT8_CHUNK (0: default)
  int oz_screen = 0;       // NAPMO_TUNE_INT8_SCREEN (0: off)
  int oz_pack = 0;         // NAPMO_TUNE_INT8_PACK: kept blocks packed per chunk
  DeviceBuffer<int> oz_ptasks, oz_pctr;  // packed tasks and their two counters
  DeviceBuffer<int> oz_pcache, oz_pccount, oz_pcev;  // per-chunk packing kept across builds
  DeviceBuffer<int> oz_perm;  // screening order of the functions (empty: identity)
  int oz_wave_blocks = OZ_MAXB;  // NAPMO_TUNE_INT8_WAVE_BLOCKS
  void ozaki_setup();
  void ozaki_contract(const double *veff, double *dF, int slot);
  // int8 mode: (upper blocks x chunks computed by the last contraction, all of them)
  std::pair<long, long> int8_screen_stats();
  int use_pass = 0;  // 0: per-tile kernels, 1: one-pass BF16x3, 2: symmetric one-pass BF16x3
  long fock_wg = 0, pass_wg = 0;  // workgroup targets of the two shapes (0 = default)
  int density_ks = 16;            // K per LDS stage of the density kernels (16 or 32)
  int density_pf = 1;             // stages in flight in the density kernels (1 or 2)
  int density_stream = 1;  // orbital form: C resident in LDS, Phi streamed (when it fits), in
                           // launch shape 1..3; 0: the LDS-staged kernels
  int density_screen = 0;  // NAPMO_TUNE_DENSITY_SCREEN (0: off)
  DeviceBuffer<double> phi_scr;        // Phi with its columns in the order dens_perm
  DeviceBuffer<int> dens_perm;         // [ld]
  DeviceBuffer<signed char> dens_eT;   // [Npad / 16][ld / 16] block exponents of phi_scr
  void density_screen_setup();
  hipStream_t stream = nullptr;   // every build launch goes here (default: the grid's)
  void reshape();
  // Launch-shape knobs (NAPMO_TUNE_*), for the autotuner; not while a build is running.
  void set_tuning(int key, long value);
  long get_tuning(int key) const;

  // Per-phase HIP-event timing of builds (on the plan's stream), for roofline reports.
  static constexpr int NPHASE = 10;
  static const char *phase_name(int k);
  int prof_cap = 0, prof_used = 0;
  std::vector<hipEvent_t> prof_events;  // prof_cap x (NPHASE + 1)
  void mark(int build_slot, int phase);

  FockPlan(DeviceGrid &g, const BasisSet &b, int lmax, int prec_mode);
  ~FockPlan();
  void build_device(const double *dP, double *dF, double *dE);
  // Orbital form: rho = eta sum_i (Phi C)_i^2 (C [n x nocc] device), P = eta C C^T.
  void build_orbitals_device(const double *dC, int nocc, double eta, double *dF, double *dE);
  void build_orbitals_host(const double *C, int nocc, double eta, double *F_out, double *E_out,
                           double *rho_out, double *v_out);
  // Stages of one build (slot: profiling slot or -1), all on the grid's stream:
  // rho from P or from orbitals (density_C also forms P = eta C C^T for the trace),
  void density_P(int slot, const double *dP);
  void density_C(int slot, const double *dC, int nocc, double eta);
  // Poisson expansion of rho (projection, radial Green solves, synthesis tables),
  void expand(int slot);
  // v = synthesis of this grid's own expansion at its own points,
  // with_scale: also write s = w bw v (the contraction's weights) in the centre-sum pass
  void synth_self(int slot, bool with_scale = false);
  // F = Phi^T diag(w bw veff) Phi and the FP64 traces with veff [N] (P: unpadded n x n).
  // s_ready: s = w bw veff was already written (synth_self with_scale, veff = v)
  void contract(int slot, const double *dP, const double *veff, double *dF, double *dE,
                bool s_ready = false);
  int next_slot();
  void build_host(const double *P, double *F_out, double *E_out, double *rho_out, double *v_out);
  // S = Phi^T diag(w bw) Phi [n x n] from the resident Phi (FP64).
  void overlap(double *S_out);
  // F = Phi^T diag(w bw v) Phi for a given host potential v [N], with the plan's precision
  // and contraction kernel (the coupling matrix of src/wavefunction/nwavefunction.cpp:77-107).
  void contract_potential(const double *v_host, double *F_out);
  // Basis centres (distinct function origins, in order of first appearance) and the
  // centre of every function; the gradient is reported per centre.
  std::vector<double> centres;   // [ncentre x 3]
  DeviceBuffer<int> centre_of;   // [n]
  DeviceBuffer<double> X, grad_part;
  bool built = false;
  // Derivatives of E = 1/2 sum_p w bw rho v = 1/2 rho^T W V rho of the last build (grid
  // fixed, every function moving with its centre), exact for the discrete operator V: with
  // ssym = (W v + V^T (W rho)) / 2 = dE/drho,
  //   dE/dP = Phi^T diag(ssym) Phi   (energy_dP, FP64 contraction, dEdP [n x n] device)
  //   dE/dR_c = -2 sum_p ssym_p sum_{mu on c} grad phi_mu(p) (Phi P)_p mu
  //                                 (gradient_device, dG [ncentre x 3] device).
  // grad Phi is recomputed in the kernel, not stored. V^T is the transposed projection,
  // radial solve, synthesis tables and synthesis (adjoint_potential: z = V^T y on device).
  DeviceBuffer<double> adj_y, adj_z, adj_T, adj_aL, adj_aR, adj_Uy, adj_Udt, adj_rt, ssym;
  // NAPMO_TUNE_BACKWARD_POINTS: 0, or the points per chunk of recollocated Phi
  long bwd_points = 0;
  DeviceBuffer<double> bwd_phi, bwd_X;
  long backward_chunk() const;
  void adjoint_potential(const double *y, double *z);
  void symmetric_weight();
  void energy_dP(double *dEdP);
  void gradient_device(const double *dP, double *dG);
  void profile_start(int max_builds);
  // Average milliseconds per phase over the builds recorded since profile_start.
  int profile_read(double *avg_ms, int nphase);
};

}  // namespace napmo
