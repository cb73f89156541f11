// Multi-species (APMO) grid Coulomb build: several species (electrons, quantum nuclei),
// each on its own Becke grid with its own basis, coupled through their Poisson potentials.
#pragma once

#include <memory>
#include <vector>

#include "fock_plan.hpp"
#include "group_plan.hpp"

namespace napmo {

// Per species s with charge q_s and density rho_s (from P_s):
//   v_eff_s = q_s^2 V[rho_s] + q_s sum_{t != s} q_t V[rho_t]   at s's points,
//   F_s     = Phi_s^T diag(w bw v_eff_s) Phi_s,
// i.e. q_s^2 times the species' own Coulomb matrix (napmo/wavefunction/psi_numeric.py:
// 152-169, G *= charge^2) plus its coupling matrix (psi_numeric.py:217-241,
// psi_hybrid.py:81-110). V[rho_t] at s's points is t's Becke-Poisson expansion
// synthesised there; when both species share one grid this is the reference's coupling.
// Every species runs on its own grid's stream; a species waits (HIP event) for the
// expansion of the others before its cross-species synthesis, so the species overlap.
struct ApmoPlan {
  struct Species {
    std::unique_ptr<FockPlan> plan;
    double charge = 0.0;
    DeviceBuffer<double> vcross, veff, E2;   // [N], [N], self trace [2]
    std::vector<std::unique_ptr<SynthPairs>> cross;  // [nspecies]: t's centres at s's points
    hipEvent_t expanded = nullptr, done = nullptr;
  };
  std::vector<Species> sp;
  DeviceBuffer<double> E;  // [4 * nspecies] for host-pointer builds
  // one launch per phase for all species (FP64): the species' atoms and points side by
  // side, own and cross-species pairs in one synthesis per lmax (GroupPlan, coupled)
  std::unique_ptr<GroupPlan> group;
  void set_one_launch(bool on);
  // One-launch mode: record one build on fixed device pointers into a HIP graph (every
  // species at once); replay() re-runs it.
  void capture(const double *const *dP, double *const *dF, double *dE);
  void replay();

  ApmoPlan(const std::vector<DeviceGrid *> &grids, const std::vector<const BasisSet *> &bases,
           const std::vector<double> &charges, const std::vector<int> &lmax, int prec);
  ~ApmoPlan();
  int nspecies() const { return (int)sp.size(); }
  // dP[s] [n_s x n_s], dF[s] [n_s x n_s], dE [4 * nspecies] on the device:
  // E[4s + 0] = 1/2 sum P_s o F_s, E[4s + 1] = 1/2 sum w bw rho_s v_eff_s,
  // E[4s + 2] = 1/2 q_s^2 sum w bw rho_s V[rho_s], E[4s + 3] = E[4s + 1] - E[4s + 2].
  void build_device(const double *const *dP, double *const *dF, double *dE);
  void synchronize();
  // Host P[s] / F[s] (any F may be NULL), E_out [4 * nspecies] (may be NULL).
  void build_host(const double *const *P, double *const *F_out, double *E_out);
};

}  // namespace napmo
