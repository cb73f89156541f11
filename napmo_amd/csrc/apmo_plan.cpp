// Multi-species (APMO) grid Coulomb build (see apmo_plan.hpp).
#include "apmo_plan.hpp"

#include <algorithm>
#include <cmath>
#include <string>

namespace napmo {

// E[2] = q^2 * E2[1] (own Coulomb energy), E[3] = E[1] - E[2] (coupling energy).
__global__ void k_apmo_energy(const double *__restrict__ E2, double q2, double *__restrict__ E) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    E[2] = q2 * E2[1];
    E[3] = E[1] - E[2];
  }
}

ApmoPlan::ApmoPlan(const std::vector<DeviceGrid *> &grids,
                   const std::vector<const BasisSet *> &bases, const std::vector<double> &charges,
                   const std::vector<int> &lmax, int prec) {
  const size_t ns = grids.size();
  NAPMO_REQUIRE(ns >= 1 && bases.size() == ns && charges.size() == ns && lmax.size() == ns,
                "apmo plan: one grid, basis, charge and lmax per species");
  sp.resize(ns);
  for (size_t s = 0; s < ns; ++s) {
    Species &S = sp[s];
    S.plan = std::make_unique<FockPlan>(*grids[s], *bases[s], lmax[s], prec);
    S.charge = charges[s];
    const long N = grids[s]->N;
    S.vcross.resize(N);
    S.veff.resize(N);
    S.E2.resize(2);
    NAPMO_HIP(hipEventCreateWithFlags(&S.expanded, hipEventDisableTiming));
    NAPMO_HIP(hipEventCreateWithFlags(&S.done, hipEventDisableTiming));
  }
  // cross-species synthesis lists: t's centres and radial grids at s's points
  for (size_t s = 0; s < ns; ++s) {
    DeviceGrid &gs = sp[s].plan->grid;
    sp[s].cross.resize(ns);
    for (size_t t = 0; t < ns; ++t) {
      if (t == s) continue;
      DeviceGrid &gt = sp[t].plan->grid;
      if (&gt == &gs) continue;  // shared grid: t's own potential is already at s's points
      sp[s].cross[t] = build_synth_pairs(gt, gs.x.get(), gs.y.get(), gs.z.get(), gs.N, gs.stream);
    }
  }
}

ApmoPlan::~ApmoPlan() {
  for (auto &S : sp) {
    if (S.expanded) (void)hipEventDestroy(S.expanded);
    if (S.done) (void)hipEventDestroy(S.done);
  }
}

void ApmoPlan::set_one_launch(bool on) {
  synchronize();
  if (!on) {
    group.reset();
    return;
  }
  std::vector<FockPlan *> plans;
  std::vector<double> charges;
  for (auto &S : sp) plans.push_back(S.plan.get()), charges.push_back(S.charge);
  group = std::make_unique<GroupPlan>(plans, &charges);
}

void ApmoPlan::capture(const double *const *dP, double *const *dF, double *dE) {
  NAPMO_REQUIRE(group != nullptr, "APMO capture: switch on the one-launch mode first");
  synchronize();
  const int ns = nspecies();
  std::vector<double *> dEs(ns);
  for (int s = 0; s < ns; ++s) dEs[s] = dE + 4 * s;
  group->capture(dP, dF, dEs.data());
}

void ApmoPlan::replay() {
  NAPMO_REQUIRE(group != nullptr, "APMO replay: switch on the one-launch mode first");
  group->replay();
}

void ApmoPlan::build_device(const double *const *dP, double *const *dF, double *dE) {
  const int ns = nspecies();
  if (group) {
    // the group's stream after the species' streams (inputs uploaded there), and the
    // species' streams after the group (outputs read there)
    for (int s = 0; s < ns; ++s) {
      NAPMO_HIP(hipEventRecord(sp[s].done, sp[s].plan->stream));
      NAPMO_HIP(hipStreamWaitEvent(group->stream, sp[s].done, 0));
    }
    std::vector<double *> dEs(ns);
    for (int s = 0; s < ns; ++s) dEs[s] = dE + 4 * s;
    group->build(dP, dF, dEs.data());
    NAPMO_HIP(hipEventRecord(sp[0].expanded, group->stream));
    for (int s = 0; s < ns; ++s) NAPMO_HIP(hipStreamWaitEvent(sp[s].plan->stream, sp[0].expanded, 0));
    return;
  }
  // densities, expansions and own potentials; each species on its own grid's stream
  for (int s = 0; s < ns; ++s) {
    FockPlan &p = *sp[s].plan;
    p.density_P(-1, dP[s]);
    p.expand(-1);
    p.synth_self(-1);
    NAPMO_HIP(hipEventRecord(sp[s].expanded, p.stream));
  }
  for (int s = 0; s < ns; ++s) {
    FockPlan &p = *sp[s].plan;
    DeviceGrid &gs = p.grid;
    hipStream_t st = p.stream;
    const double q = sp[s].charge;
    // v_cross = sum_{t != s} q_t V[rho_t] at s's points
    int first = 1;
    for (int t = 0; t < ns; ++t) {
      if (t == s) continue;
      FockPlan &pt = *sp[t].plan;
      NAPMO_HIP(hipStreamWaitEvent(st, sp[t].expanded, 0));
      if (sp[s].cross[t]) {
        pt.grid.synthesize_at(pt.lmax, pt.work, *sp[s].cross[t], gs.x.get(), gs.y.get(),
                              gs.z.get(), gs.N, sp[t].charge, first ? 0 : 1, sp[s].vcross.get(), st);
      } else {
        launch_axpby(sp[t].charge, pt.v.get(), first ? 0.0 : 1.0,
                     first ? pt.v.get() : sp[s].vcross.get(), gs.N, sp[s].vcross.get(), st);
      }
      first = 0;
    }
    if (first)
      launch_axpby(q * q, p.v.get(), 0.0, p.v.get(), gs.N, sp[s].veff.get(), st);
    else
      launch_axpby(q * q, p.v.get(), q, sp[s].vcross.get(), gs.N, sp[s].veff.get(), st);
    p.contract(-1, dP[s], sp[s].veff.get(), dF[s], dE + 4 * s);
    // own Coulomb energy 1/2 q^2 sum w bw rho V[rho_s]
    launch_traces(dP[s], dF[s], p.n, gs.wbw.get(), p.rho.get(), p.v.get(), gs.N, sp[s].E2.get(),
                  p.trace_work.get(), st);
    k_apmo_energy<<<1, 64, 0, st>>>(sp[s].E2.get(), q * q, dE + 4 * s);
    NAPMO_HIP(hipGetLastError());
    NAPMO_HIP(hipEventRecord(sp[s].done, st));
  }
  // every stream waits for every species, so a following build cannot overwrite a
  // potential or expansion another species is still reading
  for (int s = 0; s < ns; ++s)
    for (int t = 0; t < ns; ++t)
      if (t != s) NAPMO_HIP(hipStreamWaitEvent(sp[s].plan->stream, sp[t].done, 0));
}

void ApmoPlan::build_host(const double *const *P, double *const *F_out, double *E_out) {
  const int ns = nspecies();
  std::vector<DeviceBuffer<double>> dP(ns);
  std::vector<const double *> pP(ns);
  std::vector<double *> pF(ns);
  for (int s = 0; s < ns; ++s) {
    FockPlan &p = *sp[s].plan;
    for (size_t i = 0; i < (size_t)p.n * p.n; ++i)
      NAPMO_REQUIRE(std::isfinite(P[s][i]), "APMO build: P of species " + std::to_string(s) +
                                                " holds a non-finite value at row " + std::to_string(i / p.n) +
                                                ", column " + std::to_string(i % p.n));
    dP[s].upload(P[s], (size_t)p.n * p.n, p.stream);
    pP[s] = dP[s].get();
    pF[s] = p.F.get();
  }
  E.resize(4 * ns);
  build_device(pP.data(), pF.data(), E.get());
  for (int s = 0; s < ns; ++s) {
    FockPlan &p = *sp[s].plan;
    if (F_out && F_out[s]) p.F.download(F_out[s], (size_t)p.n * p.n, p.stream);
  }
  std::vector<double> Eh(4 * ns);
  E.download(Eh.data(), 4 * ns, sp[ns - 1].plan->stream);
  synchronize();
  for (double e : Eh)
    NAPMO_REQUIRE(std::isfinite(e), "APMO build: non-finite energy (a density or potential is not finite)");
  if (E_out) std::copy(Eh.begin(), Eh.end(), E_out);
}

void ApmoPlan::synchronize() {
  for (auto &S : sp) NAPMO_HIP(hipStreamSynchronize(S.plan->stream));
  if (group) group->synchronize();
}

}  // namespace napmo
