// Many independent molecules (one FockPlan each) built together, in one of two ways:
// nlanes >= 1: the plans' own launches spread over nlanes HIP streams (fork / join with
// events on one origin stream); nlanes = 0: one launch per phase for all molecules
// (GroupPlan, FP64 plans). Either way the whole batch can be captured once into a HIP
// graph and replayed per SCF iteration.
#pragma once

#include <functional>
#include <memory>
#include <vector>

#include "fock_plan.hpp"
#include "group_plan.hpp"

namespace napmo {

struct BatchPlan {
  std::vector<FockPlan *> plans;       // borrowed; must outlive the batch
  std::vector<hipStream_t> lanes;      // plan m runs on lanes[m % nlanes]
  hipStream_t origin = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<hipStream_t> saved;      // the plans' own streams, restored on destruction
  std::unique_ptr<GroupPlan> group;    // nlanes = 0

  BatchPlan(const std::vector<FockPlan *> &plans, int nlanes);
  ~BatchPlan();
  // One build of every plan: dP[m], dF[m] [n_m x n_m], dE[m] [2] (device pointers);
  // asynchronous on the origin stream.
  void build(const double *const *dP, double *const *dF, double *const *dE);
  // Orbital form (streams only): dC[m] [n_m x nocc[m]] with P_m = eta C C^T (the plans'
  // build_orbitals_device).
  void build_orbitals(const double *const *dC, const int *nocc, double eta, double *const *dF,
                      double *const *dE);
  // Records the same batch into a graph (pointers baked in); replay() launches it.
  void capture(const double *const *dP, double *const *dF, double *const *dE);
  void capture_orbitals(const double *const *dC, const int *nocc, double eta, double *const *dF,
                        double *const *dE);
  void replay();
  void synchronize();

 private:
  void record(const std::function<void()> &launch);  // capture of one batch build
};

}  // namespace napmo
