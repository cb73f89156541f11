// Batched builds of independent molecules (see batch_plan.hpp).
#include "batch_plan.hpp"

#include <functional>

namespace napmo {

BatchPlan::BatchPlan(const std::vector<FockPlan *> &p, int nlanes) : plans(p) {
  NAPMO_REQUIRE(!plans.empty(), "batch plan: no plans");
  NAPMO_REQUIRE(nlanes >= 0 && nlanes <= 16, "batch plan: 0 (one launch per phase) to 16 lanes");
  if (nlanes == 0) {
    group = std::make_unique<GroupPlan>(plans);
    return;
  }
  for (size_t m = 0; m < plans.size(); ++m) {
    FockPlan *q = plans[m];
    NAPMO_REQUIRE(q != nullptr, "batch plan: null plan");
    // plans on different lanes would race on a shared grid's synthesis buffers
    for (size_t k = 0; k < m; ++k)
      NAPMO_REQUIRE(&plans[k]->grid != &q->grid, "batch plan: every plan needs its own grid");
    // the plan's pending set-up work must be finished before its stream changes
    NAPMO_HIP(hipStreamSynchronize(q->stream));
  }
  NAPMO_HIP(hipStreamCreateWithFlags(&origin, hipStreamNonBlocking));
  lanes.resize(nlanes);
  joins.resize(nlanes);
  for (int l = 0; l < nlanes; ++l) {
    NAPMO_HIP(hipStreamCreateWithFlags(&lanes[l], hipStreamNonBlocking));
    NAPMO_HIP(hipEventCreateWithFlags(&joins[l], hipEventDisableTiming));
  }
  NAPMO_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  saved.resize(plans.size());
  for (size_t m = 0; m < plans.size(); ++m) {
    saved[m] = plans[m]->stream;
    plans[m]->stream = lanes[m % nlanes];
  }
}

BatchPlan::~BatchPlan() {
  if (group) return;
  (void)hipStreamSynchronize(origin);
  for (auto s : lanes) (void)hipStreamSynchronize(s);
  for (size_t m = 0; m < plans.size(); ++m) plans[m]->stream = saved[m];
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (auto e : joins) (void)hipEventDestroy(e);
  if (fork) (void)hipEventDestroy(fork);
  for (auto s : lanes) (void)hipStreamDestroy(s);
  if (origin) (void)hipStreamDestroy(origin);
}

void BatchPlan::build(const double *const *dP, double *const *dF, double *const *dE) {
  if (group) return group->build(dP, dF, dE);
  NAPMO_HIP(hipEventRecord(fork, origin));
  for (auto s : lanes) NAPMO_HIP(hipStreamWaitEvent(s, fork, 0));
  for (size_t m = 0; m < plans.size(); ++m) plans[m]->build_device(dP[m], dF[m], dE[m]);
  for (size_t l = 0; l < lanes.size(); ++l) {
    NAPMO_HIP(hipEventRecord(joins[l], lanes[l]));
    NAPMO_HIP(hipStreamWaitEvent(origin, joins[l], 0));
  }
}

void BatchPlan::build_orbitals(const double *const *dC, const int *nocc, double eta,
                               double *const *dF, double *const *dE) {
  NAPMO_REQUIRE(!group, "orbital-form batch: streams only (nlanes >= 1)");
  NAPMO_HIP(hipEventRecord(fork, origin));
  for (auto s : lanes) NAPMO_HIP(hipStreamWaitEvent(s, fork, 0));
  for (size_t m = 0; m < plans.size(); ++m)
    plans[m]->build_orbitals_device(dC[m], nocc[m], eta, dF[m], dE[m]);
  for (size_t l = 0; l < lanes.size(); ++l) {
    NAPMO_HIP(hipEventRecord(joins[l], lanes[l]));
    NAPMO_HIP(hipStreamWaitEvent(origin, joins[l], 0));
  }
}

void BatchPlan::capture_orbitals(const double *const *dC, const int *nocc, double eta,
                                 double *const *dF, double *const *dE) {
  NAPMO_REQUIRE(!group, "orbital-form batch: streams only (nlanes >= 1)");
  record([&] { build_orbitals(dC, nocc, eta, dF, dE); });
}

void BatchPlan::capture(const double *const *dP, double *const *dF, double *const *dE) {
  if (group) return group->capture(dP, dF, dE);
  record([&] { build(dP, dF, dE); });
}

void BatchPlan::record(const std::function<void()> &launch) {
  for (auto *q : plans)
    NAPMO_REQUIRE(q->prof_used >= q->prof_cap, "batch capture: stop per-phase profiling first");
  synchronize();
  if (exec) (void)hipGraphExecDestroy(exec), exec = nullptr;
  if (graph) (void)hipGraphDestroy(graph), graph = nullptr;
  NAPMO_HIP(hipStreamBeginCapture(origin, hipStreamCaptureModeRelaxed));
  try {
    launch();
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(origin, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  NAPMO_HIP(hipStreamEndCapture(origin, &graph));
  NAPMO_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
}

void BatchPlan::replay() {
  if (group) return group->replay();
  NAPMO_REQUIRE(exec != nullptr, "batch replay: capture first");
  NAPMO_HIP(hipGraphLaunch(exec, origin));
}

void BatchPlan::synchronize() {
  if (group) return group->synchronize();
  NAPMO_HIP(hipStreamSynchronize(origin));
}

}  // namespace napmo
