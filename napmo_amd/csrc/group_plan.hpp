// Many independent molecules built with one launch per phase (FP64 plans): the molecules'
// atoms and points side by side in one grid, so the Poisson steps (projection, radial
// Green solves, synthesis tables) are single launches and the synthesis over each
// molecule's own (point, centre) pairs is one launch per distinct lmax (the kernel is
// compiled per lmax); the per-molecule contractions (padding, density, centre sums, Fock
// partials and reduction, traces) are grouped launches whose workgroups find their
// molecule in a member table.
#pragma once

#include <vector>

#include "fock_plan.hpp"

namespace napmo {

struct GroupPlan {
  std::vector<FockPlan *> plans;  // borrowed; must outlive the group
  hipStream_t stream = nullptr;
  DeviceGrid ug;                  // every molecule's atoms and points, in plan order
  int lmax = 0;                   // the largest lmax; each molecule keeps its own (lcut)
  std::vector<long> p0;           // first point of each molecule in ug
  std::vector<int> A0;            // first atom of each molecule in ug
  DeviceBuffer<long> cbase;       // per atom of ug: its row of the per-centre sums
  DeviceBuffer<int> lcut;         // per atom of ug: its molecule's lmax
  DeviceBuffer<long> toff;        // per atom of ug: its synthesis-table rows (own lmax)
  SynthPairs pairs;               // the molecules' own pair buckets, side by side, the
                                  // molecules ordered by lmax
  struct SynthRange {             // the slots of the molecules of one lmax
    int lmax;
    long begin, count;
  };
  std::vector<SynthRange> ranges;
  DeviceBuffer<double> rho, v;    // [ug.N]
  int ks = 16, pf = 1;            // density stage shape (does not change results)
  // member tables (device) and workgroup counts of the grouped launches
  DeviceBuffer<DensityMember> dmem;
  DeviceBuffer<RhoReduceMember> rmem;
  DeviceBuffer<ScaleMember> smem;
  DeviceBuffer<FockMember> fmem;
  long dens_blocks = 0, rho_blocks = 0, scale_blocks = 0, fock_blocks = 0;
  // tables that name the build's inputs and outputs: one set for eager builds, one for
  // the captured graph (which reads its tables at every replay, so eager builds on other
  // pointers must not rewrite them)
  struct IoTables {
    DeviceBuffer<PadMember> pmem;
    DeviceBuffer<FockReduceMember> frmem;
    DeviceBuffer<TraceMember> tmem, t2mem;
    DeviceBuffer<ApmoEnergyMember> emem;
    long pad_blocks = 0, fred_blocks = 0;
    std::vector<const double *> P;  // the pointers the tables were made for
    std::vector<double *> F, E;
  };
  IoTables eager, graphed;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  // Coupled species (APMO, charges given): the members are the species of one system; the
  // pair lists also hold every other species' centres at a species' points, the per-centre
  // sums are dense rows over all points, and the centre-sum pass forms
  // veff_s = q_s^2 V[rho_s] + q_s sum_{t != s} q_t V[rho_t]; the traces add each species'
  // own Coulomb energy (E[2]) and the coupling (E[3]). Species may share a grid.
  bool coupled = false;
  std::vector<double> charge;
  DeviceBuffer<double> veff, E2;  // [ug.N], [2 x species]
  DeviceBuffer<SpeciesSum> dspecies;

  explicit GroupPlan(const std::vector<FockPlan *> &plans,
                     const std::vector<double> *charges = nullptr);
  ~GroupPlan();
  // One build of every plan: dP[m], dF[m] [n_m x n_m], dE[m] [2] (coupled: [4]) (device
  // pointers); asynchronous on the group's stream. Independent molecules: results equal the
  // plans' own builds bit for bit.
  void build(const double *const *dP, double *const *dF, double *const *dE);
  void capture(const double *const *dP, double *const *dF, double *const *dE);
  void replay();
  void synchronize();

 private:
  void io_tables(IoTables &t, const double *const *dP, double *const *dF, double *const *dE);
  void launches(const IoTables &t);
};

}  // namespace napmo
