// Many molecules, one launch per phase (see group_plan.hpp).
#include "group_plan.hpp"

#include <algorithm>
#include <map>
#include <memory>

namespace napmo {

GroupPlan::GroupPlan(const std::vector<FockPlan *> &ps, const std::vector<double> *charges)
    : plans(ps) {
  NAPMO_REQUIRE(!plans.empty(), "group plan: no plans");
  const int nm = (int)plans.size();
  coupled = charges != nullptr;
  if (coupled) {
    NAPMO_REQUIRE((int)charges->size() == nm, "group plan: one charge per species");
    charge = *charges;
  }
  std::vector<AtomInfo> all;
  std::vector<int> cut;
  p0.resize(nm + 1);
  A0.resize(nm + 1);
  long np = 0;
  for (int m = 0; m < nm; ++m) {
    FockPlan *q = plans[m];
    NAPMO_REQUIRE(q != nullptr, "group plan: null plan");
    NAPMO_REQUIRE(q->prec == NAPMO_PREC_FP64 && q->use_pass == 0,
                  "group plan: FP64 plans with the per-tile contraction only");
    NAPMO_REQUIRE(!q->grid.atoms.empty(), "group plan: every plan needs a grid of atoms");
    for (int k = 0; k < m && !coupled; ++k)
      NAPMO_REQUIRE(&plans[k]->grid != &q->grid, "group plan: every plan needs its own grid");
    NAPMO_HIP(hipStreamSynchronize(q->stream));
    p0[m] = np, A0[m] = (int)all.size();
    for (const auto &a : q->grid.atoms) all.push_back(a), cut.push_back(q->lmax);
    np += q->grid.N;
    lmax = std::max(lmax, q->lmax);
  }
  p0[nm] = np, A0[nm] = (int)all.size();
  // the union's geometry; Becke weights are each molecule's own
  ug.lcut = cut;
  ug.build_geometry(all);
  stream = ug.stream;
  ug.bw.resize(ug.N);
  ug.wbw.resize(ug.N);
  ug.becke.resize(ug.N);
  for (int m = 0; m < nm; ++m) {
    const DeviceGrid &g = plans[m]->grid;
    NAPMO_HIP(hipMemcpyAsync(ug.bw.get() + p0[m], g.bw.get(), sizeof(double) * g.N,
                             hipMemcpyDeviceToDevice, stream));
    NAPMO_HIP(hipMemcpyAsync(ug.wbw.get() + p0[m], g.wbw.get(), sizeof(double) * g.N,
                             hipMemcpyDeviceToDevice, stream));
    std::copy(g.becke.begin(), g.becke.end(), ug.becke.begin() + p0[m]);
  }
  // pair buckets: each molecule's own lists, its atoms and points shifted into the union,
  // the sources in order of lmax so each lmax's slots are one range
  std::vector<int> list, wb;
  std::vector<long> cb(ug.M), to(ug.M);
  long nwaves = 0, csize = 0, tsize = 0;
  for (int m = 0; m < nm; ++m) nwaves += plans[m]->grid.synth_pairs().nwaves;
  pairs.geo.resize((size_t)std::max<long>(nwaves, 1) * 256);
  std::vector<int> order(nm);
  for (int m = 0; m < nm; ++m) order[m] = m;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return plans[a]->lmax < plans[b]->lmax; });
  for (int m = 0; m < nm; ++m) {
    const DeviceGrid &g = plans[m]->grid;
    const long La = (long)(plans[m]->lmax + 1) * (plans[m]->lmax + 1);
    for (int a = 0; a < g.M; ++a) {
      cb[A0[m] + a] = csize + (long)a * g.N - p0[m];
      to[A0[m] + a] = tsize;
      tsize += (long)std::max(g.atoms[a].nrad - 1, 1) * La * 4;
    }
    csize += (long)g.M * g.N;
  }
  // (target molecule or species s, source t) lists: independent molecules have only s = t;
  // coupled species every pair, t's own lists when s and t share a grid
  struct Src {
    int s, t;
    SynthPairs *sp;
  };
  std::vector<Src> srcs;
  std::vector<std::unique_ptr<SynthPairs>> cross;
  for (int t : order)
    for (int s = 0; s < nm; ++s) {
      if (!coupled && s != t) continue;
      DeviceGrid &gt = plans[t]->grid, &gs = plans[s]->grid;
      if (s == t || &gs == &gt) {
        srcs.push_back({s, t, &gt.synth_pairs()});
      } else {
        cross.push_back(build_synth_pairs(gt, gs.x.get(), gs.y.get(), gs.z.get(), gs.N, gs.stream));
        srcs.push_back({s, t, cross.back().get()});
      }
    }
  if (coupled) {
    nwaves = 0;
    for (auto &r : srcs) nwaves += r.sp->nwaves;
    pairs.geo.resize((size_t)std::max<long>(nwaves, 1) * 256);
    // dense per-centre rows over every point
    for (int A = 0; A < ug.M; ++A) cb[A] = (long)A * ug.N;
    csize = (long)ug.M * ug.N;
  }
  long w0 = 0;
  for (const auto &r : srcs) {
    const int m = r.s;
    DeviceGrid &g = plans[r.t]->grid;
    SynthPairs &sp = *r.sp;
    std::vector<int> l((size_t)sp.nwaves * 64), b(sp.nwaves);
    sp.list.download(l.data(), l.size(), g.stream);
    sp.wave_bucket.download(b.data(), b.size(), g.stream);
    NAPMO_HIP(hipStreamSynchronize(g.stream));
    for (int &e : l)
      if (e >= 0) {
        NAPMO_REQUIRE(e + p0[m] < (1L << 31), "group plan: too many points for the pair lists");
        e += (int)p0[m];
      }
    for (int &e : b) {
      const int A = e / (g.max_nrad + 1), loc = e - A * (g.max_nrad + 1);
      e = (A + A0[r.t]) * (ug.max_nrad + 1) + loc;
    }
    list.insert(list.end(), l.begin(), l.end());
    wb.insert(wb.end(), b.begin(), b.end());
    NAPMO_HIP(hipMemcpyAsync(pairs.geo.get() + w0 * 256, sp.geo.get(), sizeof(double) * sp.nwaves * 256,
                             hipMemcpyDeviceToDevice, stream));
    const int lt = plans[r.t]->lmax;
    if (ranges.empty() || ranges.back().lmax != lt) ranges.push_back({lt, w0, 0});
    ranges.back().count += sp.nwaves;
    w0 += sp.nwaves;
  }
  NAPMO_HIP(hipStreamSynchronize(stream));  // the geometry copies read the cross lists
  pairs.nwaves = nwaves;
  pairs.list.upload(list, stream);
  pairs.wave_bucket.upload(wb, stream);
  pairs.contrib.resize((size_t)std::max<long>(csize, 1));
  cbase.upload(cb, stream);
  lcut.upload(cut, stream);
  toff.upload(to, stream);
  rho.resize(ug.N);
  v.resize(ug.N);
  PoissonPlan &pl = ug.plan(lmax);
  NAPMO_REQUIRE(pl.work.T.size() >= (size_t)tsize, "group plan: synthesis table size");
  // member tables that do not depend on the build's pointers
  // the density stage shape most of the plans were tuned to (it does not change results)
  std::map<std::pair<int, int>, int> votes;
  for (auto *q : plans) ++votes[{q->density_ks, q->density_pf}];
  int best = -1;
  for (const auto &kv : votes)
    if (kv.second > best) best = kv.second, ks = kv.first.first, pf = kv.first.second;
  std::vector<DensityMember> dm(nm);
  std::vector<RhoReduceMember> rm(nm);
  std::vector<ScaleMember> sm(nm);
  std::vector<FockMember> fm(nm);
  for (int m = 0; m < nm; ++m) {
    FockPlan &q = *plans[m];
    const long N = q.grid.N;
    dens_blocks += density_group_fill(dm[m], q.phi64.get(), q.ld, q.Npad, q.P.get(), q.rho_part.get(),
                                      q.n, ks, dens_blocks);
    rm[m] = RhoReduceMember{q.rho_part.get(), rho.get() + p0[m], q.Npad, N, rho_blocks, q.ld / TILE};
    rho_blocks += div_up(N, 256);
    sm[m] = ScaleMember{q.s.get(), p0[m], N, q.Npad, scale_blocks, A0[m], A0[m + 1]};
    scale_blocks += div_up(q.Npad, 256);
    fock_blocks += fock_group_fill(fm[m], q.fsh, q.phi64.get(), q.s.get(), q.F_part.get(), fock_blocks);
  }
  dmem.upload(dm, stream);
  rmem.upload(rm, stream);
  smem.upload(sm, stream);
  fmem.upload(fm, stream);
  if (coupled) {
    std::vector<SpeciesSum> ss(nm);
    for (int m = 0; m < nm; ++m) ss[m] = SpeciesSum{A0[m], A0[m + 1], charge[m]};
    dspecies.upload(ss, stream);
    veff.resize(ug.N);
    E2.resize(2 * nm);
  }
  NAPMO_HIP(hipStreamSynchronize(stream));
}

GroupPlan::~GroupPlan() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
}

void GroupPlan::io_tables(IoTables &io, const double *const *dP, double *const *dF,
                          double *const *dE) {
  const int nm = (int)plans.size();
  bool same = (int)io.P.size() == nm;
  for (int m = 0; same && m < nm; ++m)
    same = io.P[m] == dP[m] && io.F[m] == dF[m] && io.E[m] == dE[m];
  if (same) return;
  std::vector<PadMember> pm(nm);
  std::vector<FockReduceMember> frm(nm);
  std::vector<TraceMember> tm(nm), t2m(nm);
  std::vector<ApmoEnergyMember> em(nm);
  long pad_blocks = 0, fred_blocks = 0;
  for (int m = 0; m < nm; ++m) {
    FockPlan &q = *plans[m];
    NAPMO_REQUIRE(dP[m] && dF[m] && dE[m], "group build: null pointer");
    pm[m] = PadMember{dP[m], q.P.get(), pad_blocks, q.n, q.n, q.ld, q.ld};
    pad_blocks += div_up((long)q.ld * q.ld, 256);
    fred_blocks += fock_reduce_group_fill(frm[m], q.fsh, q.F_part.get(), q.n, dF[m], fred_blocks);
    const double *vs = coupled ? veff.get() : v.get();
    tm[m] = TraceMember{dP[m], dF[m], ug.wbw.get() + p0[m], rho.get() + p0[m], vs + p0[m],
                        dE[m], q.trace_work.get(), (long)q.n * q.n, q.grid.N};
    if (coupled) {
      // the species' own Coulomb energy 1/2 sum w bw rho V[rho_s] (its work buffer is free
      // again once the first traces are final: same stream)
      t2m[m] = TraceMember{nullptr, nullptr, ug.wbw.get() + p0[m], rho.get() + p0[m], v.get() + p0[m],
                           E2.get() + 2 * m, q.trace_work.get(), 0, q.grid.N};
      em[m] = ApmoEnergyMember{dE[m], E2.get() + 2 * m, charge[m] * charge[m]};
    }
  }
  // the previous build may still read the old tables: replace them in stream order, and
  // keep the host copies alive until the copies are done
  io.pmem.upload(pm, stream);
  io.frmem.upload(frm, stream);
  io.tmem.upload(tm, stream);
  if (coupled) {
    io.t2mem.upload(t2m, stream);
    io.emem.upload(em, stream);
  }
  NAPMO_HIP(hipStreamSynchronize(stream));
  io.pad_blocks = pad_blocks, io.fred_blocks = fred_blocks;
  io.P.assign(dP, dP + nm);
  io.F.assign(dF, dF + nm);
  io.E.assign(dE, dE + nm);
}

void GroupPlan::launches(const IoTables &io) {
  const int nm = (int)plans.size();
  launch_pad_group(io.pmem.get(), nm, io.pad_blocks, stream);
  launch_density_group(dmem.get(), nm, dens_blocks, ks, pf, stream);
  launch_rho_reduce_group(rmem.get(), nm, rho_blocks, stream);
  PoissonPlan &pl = ug.plan(lmax);
  AtomTables at = ug.tables(pl);
  at.cbase = cbase.get();
  at.lcut = lcut.get();
  at.toff = toff.get();
  at.amp_stride = (lmax + 1) * (lmax + 1);
  PoissonWork &w = pl.work;
  launch_project(at, rho.get(), ug.bw.get(), pl.yhat.get(), lmax, w.rho_lm.get(), stream);
  launch_radial(at, w.rho_lm.get(), pl.green.get(), lmax, w.Uy.get(), w.Udt.get(), stream);
  launch_synth_tables(at, w.Uy.get(), w.Udt.get(), lmax, w.T.get(), w.ampL.get(), w.ampR.get(), stream);
  for (const auto &r : ranges)
    launch_synth2(at, pairs.geo.get() + r.begin * 256, ug.N, pairs.list.get() + r.begin * 64,
                  pairs.wave_bucket.get() + r.begin, r.count, w.T.get(), w.ampL.get(), w.ampR.get(),
                  r.lmax, pairs.contrib.get(), stream, 2);
  if (coupled)
    launch_apmo_reduce_group(smem.get(), nm, scale_blocks, dspecies.get(), pairs.contrib.get(), ug.N,
                             ug.wbw.get(), v.get(), veff.get(), stream);
  else
    launch_contrib_reduce_scale_group(smem.get(), nm, scale_blocks, pairs.contrib.get(), cbase.get(),
                                      ug.wbw.get(), v.get(), stream);
  launch_fock_group(fmem.get(), nm, fock_blocks, stream);
  launch_fock_reduce_group(io.frmem.get(), nm, io.fred_blocks, stream);
  launch_traces_group(io.tmem.get(), nm, stream);
  if (coupled) {
    launch_traces_group(io.t2mem.get(), nm, stream);
    launch_apmo_energy_group(io.emem.get(), nm, stream);
  }
}

void GroupPlan::build(const double *const *dP, double *const *dF, double *const *dE) {
  io_tables(eager, dP, dF, dE);
  launches(eager);
}

void GroupPlan::capture(const double *const *dP, double *const *dF, double *const *dE) {
  synchronize();
  graphed.P.clear();  // a new capture: fresh tables even for the same pointers
  io_tables(graphed, dP, dF, dE);
  if (exec) (void)hipGraphExecDestroy(exec), exec = nullptr;
  if (graph) (void)hipGraphDestroy(graph), graph = nullptr;
  NAPMO_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
  try {
    launches(graphed);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  NAPMO_HIP(hipStreamEndCapture(stream, &graph));
  NAPMO_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
}

void GroupPlan::replay() {
  NAPMO_REQUIRE(exec != nullptr, "group replay: capture first");
  NAPMO_HIP(hipGraphLaunch(exec, stream));
}

void GroupPlan::synchronize() { NAPMO_HIP(hipStreamSynchronize(stream)); }

}  // namespace napmo
