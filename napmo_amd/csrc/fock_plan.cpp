// Device-resident grid Fock build (see fock_plan.hpp).
#include "fock_plan.hpp"

#include <numeric>

#include <algorithm>
#include <cmath>
#include <vector>

namespace napmo {

void DeviceBasis::upload(const FlatBasis &fb, hipStream_t s) {
  nbasis = fb.nbasis;
  origin.upload(fb.origin, s);
  norm.upload(fb.norm, s);
  zeta.upload(fb.zeta, s);
  coef.upload(fb.coef, s);
  lxyz.upload(fb.lxyz, s);
  first.upload(fb.first, s);
  nprim.upload(fb.nprim, s);
}

void DeviceBasis::collocate(const double *x, const double *y, const double *z, long N, long Npad,
                            int ld, double *out64, float *out32, hipStream_t s) const {
  launch_collocate(x, y, z, N, Npad, nbasis, ld, origin.get(), lxyz.get(), first.get(),
                   nprim.get(), norm.get(), zeta.get(), coef.get(), out64, out32, s);
}

FockPlan::FockPlan(DeviceGrid &g, const BasisSet &b, int lmax_, int prec_mode)
    : grid(g), lmax(lmax_), prec(prec_mode) {
  NAPMO_REQUIRE(prec == NAPMO_PREC_FP64 || prec == NAPMO_PREC_BF16X3 || prec == NAPMO_PREC_INT8,
                "precision mode not available: use NAPMO_PREC_FP64 (0), NAPMO_PREC_BF16X3 (2) or "
                "NAPMO_PREC_INT8 (3)");
  FlatBasis fb = flatten(b);
  n = fb.nbasis;
  NAPMO_REQUIRE(n > 0, "empty basis");
  std::vector<int> cof(n);
  for (int mu = 0; mu < n; ++mu) {
    const double *o = &fb.origin[3 * mu];
    int c = 0;
    const int nc = (int)centres.size() / 3;
    while (c < nc && !(centres[3 * c] == o[0] && centres[3 * c + 1] == o[1] && centres[3 * c + 2] == o[2])) ++c;
    if (c == nc) centres.insert(centres.end(), o, o + 3);
    cof[mu] = c;
  }
  centre_of.upload(cof, grid.stream);
  ld = (int)round_up(n, TILE);
  Npad = round_up(grid.N, 256);  // multiple of every point tile (64, 256)
  hipStream_t st = grid.stream;
  stream = st;
  basis.upload(fb, st);
  phi64.resize((size_t)Npad * ld);
  basis.collocate(grid.x.get(), grid.y.get(), grid.z.get(), grid.N, Npad, ld, phi64.get(), nullptr,
                  st);
  if (prec == NAPMO_PREC_BF16X3) {
    phiT_hi.resize((size_t)ld * Npad);
    phiT_lo.resize((size_t)ld * Npad);
    phiT32.resize((size_t)ld * Npad);
    launch_split_phi(phi64.get(), ld, Npad, phiT_hi.get(), phiT_lo.get(), phiT32.get(), st);
  }
  if (prec == NAPMO_PREC_INT8) ozaki_setup();
  grid.plan(lmax);     // Poisson tables (projection, Green matrices) built once per grid
  grid.alloc_work(lmax, work);
  grid.synth_pairs();  // synthesis pair buckets built once per geometry
  P.resize((size_t)ld * ld);
  NAPMO_HIP(hipMemsetAsync(P.get(), 0, sizeof(double) * ld * ld, st));
  rho_part.resize((size_t)(ld / TILE) * Npad);
  rho.resize(grid.N);
  v.resize(grid.N);
  s.resize(Npad);
  use_pass = (prec == NAPMO_PREC_BF16X3 && fock_pass_supported(ld)) ? 1 : 0;
  reshape();
  F.resize((size_t)n * n);
  E.resize(2);
  trace_work.resize(2 * TRACE_WORK);
  NAPMO_HIP(hipStreamSynchronize(st));
}

void FockPlan::reshape() {
  fsh = fock_shape(ld, Npad, true, fock_wg);
  psh = fock_pass_shape(ld, Npad, pass_wg);
  F_part.resize((size_t)std::max(fsh.nchunk, use_pass ? psh.nchunk : 0) * fsh.cstride());
}

void FockPlan::set_tuning(int key, long value) {
  NAPMO_REQUIRE(value >= 0, "tuning value must be non-negative");
  switch (key) {
    case NAPMO_TUNE_FOCK_WG: fock_wg = value; break;
    case NAPMO_TUNE_PASS_WG: pass_wg = value; break;
    case NAPMO_TUNE_USE_PASS:
      NAPMO_REQUIRE(value >= 0 && value <= 2, "one-pass kernel choice must be 0, 1 or 2");
      NAPMO_REQUIRE(value == 0 || (prec == NAPMO_PREC_BF16X3 &&
                                   (value == 1 ? fock_pass_supported(ld) : fock_syrk_supported(ld))),
                    "one-pass kernel not available for this plan");
      use_pass = (int)value;
      break;
    case NAPMO_TUNE_DENSITY_PF:
      NAPMO_REQUIRE(value == 0 || value == 1 || value == 2, "density prefetch depth must be 1 or 2");
      density_pf = value ? (int)value : 1;
      break;
    case NAPMO_TUNE_DENSITY_KS:
      NAPMO_REQUIRE(value == 0 || value == 16 || (value == 32 && ld % 32 == 0),
                    "density stage depth must be 16 or 32");
      density_ks = value ? (int)value : 16;
      break;
    case NAPMO_TUNE_SYNTH_NP:
      NAPMO_REQUIRE(value == 0 || value == 1 || value == 2, "synthesis pairs per lane must be 1 or 2");
      grid.synth_np = value ? (int)value : 2;
      break;
    case NAPMO_TUNE_SYNTH_MFMA_MIN:
      NAPMO_REQUIRE(value >= 0 && value <= 17, "MFMA synthesis threshold must be 0 to 17");
      grid.synth_mfma_min = value ? (int)value : SYNTH_MFMA_MIN;
      break;
    case NAPMO_TUNE_BACKWARD_POINTS: bwd_points = value; break;
    case NAPMO_TUNE_DENSITY_STREAM:
      NAPMO_REQUIRE(value >= 0 && value <= 3, "density streaming choice must be 0 to 3");
      density_stream = (int)value;
      break;
    case NAPMO_TUNE_DENSITY_SCREEN:
      NAPMO_REQUIRE(value >= 0 && value <= 1000, "density screening threshold must be 0 (off) to 1000");
      NAPMO_REQUIRE(value == 0 || ld <= 1024, "density screening needs at most 1024 functions");
      NAPMO_HIP(hipStreamSynchronize(stream));
      density_screen = (int)value;
      density_screen_setup();
      break;
    case NAPMO_TUNE_INT8_PACK:
      NAPMO_REQUIRE(prec == NAPMO_PREC_INT8, "int8 tuning keys need an NAPMO_PREC_INT8 plan");
      NAPMO_REQUIRE(value == 0 || value == 1, "int8 packing must be 0 or 1");
      NAPMO_REQUIRE(value == 0 || (oz.pack_rows >= 1 && oz.nb <= OZ_MAXROWS),
                    "int8 packing: the basis does not fit a packed task");
      oz_pack = (int)value;
      break;
    case NAPMO_TUNE_INT8_CHUNK:
    case NAPMO_TUNE_INT8_WAVE_BLOCKS:
    case NAPMO_TUNE_INT8_SCREEN:
      NAPMO_REQUIRE(prec == NAPMO_PREC_INT8, "int8 tuning keys need an NAPMO_PREC_INT8 plan");
      NAPMO_REQUIRE(key != NAPMO_TUNE_INT8_SCREEN || (value >= 0 && value <= 1000),
                    "int8 screening threshold must be 0 (off) to 1000");
      NAPMO_REQUIRE(key != NAPMO_TUNE_INT8_CHUNK || value == 0 ||
                        (value % OZ_STAGE == 0 && value <= (long)OZ_MAX_STAGES * OZ_STAGE),
                    "int8 chunk must be a multiple of 64 points, at most 4032");
      NAPMO_REQUIRE(key != NAPMO_TUNE_INT8_WAVE_BLOCKS || (value >= 0 && value <= OZ_MAXB),
                    "int8 blocks per wave must be 1 to " + std::to_string(OZ_MAXB) + " (0: the most)");
      NAPMO_HIP(hipStreamSynchronize(stream));
      if (key == NAPMO_TUNE_INT8_CHUNK)
        oz_chunk = value;
      else if (key == NAPMO_TUNE_INT8_WAVE_BLOCKS)
        oz_wave_blocks = value ? (int)value : OZ_MAXB;
      else
        oz_screen = (int)value;
      ozaki_setup();
      break;
    default: throw Error("unknown tuning key " + std::to_string(key));
  }
  NAPMO_HIP(hipStreamSynchronize(stream));
  reshape();
}

long FockPlan::get_tuning(int key) const {
  switch (key) {
    case NAPMO_TUNE_FOCK_WG: return (long)fsh.nchunk * fsh.ntiles;
    case NAPMO_TUNE_PASS_WG: return (long)psh.nchunk;
    case NAPMO_TUNE_USE_PASS: return use_pass;
    case NAPMO_TUNE_DENSITY_KS: return density_ks;
    case NAPMO_TUNE_DENSITY_PF: return density_pf;
    case NAPMO_TUNE_SYNTH_NP: return grid.synth_np;
    case NAPMO_TUNE_SYNTH_MFMA_MIN: return grid.synth_mfma_min;
    case NAPMO_TUNE_BACKWARD_POINTS: return bwd_points ? backward_chunk() : 0;
    case NAPMO_TUNE_DENSITY_STREAM: return density_stream;
    case NAPMO_TUNE_DENSITY_SCREEN: return density_screen;
    case NAPMO_TUNE_INT8_CHUNK: return prec == NAPMO_PREC_INT8 ? oz.chunk : 0;
    case NAPMO_TUNE_INT8_WAVE_BLOCKS: return prec == NAPMO_PREC_INT8 ? oz_wave_blocks : 0;
    case NAPMO_TUNE_INT8_SCREEN: return prec == NAPMO_PREC_INT8 ? oz_screen : 0;
    case NAPMO_TUNE_INT8_PACK: return prec == NAPMO_PREC_INT8 ? oz_pack : 0;
    default: throw Error("unknown tuning key " + std::to_string(key));
  }
}

FockPlan::~FockPlan() {
  for (auto e : prof_events) (void)hipEventDestroy(e);
}

const char *FockPlan::phase_name(int k) {
  static const char *names[NPHASE] = {"density",      "rho_reduce", "project", "radial",
                                      "synth_tables", "synth",      "scale",   "fock",
                                      "fock_reduce",  "traces"};
  return (k >= 0 && k < NPHASE) ? names[k] : "";
}

void FockPlan::mark(int slot, int phase) {
  if (slot < 0) return;
  NAPMO_HIP(hipEventRecord(prof_events[(size_t)slot * (NPHASE + 1) + phase], stream));
}

void FockPlan::profile_start(int max_builds) {
  for (auto e : prof_events) (void)hipEventDestroy(e);
  prof_events.clear();
  prof_cap = std::max(0, max_builds);
  prof_used = 0;
  prof_events.resize((size_t)prof_cap * (NPHASE + 1));
  for (auto &e : prof_events) NAPMO_HIP(hipEventCreate(&e));
}

int FockPlan::profile_read(double *avg_ms, int nphase) {
  NAPMO_HIP(hipStreamSynchronize(stream));
  const int k = std::min(nphase, NPHASE);
  for (int p = 0; p < k; ++p) avg_ms[p] = 0.0;
  for (int b = 0; b < prof_used; ++b)
    for (int p = 0; p < k; ++p) {
      float ms = 0.f;
      NAPMO_HIP(hipEventElapsedTime(&ms, prof_events[(size_t)b * (NPHASE + 1) + p],
                                    prof_events[(size_t)b * (NPHASE + 1) + p + 1]));
      avg_ms[p] += ms;
    }
  if (prof_used)
    for (int p = 0; p < k; ++p) avg_ms[p] /= prof_used;
  return prof_used;
}

void FockPlan::density_P(int slot, const double *dP) {
  hipStream_t st = stream;
  // P (n x n) into the zero-padded ld x ld operand (a kernel, so the build can be captured)
  launch_pad2d(dP, n, n, P.get(), ld, ld, st);
  mark(slot, 0);
  launch_density_f64(phi64.get(), ld, Npad, P.get(), rho_part.get(), st, density_ks, density_pf, n);
  mark(slot, 1);
  launch_rho_reduce(rho_part.get(), ld / TILE, Npad, grid.N, rho.get(), st);
}

void FockPlan::density_screen_setup() {
  hipStream_t st = stream;
  if (density_screen <= 0) {
    phi_scr.resize(0), dens_perm.resize(0), dens_eT.resize(0);
    return;
  }
  // per (16-row tile, function) exponents of max |Phi|; the compact functions first (by the
  // number of tiles where they reach the largest exponent less 30), then the block exponents
  // of Phi in that order
  const long nt = Npad / 16;
  const int ng = ld / 16;
  DeviceBuffer<signed char> e16((size_t)nt * ld);
  launch_phi_tile_exp(phi64.get(), ld, Npad, e16.get(), st);
  std::vector<signed char> e((size_t)nt * ld);
  e16.download(e.data(), e.size(), st);
  NAPMO_HIP(hipStreamSynchronize(st));
  int top = -128;
  for (long t = 0; t < nt; ++t)
    for (int f = 0; f < n; ++f) top = std::max(top, (int)e[(size_t)t * ld + f]);
  std::vector<long> reach(n, 0);
  for (long t = 0; t < nt; ++t)
    for (int f = 0; f < n; ++f) reach[f] += e[(size_t)t * ld + f] != -128 && e[(size_t)t * ld + f] >= top - 30;
  std::vector<int> order(ld);
  for (int f = 0; f < ld; ++f) order[f] = f;
  std::stable_sort(order.begin(), order.begin() + n, [&](int x, int y) { return reach[x] < reach[y]; });
  std::vector<signed char> eT((size_t)nt * ng);
  for (long t = 0; t < nt; ++t)
    for (int g = 0; g < ng; ++g) {
      int m = -128;
      for (int i = 0; i < 16; ++i) m = std::max(m, (int)e[(size_t)t * ld + order[16 * g + i]]);
      eT[(size_t)t * ng + g] = (signed char)m;
    }
  dens_perm.upload(order, st);
  dens_eT.upload(eT, st);
  phi_scr.resize((size_t)Npad * ld);
  launch_permute_cols(phi64.get(), ld, Npad, dens_perm.get(), phi_scr.get(), st);
  NAPMO_HIP(hipStreamSynchronize(st));
}

void FockPlan::density_C(int slot, const double *dC, int nocc, double eta) {
  NAPMO_REQUIRE(nocc > 0 && nocc <= n, "orbital count out of range");
  hipStream_t st = stream;
  const int ldc = density_orb_cols(nocc), nct = density_orb_tiles(ldc);
  Pfull.resize((size_t)n * n);
  if (nct == 1 && density_stream && density_orb_stream_supported(ld, ldc, n)) {
    // C resident in LDS (read as given, no padded copy), Phi streamed; the epilogue writes
    // rho = eta sum_i X_pi^2 itself. P = eta C C^T feeds the energy trace only.
    launch_gram(dC, n, nocc, eta, Pfull.get(), st);
    mark(slot, 0);
    DensityScreen sc;
    if (density_screen > 0) sc.perm = dens_perm.get(), sc.eT = dens_eT.get(), sc.S = density_screen;
    launch_density_orb_stream(density_screen > 0 ? phi_scr.get() : phi64.get(), ld, Npad, dC, n, nocc,
                              rho.get(), st, eta, grid.N, density_stream, sc);
    mark(slot, 1);
    return;
  }
  Cpad.resize((size_t)ld * ldc);
  if (rho_part.size() < (size_t)nct * Npad) rho_part.resize((size_t)nct * Npad);
  launch_pad2d(dC, n, nocc, Cpad.get(), ld, ldc, st);
  launch_gram(dC, n, nocc, eta, Pfull.get(), st);
  mark(slot, 0);
  if (nct == 1) {
    // one column tile: the kernel's epilogue writes rho = eta sum_i X_pi^2 itself
    launch_density_orb_f64(phi64.get(), ld, Npad, Cpad.get(), ldc, rho.get(), st, density_ks,
                           density_pf, n, eta, grid.N);
    mark(slot, 1);
  } else {
    launch_density_orb_f64(phi64.get(), ld, Npad, Cpad.get(), ldc, rho_part.get(), st, density_ks,
                           density_pf, n);
    mark(slot, 1);
    launch_rho_reduce(rho_part.get(), nct, Npad, grid.N, rho.get(), st, eta);
  }
}

void FockPlan::expand(int slot) {
  hipStream_t st = stream;
  PoissonPlan &pl = grid.plan(lmax);
  AtomTables at = grid.tables(pl);
  mark(slot, 2);
  launch_project(at, rho.get(), grid.bw.get(), pl.yhat.get(), lmax, work.rho_lm.get(), st);
  mark(slot, 3);
  launch_radial(at, work.rho_lm.get(), pl.green.get(), lmax, work.Uy.get(), work.Udt.get(), st);
  mark(slot, 4);
  launch_synth_tables(at, work.Uy.get(), work.Udt.get(), lmax, work.T.get(), work.ampL.get(),
                      work.ampR.get(), st);
}

void FockPlan::synth_self(int slot, bool with_scale) {
  hipStream_t st = stream;
  PoissonPlan &pl = grid.plan(lmax);
  AtomTables at = grid.tables(pl);
  SynthPairs &sp = grid.synth_pairs();
  mark(slot, 5);
  launch_synth2(at, sp.geo.get(), grid.N, sp.list.get(),
                sp.wave_bucket.get(), sp.nwaves, work.T.get(), work.ampL.get(), work.ampR.get(), lmax,
                sp.contrib.get(), st, grid.synth_np);
  mark(slot, 6);  // the "scale" phase: centre sum, then w bw v
  if (with_scale)
    launch_contrib_reduce_scale(sp.contrib.get(), grid.M, grid.N, Npad, grid.wbw.get(), v.get(),
                                s.get(), st);
  else
    launch_contrib_reduce(sp.contrib.get(), grid.M, grid.N, 0, v.get(), st);
}

void FockPlan::ozaki_setup() {
  hipStream_t st = stream;
  oz = ozaki_layout(grid.N, n, ld, oz_wave_blocks, max_lds_per_block(), oz_chunk);
  oz_planes.resize((size_t)OZ_DIGITS * ld * oz.Nq);
  oz_eA.resize((size_t)ld * oz.nchunk);
  oz_maxd1.resize((size_t)ld * (oz.Nq / OZ_STAGE));
  oz_vs.resize(oz.Nq);
  oz_part.resize((size_t)oz.nchunk * oz.partial_stride());
  oz_ptasks.resize((size_t)oz.nchunk * oz.nb * OZ_PTASK);
  oz_pctr.resize(2);
  oz_pcache.resize((size_t)oz.nchunk * oz.nb * OZ_PTASK);
  oz_pccount.resize(oz.nchunk);
  const std::vector<int> no_cache(oz.nchunk, OZ_NO_CACHE);  // kept until the synchronisation below
  oz_pcev.upload(no_cache, st);
  std::vector<int> tab = oz.tasks;
  tab.insert(tab.end(), oz.row0.begin(), oz.row0.end());
  oz_tasks.upload(tab, st);
  oz_perm.resize(0);
  std::vector<int> order;  // kept until the stream is synchronised below
  if (oz_screen > 0) {     // the compact functions first, from their chunk exponents
    launch_ozaki_scales(oz, phi64.get(), grid.wbw.get(), nullptr, oz_eA.get(), st);
    std::vector<int> eA((size_t)ld * oz.nchunk);
    NAPMO_HIP(hipMemcpyAsync(eA.data(), oz_eA.get(), eA.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    NAPMO_HIP(hipStreamSynchronize(st));
    order = ozaki_compact_order(oz, eA);
    oz_perm.upload(order, st);
  }
  launch_ozaki_setup(oz, phi64.get(), grid.wbw.get(), oz_perm.get(), oz_planes.get(), oz_eA.get(),
                     oz_maxd1.get(), st);
  NAPMO_HIP(hipStreamSynchronize(st));
}

std::pair<long, long> FockPlan::int8_screen_stats() {
  NAPMO_REQUIRE(prec == NAPMO_PREC_INT8, "int8 screening statistics need an NAPMO_PREC_INT8 plan");
  std::vector<int> eA((size_t)ld * oz.nchunk);
  std::vector<double> vs(oz.Nq);
  NAPMO_HIP(hipStreamSynchronize(stream));
  oz_eA.download(eA.data(), eA.size(), stream);
  oz_vs.download(vs.data(), vs.size(), stream);
  NAPMO_HIP(hipStreamSynchronize(stream));
  return {ozaki_kept_blocks(oz, eA, vs, oz_screen), (long)oz.nchunk * oz.nblk};
}

void FockPlan::ozaki_contract(const double *veff, double *dF, int slot) {
  hipStream_t st = stream;
  launch_ozaki_vs(veff, grid.wbw.get(), grid.N, oz.Nq, oz_vs.get(), st);
  mark(slot, 7);
  if (oz_pack)
    launch_ozaki_fock_packed(oz, oz_planes.get(), oz_eA.get(), oz_maxd1.get(), oz_vs.get(), oz_screen,
                             oz_part.get(), oz_ptasks.get(), oz_pctr.get(), oz_pcache.get(),
                             oz_pccount.get(), oz_pcev.get(), st);
  else
    launch_ozaki_fock(oz, oz_tasks.get(), oz_planes.get(), oz_eA.get(), oz_maxd1.get(), oz_vs.get(),
                      oz_screen, oz_part.get(), st);
  mark(slot, 8);
  launch_ozaki_reduce(oz, oz_part.get(), oz_perm.get(), dF, st);
}

void FockPlan::contract(int slot, const double *dP, const double *veff, double *dF, double *dE,
                        bool s_ready) {
  hipStream_t st = stream;
  if (prec == NAPMO_PREC_INT8) {
    ozaki_contract(veff, dF, slot);
    mark(slot, 9);
    launch_traces(dP, dF, n, grid.wbw.get(), rho.get(), veff, grid.N, dE, trace_work.get(), st);
    mark(slot, 10);
    built = true;
    return;
  }
  if (!s_ready) launch_scale3(veff, grid.wbw.get(), nullptr, grid.N, Npad, s.get(), st);
  mark(slot, 7);
  // the one-pass kernel writes FP32 partials into the same buffer
  float *part32 = reinterpret_cast<float *>(F_part.get());
  if (use_pass == 2)
    launch_fock_syrk_pass(psh, phiT32.get(), s.get(), part32, st);
  else if (use_pass == 1)
    launch_fock_bf16x3_pass(psh, phiT_hi.get(), phiT_lo.get(), s.get(), part32, st);
  else if (prec == NAPMO_PREC_BF16X3)
    launch_fock_bf16x3(fsh, phiT_hi.get(), phiT_lo.get(), phiT32.get(), s.get(), F_part.get(), st);
  else
    launch_fock_f64(fsh, phi64.get(), phi64.get(), s.get(), F_part.get(), st);
  mark(slot, 8);
  if (use_pass)
    launch_fock_reduce_f32(psh, part32, n, dF, st);
  else
    launch_fock_reduce(fsh, F_part.get(), n, 0, 1.0, 0, dF, st);
  mark(slot, 9);
  launch_traces(dP, dF, n, grid.wbw.get(), rho.get(), veff, grid.N, dE, trace_work.get(), st);
  mark(slot, 10);
  built = true;
}

void FockPlan::adjoint_potential(const double *y, double *z) {
  hipStream_t st = stream;
  PoissonPlan &pl = grid.plan(lmax);
  AtomTables at = grid.tables(pl);
  SynthPairs &sp = grid.synth_pairs();
  const long L = (long)(lmax + 1) * (lmax + 1);
  long nrad_total = 0;
  for (const auto &a : grid.atoms) nrad_total += a.nrad;
  adj_T.resize(std::max<long>(synth_table_size(at, lmax), 1));
  adj_aL.resize((size_t)grid.M * L);
  adj_aR.resize((size_t)grid.M * L);
  adj_Uy.resize((size_t)nrad_total * L);
  adj_Udt.resize((size_t)nrad_total * L);
  adj_rt.resize((size_t)nrad_total * L);
  launch_poisson_adjoint(at, lmax, sp.geo.get(), sp.list.get(), sp.slot0.get(), grid.owner.get(),
                         grid.bw.get(), pl.yhat.get(), pl.green.get(), grid.N, y, adj_T.get(),
                         adj_aL.get(), adj_aR.get(), adj_Uy.get(), adj_Udt.get(), adj_rt.get(), z,
                         st);
}

void FockPlan::symmetric_weight() {
  NAPMO_REQUIRE(built, "energy derivatives: run a build with this density first");
  hipStream_t st = stream;
  adj_y.resize(Npad);
  adj_z.resize(Npad);
  ssym.resize(Npad);
  // y = W rho, z = V^T y, ssym = (W v + z) / 2 with W v = s of the build
  launch_scale3(rho.get(), grid.wbw.get(), nullptr, grid.N, Npad, adj_y.get(), st);
  adjoint_potential(adj_y.get(), adj_z.get());
  launch_half_sum(s.get(), adj_z.get(), grid.N, Npad, ssym.get(), st);
}

// Points per chunk of the recollocating backward: a multiple of the Fock chunk, of the
// gradient's 4096-point chunk and of the density tile, so every partial sum covers the same
// points as in the stored-Phi path.
long FockPlan::backward_chunk() const {
  const long unit = std::lcm(std::lcm(fsh.chunk, grad_chunk_points()), 256L);
  return std::max<long>(1, div_up(std::max<long>(bwd_points, 1), unit)) * unit;
}

void FockPlan::energy_dP(double *dEdP) {
  symmetric_weight();
  hipStream_t st = stream;
  if (bwd_points <= 0) {
    launch_fock_f64(fsh, phi64.get(), phi64.get(), ssym.get(), F_part.get(), st);
  } else {
    // Phi recollocated chunk by chunk; each chunk's Fock partials land where the stored-Phi
    // launch writes them
    const long C = backward_chunk();
    bwd_phi.resize((size_t)std::min(C, Npad) * ld);
    for (long p0 = 0; p0 < Npad; p0 += C) {
      const long cn = std::max<long>(0, std::min(C, grid.N - p0)), cpad = std::min(C, Npad - p0);
      basis.collocate(grid.x.get() + p0, grid.y.get() + p0, grid.z.get() + p0, cn, cpad, ld,
                      bwd_phi.get(), nullptr, st);
      FockShape cs = fsh;
      cs.Npad = cpad;
      cs.nchunk = (int)div_up(cpad, fsh.chunk);
      launch_fock_f64(cs, bwd_phi.get(), bwd_phi.get(), ssym.get() + p0,
                      F_part.get() + (p0 / fsh.chunk) * fsh.cstride(), st);
    }
  }
  launch_fock_reduce(fsh, F_part.get(), n, 0, 1.0, 0, dEdP, st);
}

void FockPlan::gradient_device(const double *dP, double *dG) {
  symmetric_weight();
  hipStream_t st = stream;
  launch_pad2d(dP, n, n, P.get(), ld, ld, st);
  const int ncentre = (int)centres.size() / 3;
  grad_part.resize((size_t)n * grad_chunks(grid.N) * 3);
  if (bwd_points <= 0) {
    X.resize((size_t)Npad * ld);
    launch_phi_times(phi64.get(), ld, Npad, P.get(), ld, X.get(), st);
    launch_grad_centres(grid.x.get(), grid.y.get(), grid.z.get(), grid.N, ssym.get(), X.get(), ld,
                        n, basis.origin.get(), basis.lxyz.get(), basis.first.get(),
                        basis.nprim.get(), basis.norm.get(), basis.zeta.get(), basis.coef.get(),
                        centre_of.get(), ncentre, grad_part.get(), dG, st);
    return;
  }
  // Phi recollocated and Phi P formed chunk by chunk; the partial sums of every 4096-point
  // chunk land where the one-pass kernel writes them, then one reduction
  const long C = backward_chunk();
  bwd_phi.resize((size_t)std::min(C, Npad) * ld);
  bwd_X.resize((size_t)std::min(C, Npad) * ld);
  for (long p0 = 0; p0 < grid.N; p0 += C) {
    const long cn = std::min(C, grid.N - p0), cpad = std::min(C, Npad - p0);
    basis.collocate(grid.x.get() + p0, grid.y.get() + p0, grid.z.get() + p0, cn, cpad, ld,
                    bwd_phi.get(), nullptr, st);
    launch_phi_times(bwd_phi.get(), ld, cpad, P.get(), ld, bwd_X.get(), st);
    launch_grad_basis_part(grid.x.get(), grid.y.get(), grid.z.get(), cn, p0, grid.N, ssym.get(),
                           bwd_X.get(), ld, n, basis.origin.get(), basis.lxyz.get(),
                           basis.first.get(), basis.nprim.get(), basis.norm.get(),
                           basis.zeta.get(), basis.coef.get(), grad_part.get(), st);
  }
  launch_grad_reduce(grad_part.get(), n, grid.N, centre_of.get(), ncentre, dG, st);
}

int FockPlan::next_slot() { return (prof_used < prof_cap) ? prof_used++ : -1; }

void FockPlan::build_device(const double *dP, double *dF, double *dE) {
  const int slot = next_slot();
  density_P(slot, dP);
  expand(slot);
  synth_self(slot, true);
  contract(slot, dP, v.get(), dF, dE, true);
}

void FockPlan::build_orbitals_device(const double *dC, int nocc, double eta, double *dF,
                                     double *dE) {
  const int slot = next_slot();
  density_C(slot, dC, nocc, eta);
  expand(slot);
  synth_self(slot, true);
  contract(slot, Pfull.get(), v.get(), dF, dE, true);
}

// Host-pointer builds check their inputs and the energies they return: a non-finite value
// in P (or C, eta, v) is reported with its position, and a non-finite energy (a density or
// potential that is not finite, e.g. from a basis with a non-finite exponent or coefficient)
// is reported instead of being returned.
static void require_finite(const double *x, size_t count, long ncol, const char *what) {
  for (size_t i = 0; i < count; ++i)
    if (!std::isfinite(x[i]))
      throw Error(std::string(what) + " holds a non-finite value at row " + std::to_string(i / ncol) +
                  ", column " + std::to_string(i % ncol));
}

static void require_finite_energy(const double *E) {
  if (!std::isfinite(E[0]) || !std::isfinite(E[1]))
    throw Error("grid Fock build: non-finite energy (the density or the potential on the grid is "
                "not finite)");
}

void FockPlan::build_orbitals_host(const double *C, int nocc, double eta, double *F_out,
                                   double *E_out, double *rho_out, double *v_out) {
  hipStream_t st = stream;
  NAPMO_REQUIRE(nocc >= 0, "orbital build: negative orbital count");
  require_finite(C, (size_t)n * nocc, std::max(nocc, 1), "C");
  require_finite(&eta, 1, 1, "eta");
  DeviceBuffer<double> dC;
  dC.upload(C, (size_t)n * nocc, st);
  build_orbitals_device(dC.get(), nocc, eta, F.get(), E.get());
  double Eh[2];
  E.download(Eh, 2, st);
  if (F_out) F.download(F_out, (size_t)n * n, st);
  if (rho_out) rho.download(rho_out, grid.N, st);
  if (v_out) v.download(v_out, grid.N, st);
  NAPMO_HIP(hipStreamSynchronize(st));
  require_finite_energy(Eh);
  if (E_out) E_out[0] = Eh[0], E_out[1] = Eh[1];
}

void FockPlan::contract_potential(const double *v_host, double *F_out) {
  hipStream_t st = stream;
  require_finite(v_host, grid.N, 1, "v");
  DeviceBuffer<double> dv;
  dv.upload(v_host, grid.N, st);
  if (prec == NAPMO_PREC_INT8) {
    ozaki_contract(dv.get(), F.get(), -1);
    F.download(F_out, (size_t)n * n, st);
    NAPMO_HIP(hipStreamSynchronize(st));
    return;
  }
  launch_scale3(dv.get(), grid.wbw.get(), nullptr, grid.N, Npad, s.get(), st);
  float *part32 = reinterpret_cast<float *>(F_part.get());
  if (use_pass == 2)
    launch_fock_syrk_pass(psh, phiT32.get(), s.get(), part32, st);
  else if (use_pass == 1)
    launch_fock_bf16x3_pass(psh, phiT_hi.get(), phiT_lo.get(), s.get(), part32, st);
  else if (prec == NAPMO_PREC_BF16X3)
    launch_fock_bf16x3(fsh, phiT_hi.get(), phiT_lo.get(), phiT32.get(), s.get(), F_part.get(), st);
  else
    launch_fock_f64(fsh, phi64.get(), phi64.get(), s.get(), F_part.get(), st);
  if (use_pass)
    launch_fock_reduce_f32(psh, part32, n, F.get(), st);
  else
    launch_fock_reduce(fsh, F_part.get(), n, 0, 1.0, 0, F.get(), st);
  F.download(F_out, (size_t)n * n, st);
  NAPMO_HIP(hipStreamSynchronize(st));
}

void FockPlan::overlap(double *S_out) {
  hipStream_t st = stream;
  launch_scale3(grid.w.get(), grid.bw.get(), nullptr, grid.N, Npad, s.get(), st);
  launch_fock_f64(fsh, phi64.get(), phi64.get(), s.get(), F_part.get(), st);
  launch_fock_reduce(fsh, F_part.get(), n, 0, 1.0, 0, F.get(), st);
  F.download(S_out, (size_t)n * n, st);
  NAPMO_HIP(hipStreamSynchronize(st));
}

void FockPlan::build_host(const double *hP, double *F_out, double *E_out, double *rho_out,
                          double *v_out) {
  hipStream_t st = stream;
  require_finite(hP, (size_t)n * n, n, "P");
  DeviceBuffer<double> dP;
  dP.upload(hP, (size_t)n * n, st);
  build_device(dP.get(), F.get(), E.get());
  double Eh[2];
  E.download(Eh, 2, st);
  if (F_out) F.download(F_out, (size_t)n * n, st);
  if (rho_out) rho.download(rho_out, grid.N, st);
  if (v_out) v.download(v_out, grid.N, st);
  NAPMO_HIP(hipStreamSynchronize(st));
  require_finite_energy(Eh);
  if (E_out) E_out[0] = Eh[0], E_out[1] = Eh[1];
}

}  // namespace napmo
