// Launchers of the gfx950 kernels (kernels.hip). All pointers are device pointers;
// every launcher enqueues on `stream` and returns without synchronising.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace napmo {

constexpr int MAX_LMAX = 24;   // synthesis / projection instantiated for lmax <= MAX_LMAX

// Scaled PI_l^m recurrence of the synthesis: PI_l^m(z) = sigma(l, m) Pt_l^m(z) with
// Pt_m^m = 1, Pt_{m+1}^m = z, Pt_l^m = z Pt_{l-1}^m - beta(l, m) Pt_{l-2}^m (r = 1), which
// needs one multiply and one FMA per step; sigma(l, m) is folded into the spline tables.
// From PI_l^m = ((2l-1) z PI_{l-1}^m - (l+m-1) PI_{l-2}^m) / (l-m), PI_m^m = (2m-1)!!,
// PI_{m+1}^m = (2m+1) z PI_m^m.
constexpr double pi_sigma(int l, int m) {
  double s = 1.0;
  for (int k = 1; k <= m; ++k) s *= (2 * k - 1);
  if (l > m) s *= (2 * m + 1);
  for (int k = m + 2; k <= l; ++k) s *= (double)(2 * k - 1) / (double)(k - m);
  return s;
}
constexpr double pi_beta(int l, int m) {
  return (double)(l + m - 1) / (double)(l - m) * pi_sigma(l - 2, m) / pi_sigma(l, m);
}

constexpr int TILE = 64;       // MFMA output tile edge of the contractions

// Becke partition (src/grid/becke.cpp:79-176) for every point; owner[p] = atom of p.
// R: [M*M] centre distances; As: [M*M] clamped size adjustments with the sign of the
// reference evaluation (negative when i < j).
void launch_becke(const double *x, const double *y, const double *z, long N, const int *owner,
                  int M, const double *center, const double *R, const double *As, double *bw,
                  hipStream_t s);

// Collocation Phi[p, mu] (src/system/gto.cpp:41-51, 182-192) into out64 [Npad x ld]
// (and out32 when non-null). Rows p >= N and columns mu >= n are written as zero.
void launch_collocate(const double *x, const double *y, const double *z, long N, long Npad, int n,
                      int ld, const double *fn_origin, const int *fn_l, const int *fn_first,
                      const int *fn_nprim, const double *fn_norm, const double *prim_zeta,
                      const double *prim_coef, double *out64, float *out32, hipStream_t s);

// Per-atom geometry tables shared by the Poisson kernels.
struct AtomTables {
  int M;
  const long *point_off;    // [M+1] first point of each atom
  const int *nrad;          // [M]
  const int *nang;          // [M]
  const long *rad_off;      // [M] first radial shell of the atom in concatenated radial arrays
  const long *yhat_off;     // [M] offset of the atom's Lebedev projection table
  const long *green_off;    // [M] offset of the atom's Green matrices (per l: y then dt)
  const int *rkind;         // [M] radial transform kind
  const double *rp0, *rp1;  // [M] transform parameters
  const double *first_x, *last_x;  // [M] spline domain
  const double *center;     // [M*3]
  const double *r;          // [sum nrad] radial nodes
  int max_nrad;
  // first element of centre A's row in the synthesis' per-centre sums (contrib); null: A N
  // (a grid of one molecule; several molecules side by side give each its own [M_m x N_m])
  const long *cbase = nullptr;
  // per centre, the largest l of its expansion (null: lmax): the projection writes zeros
  // above it without summing them, the radial solves skip those l, and the synthesis
  // tables hold the centre's own harmonics only, in its own synthesis order
  const int *lcut = nullptr;
  // per centre, its first synthesis-table element (null: A (max_nrad - 1) L 4), rows of
  // (lcut + 1)^2 x 4 coefficients; amplitude rows amp_stride apart (0: L)
  const long *toff = nullptr;
  int amp_stride = 0;
};

// rho_lm[A][i][lm] = sum_j f_Aij Yhat[j][lm], f = dens * bw  (spherical expansion).
// nset densities at once: density k at dens + k dstride, its projections at rho_lm + k ostride.
void launch_project(const AtomTables &at, const double *dens, const double *bw, const double *yhat,
                    int lmax, double *rho_lm, hipStream_t s, int nset = 1, long dstride = 0,
                    long ostride = 0);
// U[A] (y and dt at the radial nodes, lm reference order) = G_{A,l} rho_lm[A].
// nset sets at once: set k reads rho_lm + k stride and writes Uy, Udt + k stride.
void launch_radial(const AtomTables &at, const double *rho_lm, const double *green, int lmax,
                   double *Uy, double *Udt, hipStream_t s, int nset = 1, long stride = 0);
// Synthesis tables: T[A][j][q][4] Horner coefficients in synthesis order q, amp_left /
// amp_right [A][q], all scaled by sqrt(2l+1)/sqrt(4 pi).
// write_amps = 0 keeps ampL / ampR as provided (non-potential extrapolations).
// nset sets at once: set k reads Uy, Udt + k ustride and writes T + k synth_table_size,
// ampL, ampR + k M L.
long synth_table_size(const AtomTables &at, int lmax);
// z [N] = V^T y [N]: the transposed Becke-Poisson operator of the grid (its (point, centre)
// pair buckets: geo, list and the first slot of every bucket slot0 [M (max_nrad + 1) + 1];
// owner [N] the atom of every point). Work arrays: Tt (synth_table_size), aLt, aRt [M L],
// Uyt, Udtt, rt [sum nrad x L].
void launch_poisson_adjoint(const AtomTables &at, int lmax, const double *geo, const int *list,
                            const long *slot0, const int *owner, const double *bw,
                            const double *yhat, const double *green, long N, const double *y,
                            double *Tt, double *aLt, double *aRt, double *Uyt, double *Udtt,
                            double *rt, double *z, hipStream_t s);
void launch_synth_tables(const AtomTables &at, const double *Uy, const double *Udt, int lmax,
                         double *T, double *ampL, double *ampR, hipStream_t s, int write_amps = 1,
                         int nset = 1, long ustride = 0);
// Synthesis-order index q -> (l, m, cos/sin) and its coefficient scale (host copy).
void synth_order(int lmax, std::vector<int> &l, std::vector<int> &m, std::vector<int> &cs,
                 std::vector<double> &scale);
// v[p] (+)= sum_A sum_lm U^A_lm(|p - R_A|) S_lm(unit) (src/horton/evaluate.cpp:129-223),
// for points given as SoA (x, y, z). accumulate = add into v instead of overwriting.
// ext_mode 0: PotentialExtrapolation (ampL, ampR); 1: no tail, cusp left (a0, b0).
void launch_synth(const AtomTables &at, const double *x, const double *y, const double *z, long N,
                  const double *T, const double *ampL, const double *ampR, int lmax, int accumulate,
                  double *v, hipStream_t s, int ext_mode = 0);

template <int LMAX>
void launch_synth_t(dim3 g, dim3 b, hipStream_t s, const AtomTables &at, const double *x,
                    const double *y, const double *z, long N, const double *T, const double *aL,
                    const double *aR, int acc, int ext_mode, double *v);

template <int LMAX>
void launch_synth2_t(int np, hipStream_t s, const AtomTables &at, const double *geo, long N,
                     const int *list, const int *wave_bucket, long nwaves, const double *T,
                     const double *aL, const double *aR, double *contrib);

// Pair buckets of the bucketed synthesis (per geometry): count[gb] += 1 for every
// (point, centre); then list[off[gb] + k] = p (slots pre-filled with -1).
void launch_pair_count(const AtomTables &at, const double *x, const double *y, const double *z,
                       long N, int *count, hipStream_t s);
void launch_pair_fill(const AtomTables &at, const double *x, const double *y, const double *z,
                      long N, const long *off, int *cursor, int *list, hipStream_t s);
// Per-pair geometry of the bucketed synthesis, geo [nwaves][4][64] in list order: u (interval
// buckets) or d (extrapolation buckets), then the unit vector (x, y, z) from the centre.
void launch_pair_geometry(const AtomTables &at, const double *x, const double *y,
                          const double *z, const int *list, const int *wave_bucket, long nwaves,
                          double *geo, hipStream_t s);
// Bucketed synthesis into contrib [M x N], then v[p] (+)= scale * sum_A contrib[A][p].
// np = (point, centre) pairs per lane, 1 or 2 (the lists are padded to 128 per bucket).
constexpr int SYNTH_PAD = 128;
void launch_synth2(const AtomTables &at, const double *geo, long N,
                   const int *list, const int *wave_bucket, long nwaves, const double *T,
                   const double *ampL, const double *ampR, int lmax, double *contrib,
                   hipStream_t s, int np = 2);
void launch_contrib_reduce(const double *contrib, int M, long N, int accumulate, double *v,
                           hipStream_t s, double scale = 1.0);
// Synthesis of many densities on FP64 MFMA (synth_mfma.hpp): 16 densities per tile.
// Tables in the density-interleaved layout of ndp = 16 ntile columns (sets nset .. ndp - 1
// written as zeros): T [A][interval][q][ndp][4], ampL / ampR [A][q][ndp], from the radial
// solutions of set s at Uy, Udt + s ustride.
void launch_synth_tables_cols(const AtomTables &at, const double *Uy, const double *Udt, int lmax,
                              int nset, long ustride, int ndp, double *T, double *ampL,
                              double *ampR, hipStream_t s);
// contrib tile t [M][N][16] at contrib + t M N 16 for t < ntile (pairs as for launch_synth2).
void launch_synth_mfma(const AtomTables &at, const double *geo, long N, const int *list,
                       const int *wave_bucket, long nwaves, int ntile, const double *T,
                       const double *ampL, const double *ampR, int lmax, double *contrib,
                       hipStream_t s);
template <int LMAX>
void launch_synth_mfma_t(hipStream_t s, const AtomTables &at, const double *geo, long N,
                         const int *list, const int *wave_bucket, long nslots, int ntile,
                         int ndp, const double *T, const double *aL, const double *aR,
                         long tstride, double *contrib);
// out[k ostride + p] = scale sum_A tile[A][p][k] for the nvalid densities of one tile.
void launch_contrib_reduce_tile(const double *tile, int M, long N, int nvalid, double scale,
                                double *out, long ostride, hipStream_t s);
// v = sum_A contrib[A] and sv = v wbw in one pass (sv [Npad], zero beyond N): the same
// values as launch_contrib_reduce followed by launch_scale3(v, wbw).
void launch_contrib_reduce_scale(const double *contrib, int M, long N, long Npad, const double *wbw,
                                 double *v, double *sv, hipStream_t s);
// Point-charge potential V[p] = sum_A q_A / |p - R_A| (o: [nq x 3]).
void launch_nuclear(const double *x, const double *y, const double *z, long N, const double *q,
                    const double *o, int nq, double *V, hipStream_t s);
// Exchange potential pieces: chi = D^T psi; out = a o b; K[v] += aux chi[s] (and K[s] +=
// aux chi[v] for s != v), all [ndim x N] function-major.
void launch_exchange_chi(const double *psi, const double *D, int ndim, long N, double *chi,
                         hipStream_t st);
void launch_mul(const double *a, const double *b, long N, double *out, hipStream_t st);
void launch_exchange_acc(const double *aux, const double *chi, int s, int v, long N, double *K,
                         hipStream_t st);
// v = a x + b y over N points.
void launch_axpby(double a, const double *x, double b, const double *y, long N, double *v,
                  hipStream_t s);

// ---- helmholtz.hip: the Becke-Dickson orbital correction (napmo/wavefunction/ndpsi.py)
// Finite-element tables of one radial grid (helmholtz_tables in poisson_plan.cpp), per atom
// offsets into concatenated arrays: c0 [2n x 7] band of the Galerkin matrix without the
// a term, oa [2n x 7 x 6] a-term integrals, o2 [2n x 6] right-hand-side integrals of the
// merged source f, j1 / j2 [sum nrad] (indexed like AtomTables::r).
struct HelmholtzTables {
  const double *c0, *oa, *o2, *j1, *j2;
  const long *c0_off, *oa_off, *o2_off;
};
// LDS of one radial system of n points: 2n x 13 band and multiplier doubles, 2n spline
// factors and n/2 pivot words, and at least 3 columns of 2n: 280 n bytes (143 KB at 512)
constexpr int HELMHOLTZ_MAX_NRAD = 512;
// out[o][p] = a[o][p] * s[o] - b[o][p] for norb rows of N points (s: device [norb]).
void launch_scaled_diff(const double *a, const double *s, const double *b, int norb, long N,
                        double *out, hipStream_t st);
// Vsph[o][rad_off[A] + i] = sum_j f w bw / sum_j w over shell i of atom A (f = V[o]).
void launch_shell_average(const AtomTables &at, const double *V, int norb, long N,
                          const double *w, const double *bw, long nrad_total, double *Vsph,
                          hipStream_t st);
// Per (atom, l, orbital): assemble u'' + (2/r) u' + a u = c mu (c = -2/m_inv,
// a = c (l(l+1)/r^2 + Vsph - oi - doi)), factor it (band LU, partial pivoting) and solve
// the 2l+1 right-hand sides: Uy / Udt [norb][nrad_total x L]; status[0] = 1 on a zero pivot.
void launch_helmholtz_radial(const AtomTables &at, const HelmholtzTables &ht, int lmax, int norb,
                             long nrad_total, const double *Vsph, const double *oi,
                             const double *doi, double mass_inv, const double *mu, double *Uy,
                             double *Udt, int *status, hipStream_t st);
// Residual R[o] = T[o] + V[o] - e[o] psi[o] and de[o] = sum_p R psi w bw (fixed order).
// work: device scratch of norb * RESID_BLOCKS doubles.
constexpr int RESID_BLOCKS = 256;
void launch_residual(const double *psi, const double *T, const double *V, const double *e,
                     const double *wbw, int norb, long N, double *R, double *de, double *work,
                     hipStream_t st);

// ---- epc.hip: epc17-2 electron-proton correlation (napmo/functional/interspecies_correlation.py)
// Elementwise: ene, potE, potN [n] of (rE, rN) [n]; flag: device int set to 0 beforehand.
void launch_epc17_2(const double *rE, const double *rN, long n, int *flag, double *ene,
                    double *potE, double *potN, hipStream_t st);
// On the common points index [n x 2] of two grids: vE[i] += potE, vN[j] += potN and
// energy[0] = sum (wbwN[j] ene) rE[i]; part: EPC_PART doubles of scratch.
// Index rows outside [0, Na) x [0, Nb) are skipped.
constexpr int EPC_PART = 256;
void launch_epc17_2_pairs(const double *rE, const double *rN, const int *index, long n, long Na,
                          long Nb, const double *wbwN, int *flag, double *vE, double *vN,
                          double *part, double *energy, hipStream_t st);
// The same with the flag and partial energies in the file's static device scratch (no
// allocation; the scratch address is looked up once per device).
void launch_epc17_2_pairs_noalloc(const double *rE, const double *rN, const int *index, long n,
                                  long Na, long Nb, const double *wbwN, double *vE, double *vN,
                                  double *energy, hipStream_t st);

// rho[p] = Phi_p^T P Phi_p with FP64 MFMA (P [ld x ld] zero padded). partial [ntile x Npad].
// nreal: functions actually present (columns nreal .. ld - 1 of Phi and rows of P are zero),
// 0 = ld.
void launch_density_f64(const double *phi, int ld, long Npad, const double *P, double *partial,
                        hipStream_t s, int ks = 16, int pf = 1, int nreal = 0);
// rho[p] = scale * sum_t partial[t][p] for p < N.
void launch_rho_reduce(const double *partial, int ntile, long Npad, long N, double *rho,
                       hipStream_t s, double scale = 1.0);

// out[p] = (a[p] + b[p]) / 2 for p < N, zero for p < Npad beyond.
void launch_half_sum(const double *a, const double *b, long N, long Npad, double *out, hipStream_t st);
// Scale vector s[p] = a[p] * b[p] * c[p] (c may be null), zero for p >= N.
void launch_scale3(const double *a, const double *b, const double *c, long N, long Npad, double *s,
                   hipStream_t st);

// F = X^T diag(s) Y over Npad points. symmetric: only upper tiles of F (X == Y).
// partial buffer: [nchunk x ntiles x TILE x TILE]. Returns nothing; F written by reduce.
struct FockShape {
  int ld, nt, ntiles, nchunk;
  long Npad, chunk;
  bool symmetric;
  // doubles between consecutive chunks' partials: ntiles * TILE * TILE plus 256 B, so the
  // chunks of one element do not all fall on the same few L2 channels
  long cstride() const { return (long)ntiles * TILE * TILE + 32; }
};
FockShape fock_shape(int ld, long Npad, bool symmetric, long wg_target = 0);
void launch_fock_f64(const FockShape &sh, const double *X, const double *Y, const double *s,
                     double *partial, hipStream_t st);
// F_out [n x n] row-major from the partials; symmetric mirrors upper tiles;
// upper_mirror: general tiles, then F_ij := F_ji for i > j from the upper triangle
// (the exchange block of src/wavefunction/nwavefunction.cpp:53-62).
void launch_fock_reduce(const FockShape &sh, const double *partial, int n, int upper_mirror,
                        double scale, int accumulate, double *F, hipStream_t st);
// Same for FP32 chunk partials (one-pass mixed kernel); symmetric shapes only.
void launch_fock_reduce_f32(const FockShape &sh, const float *partial, int n, double *F,
                            hipStream_t st);

// E[0] = 0.5 * sum_ij P_ij F_ij (FP64), E[1] = 0.5 * sum_p w_p rho_p v_p.
// work: device scratch of 2 * TRACE_WORK doubles.
constexpr int TRACE_WORK = 512;
void launch_traces(const double *P, const double *F, int n, const double *wbw, const double *rho,
                   const double *v, long N, double *E, double *work, hipStream_t st);

// ---- kernels_mixed.hip
// phi64 [Npad x ld] -> function-major bf16 hi / lo bits and float32 copies [ld x Npad].
void launch_split_phi(const double *phi64, int ld, long Npad, unsigned short *hi, unsigned short *lo,
                      float *f32, hipStream_t s);
// F partials with the 3-term bf16 split (see kernels_mixed.hip); same shape as the FP64 one.
void launch_fock_bf16x3(const FockShape &sh, const unsigned short *Ahi, const unsigned short *Alo,
                        const float *B32, const double *s, double *partial, hipStream_t st);
// One pass over Phi (all upper tiles per chunk of points); same partial layout as
// fock_shape(ld, Npad, true) with the pass's own chunking.
FockShape fock_pass_shape(int ld, long Npad, long wg_target = 0);
bool fock_pass_supported(int ld);
// F = G^T diag(sign s) G, G = sqrt|s| Phi, on the same shape and FP32 partial layout
// (phi32: the function-major FP32 copy of Phi).
bool fock_syrk_supported(int ld);
void launch_fock_syrk_pass(const FockShape &sh, const float *phi32, const double *s,
                           float *partial, hipStream_t st);
void launch_fock_bf16x3_pass(const FockShape &sh, const unsigned short *Ahi,
                             const unsigned short *Alo, const double *s, float *partial,
                             hipStream_t st);
// partial[ct][p] = sum over orbital tile ct of (Phi C)_pi^2, C zero-padded [ld x ldc];
// ldc = density_orb_cols(nocc), density_orb_tiles(ldc) partial rows.
int density_orb_cols(int nocc);
// X [Npad x ldb] = Phi [Npad x ld] B [ld x ldb] on FP64 MFMA (ld, ldb multiples of 64).
void launch_phi_times(const double *phi, int ld, long Npad, const double *B, int ldb, double *X,
                      hipStream_t s);
// Centre gradient of 1/2 sum w bw rho v (gradient.hip): grad [ncentre x 3]; s = w bw v,
// X = Phi P [N x ldx]; basis arrays as for collocation; partial [n x grad_chunks(N) x 3].
int grad_chunks(long N);
void launch_grad_centres(const double *x, const double *y, const double *z, long N, const double *s,
                         const double *X, int ldx, int n, const double *fo, const int *fl,
                         const int *ff, const int *fnp, const double *fnorm, const double *pz,
                         const double *pc, const int *centre_of, int ncentre, double *partial,
                         double *grad, hipStream_t st);
// The same in pieces: partial sums of the points p0 .. p0 + N - 1 of an Ntotal-point grid
// (p0 a multiple of grad_chunk_points(); X holds those points' rows), then the reduction
// over all chunks: equal, bit for bit, to one launch_grad_centres over all points.
long grad_chunk_points();
void launch_grad_basis_part(const double *x, const double *y, const double *z, long N, long p0,
                            long Ntotal, const double *s, const double *X, int ldx, int n,
                            const double *fo, const int *fl, const int *ff, const int *fnp,
                            const double *fnorm, const double *pz, const double *pc,
                            double *partial, hipStream_t st);
void launch_grad_reduce(const double *partial, int n, long Ntotal, const int *centre_of,
                        int ncentre, double *grad, hipStream_t st);
int density_orb_tiles(int ldc);
// With one column tile (ldc <= 64), nvalid >= 0 makes the kernel write
// partial[p] = scale * sum for p < nvalid directly (rho itself, no reduction pass).
void launch_density_orb_f64(const double *phi, int ld, long Npad, const double *C, int ldc,
                            double *partial, hipStream_t s, int ks = 16, int pf = 1, int nreal = 0,
                            double scale = 1.0, long nvalid = -1);
// Orbital form with all of C [nrow x ncol] (row-major as given, ncol <= 64, zero-padded to
// 16-column blocks in LDS) resident in LDS, Phi streamed into MFMA operands:
// out[p] = scale sum_c X_pc^2 for p < nvalid (nvalid < 0: Npad). shape 1..3: (rows per
// wave, waves per workgroup, K groups in flight) = (32, 12, 3), (48, 8, 2), (48, 8, 3).
bool density_orb_stream_supported(int ld, int ldc, int nreal);
// Screening of the streamed orbital-form density (S > 0): phi is then the plan's Phi with its
// columns in the order perm (row k of C taken as C[perm[k]]), eT [Npad / 16][ld / 16] int8 with
// every |phi| of a 16-row, 16-column block <= 2^eT; the (16-row tile, 16-function group)
// products with eT + eC <= -S, 2^eC >= sum over the group of max_c |C|, are skipped, so each
// X_pc moves by at most (ld / 16) 2^-S.
struct DensityScreen {
  const int *perm = nullptr;
  const signed char *eT = nullptr;
  int S = 0;
};
void launch_density_orb_stream(const double *phi, int ld, long Npad, const double *C, int nrow,
                               int ncol, double *out, hipStream_t s, double scale, long nvalid,
                               int shape = 1, const DensityScreen &sc = DensityScreen());
// Per (16-row tile, column) exponent of max |phi| over the 16 rows: out [Npad / 16][ld] int8,
// clamped to [-127, 127], -128 for an all-zero tile column.
void launch_phi_tile_exp(const double *phi, int ld, long Npad, signed char *out, hipStream_t s);
// out [Npad x ld] = phi with its columns in the order perm (out[:, k] = phi[:, perm[k]]).
void launch_permute_cols(const double *phi, int ld, long Npad, const int *perm, double *out, hipStream_t s);
void launch_gram(const double *C, int n, int nocc, double eta, double *P, hipStream_t s);
void launch_pad2d(const double *src, int rows, int cols, double *dst, int drows, int dcols,
                  hipStream_t s);
void launch_selftest_mfma(int which, double *out, hipStream_t s);

// ---- grouped launches: many plans (independent molecules) in one launch per phase. Each
// member carries its own arguments; block0 is its first workgroup (a multiple of 8 where
// the kernel orders tiles by XCD), and a workgroup finds its member by binary search.
struct DensityMember {   // P-form density of one plan
  const double *phi, *B;
  double *partial;
  long Npad, nvalid, block0;
  int ld, ldb, kend;
  double scale;
};
struct FockMember {      // FP64 Fock partials of one plan
  const double *X, *s;
  double *partial;
  long Npad, chunk, block0;
  int ld, nt, ntiles, nchunk;
};
struct FockReduceMember {
  const double *partial;
  double *F;
  long block0;
  int nt, ntiles, nchunk, n;
};
struct TraceMember {     // E[0] = 1/2 sum P o F, E[1] = 1/2 sum wbw rho v
  const double *P, *F, *wbw, *rho, *v;
  double *E, *work;
  long nn, N;
};
struct PadMember {       // src [rows x cols] -> dst [drows x dcols], zero padded
  const double *src;
  double *dst;
  long block0;
  int rows, cols, drows, dcols;
};
struct RhoReduceMember {  // rho[p] = sum_t partial[t][p], p < N
  const double *partial;
  double *rho;
  long Npad, N, block0;
  int ntile;
};
struct ScaleMember {     // union points p0 .. p0 + N - 1, atoms A0 .. A1 - 1; s [Npad]
  double *s;
  long p0, N, Npad, block0;
  int A0, A1;
};
template <class T>
__host__ __device__ inline int group_member(const T *mem, int nmem, long g) {
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mem[mid].block0 <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}
// Fills d for one plan's P-form density (ks as for launch_density_f64); returns its
// workgroup count.
long density_group_fill(DensityMember &d, const double *phi, int ld, long Npad, const double *P,
                        double *partial, int nreal, int ks, long block0);
void launch_density_group(const DensityMember *d_mem, int nmem, long nblocks, int ks, int pf,
                          hipStream_t s);
long fock_group_fill(FockMember &f, const FockShape &sh, const double *X, const double *s,
                     double *partial, long block0);
void launch_fock_group(const FockMember *d_mem, int nmem, long nblocks, hipStream_t st);
long fock_reduce_group_fill(FockReduceMember &f, const FockShape &sh, const double *partial, int n,
                            double *F, long block0);
void launch_fock_reduce_group(const FockReduceMember *d_mem, int nmem, long nblocks, hipStream_t st);
void launch_traces_group(const TraceMember *d_mem, int nmem, hipStream_t st);
void launch_pad_group(const PadMember *d_mem, int nmem, long nblocks, hipStream_t st);
void launch_rho_reduce_group(const RhoReduceMember *d_mem, int nmem, long nblocks, hipStream_t st);
// Coupled species (APMO), member m = species m with charge q_m: per union point p of m,
// S_t = sum over t's centres of contrib[A Nu + p] (every species t, centres in order),
// vown = S_m, vcross = sum_{t != m} q_t S_t (in species order), veff = q_m^2 vown +
// q_m vcross; s_m[p - p0] = veff wbw[p] (zero from N to Npad).
struct SpeciesSum {
  int A0, A1;
  double q;
};
void launch_apmo_reduce_group(const ScaleMember *d_mem, int nmem, long nblocks,
                              const SpeciesSum *sp, const double *contrib, long Nu,
                              const double *wbw, double *vown, double *veff, hipStream_t st);
// E[2] = q2 E2[1] (own Coulomb energy), E[3] = E[1] - E[2] (coupling), per species.
struct ApmoEnergyMember {
  double *E;
  const double *E2;
  double q2;
};
void launch_apmo_energy_group(const ApmoEnergyMember *d_mem, int ns, hipStream_t st);
// v[p] = sum_{A0 <= A < A1} contrib[cbase[A] + p] (union point p of member m, its atoms in
// order) and s_m[p - p0] = v[p] wbw[p]; s_m zero from N to Npad.
void launch_contrib_reduce_scale_group(const ScaleMember *d_mem, int nmem, long nblocks,
                                       const double *contrib, const long *cbase,
                                       const double *wbw, double *v, hipStream_t st);
void launch_copy(const double *src, double *dst, long n, hipStream_t s);

}  // namespace napmo
