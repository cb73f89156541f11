// Auxiliary basis of the grid-based HF with orbital optimisation
// (src/grid/auxiliary_basis.cpp:10-280, bound at napmo/__init__.py:422-432, used by
// napmo/wavefunction/psi_optimization.py:230): per centre, Slater-type radial functions
// exp(-ln2 / r_i * |r - R|) for every radial node r_i, made orthonormal on the centre's
// atomic grid (canonical orthogonalisation with the threshold abldep); for l = 1..ablmax the
// s block times real spherical harmonics; then the whole set is orthonormalised on the
// molecular grid. A set-up step on the host: it runs once per grid, never in an SCF
// iteration, and its eigenproblems are small (at most ncenter nrad (ablmax + 1)^2).
//
// Kept as the reference computes it, including that the l > 0 blocks of every centre scale
// the *first* centre's s block (columns 0 .. nbas[ip] - 1 of the point's row), which only
// matters for ablmax > 0 with two or more centres.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "host.hpp"
#include "poisson_plan.hpp"

namespace napmo {

namespace {

// Symmetric eigenproblem by cyclic Jacobi rotations: A [n x n] row-major (destroyed),
// eigenvalues ascending in w, eigenvectors as the columns of V [n x n] row-major.
void jacobi_eigen(std::vector<double> &A, int n, std::vector<double> &w, std::vector<double> &V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; ++i) {
      diag += a(i, i) * a(i, i);
      for (int j = i + 1; j < n; ++j) off += a(i, j) * a(i, j);
    }
    if (off <= 1e-30 * diag || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a(p, q);
        if (apq == 0.0) continue;
        const double theta = (a(q, q) - a(p, p)) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int i, int j) { return a(i, i) < a(j, j); });
  std::vector<double> Vs((size_t)n * n);
  w.resize(n);
  for (int c = 0; c < n; ++c) {
    w[c] = a(order[c], order[c]);
    for (int k = 0; k < n; ++k) Vs[(size_t)k * n + c] = V[(size_t)k * n + order[c]];
  }
  V.swap(Vs);
}

// Columns of V [n x n] made orthonormal left to right (classical Gram-Schmidt against the
// already normalised columns, then normalised), as the reference re-cleans its eigenvectors.
void gram_schmidt_columns(std::vector<double> &V, int n) {
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < i; ++j) {
      double d = 0.0, c = 0.0;
      for (int k = 0; k < n; ++k) {
        d += V[(size_t)k * n + j] * V[(size_t)k * n + j];
        c += V[(size_t)k * n + i] * V[(size_t)k * n + j];
      }
      for (int k = 0; k < n; ++k) V[(size_t)k * n + i] -= c / d * V[(size_t)k * n + j];
    }
    double nr = 0.0;
    for (int k = 0; k < n; ++k) nr += V[(size_t)k * n + i] * V[(size_t)k * n + i];
    nr = std::sqrt(nr);
    for (int k = 0; k < n; ++k) V[(size_t)k * n + i] /= nr;
  }
}

// Columns of the canonical orthogonalisation of the overlap S [n x n]: v_i / sqrt(w_i) for
// the eigenpairs with w_i > thresh, ascending. Returns [n x nkeep] row-major.
std::vector<double> canonical_columns(std::vector<double> S, int n, double thresh, int &nkeep) {
  std::vector<double> w, V;
  jacobi_eigen(S, n, w, V);
  gram_schmidt_columns(V, n);
  std::vector<int> keep;
  for (int i = 0; i < n; ++i)
    if (w[i] > thresh) keep.push_back(i);
  nkeep = (int)keep.size();
  std::vector<double> C((size_t)n * nkeep);
  for (int c = 0; c < nkeep; ++c) {
    const double f = 1.0 / std::sqrt(w[keep[c]]);
    for (int k = 0; k < n; ++k) C[(size_t)k * nkeep + c] = V[(size_t)k * n + keep[c]] * f;
  }
  return C;
}

}  // namespace

AuxBasis *make_aux_basis(const DeviceGrid &g, int ablmax, double abldep) {
  NAPMO_REQUIRE(!g.atoms.empty(), "AuxiliaryBasis: the grid has no atomic grids");
  NAPMO_REQUIRE(ablmax >= 0, "AuxiliaryBasis: ablmax can not be negative");
  const int nc = g.M;
  const double ln2 = std::log(2.0);
  std::vector<std::vector<double>> aoexp(nc), C(nc);
  std::vector<int> nbas(nc);
  int nbastotal = 0;
  auto dist = [&](long p, const double *o) {
    const double dx = g.points[3 * p] - o[0], dy = g.points[3 * p + 1] - o[1], dz = g.points[3 * p + 2] - o[2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
  };
  for (int ip = 0; ip < nc; ++ip) {
    const AtomInfo &at = g.atoms[ip];
    const int nr = at.nrad;
    aoexp[ip].resize(nr);
    for (int i = 0; i < nr; ++i) aoexp[ip][i] = ln2 / at.r[i];
    std::vector<double> S((size_t)nr * nr, 0.0), e(nr);
    for (long p = g.offsets[ip]; p < g.offsets[ip + 1]; ++p) {
      const double R = dist(p, at.origin), wgt = g.weights[p];
      for (int i = 0; i < nr; ++i) e[i] = std::exp(-aoexp[ip][i] * R);
      for (int i = 0; i < nr; ++i)
        for (int j = 0; j < nr; ++j) S[(size_t)i * nr + j] += e[i] * (e[j] * wgt);
    }
    C[ip] = canonical_columns(S, nr, abldep, nbas[ip]);
    nbastotal += nbas[ip];
  }
  const int L = (ablmax + 1) * (ablmax + 1);
  const int ncol = nbastotal * L;
  // primitive auxiliary functions at every grid point: T [N x ncol]
  std::vector<double> T((size_t)g.N * ncol, 0.0), e, ylm(L);
  int norb = 0;
  for (int ip = 0; ip < nc; ++ip) {
    const AtomInfo &at = g.atoms[ip];
    const int nr = at.nrad, nb = nbas[ip];
    e.resize(nr);
    for (long p = 0; p < g.N; ++p) {
      const double R = dist(p, at.origin);
      for (int i = 0; i < nr; ++i) e[i] = std::exp(-aoexp[ip][i] * R);
      double *row = &T[(size_t)p * ncol];
      for (int c = 0; c < nb; ++c) {
        double s = 0.0;
        for (int i = 0; i < nr; ++i) s += e[i] * C[ip][(size_t)i * nb + c];
        row[norb + c] = s;
      }
    }
    norb += nb;
    for (int l = 1; l <= ablmax; ++l) {
      // angular factors of every point about centre ip: m = 0, then (cos, sin) per m
      const int first = norb;
      for (long p = 0; p < g.N; ++p) {
        const double dx = g.points[3 * p] - at.origin[0], dy = g.points[3 * p + 1] - at.origin[1],
                     dz = g.points[3 * p + 2] - at.origin[2];
        const double R = std::sqrt(dx * dx + dy * dy + dz * dz);
        real_ylm(std::acos(dz / R), std::atan2(dy, dx), l, ylm.data());
        double *row = &T[(size_t)p * ncol];
        const int base = l * l;  // index of (l, m = -l) in real_ylm's order
        int col = first;
        for (int c = 0; c < nb; ++c) row[col + c] = row[c] * ylm[base + l];  // m = 0
        col += nb;
        for (int m = 1; m <= l; ++m) {
          for (int c = 0; c < nb; ++c) {
            row[col + c] = row[c] * ylm[base + l + m];           // sqrt2 Q cos(m phi)
            row[col + nb + c] = row[c] * ylm[base + l - m];      // sqrt2 Q sin(m phi)
          }
          col += 2 * nb;
        }
      }
      norb += (2 * l + 1) * nb;
    }
  }
  // overlap of the whole set on the molecular (atomic) weights, then canonical orthogonalisation
  std::vector<double> S((size_t)norb * norb, 0.0);
  for (long p = 0; p < g.N; ++p) {
    const double *row = &T[(size_t)p * ncol];
    const double wgt = g.weights[p];
    for (int i = 0; i < norb; ++i) {
      const double ri = row[i] * wgt;
      if (ri == 0.0) continue;
      double *Si = &S[(size_t)i * norb];
      for (int j = 0; j < norb; ++j) Si[j] += ri * row[j];
    }
  }
  int nao = 0;
  std::vector<double> X = canonical_columns(S, norb, abldep, nao);
  auto *ab = new AuxBasis;
  ab->ncenter = nc;
  ab->nao = nao;
  ab->aobasis.assign((size_t)g.N * std::max(nao, 1), 0.0);
  for (long p = 0; p < g.N; ++p) {
    const double *row = &T[(size_t)p * ncol];
    double *out = &ab->aobasis[(size_t)p * nao];
    for (int k = 0; k < norb; ++k) {
      const double t = row[k];
      if (t == 0.0) continue;
      const double *Xk = &X[(size_t)k * nao];
      for (int i = 0; i < nao; ++i) out[i] += t * Xk[i];
    }
  }
  return ab;
}

}  // namespace napmo
