// Common points of two species' Becke grids (src/grid/multi_grid.cpp:14-132): the index
// pairs (i, j) with identical coordinates, i in the larger grid. The reference compares all
// Na x Nb coordinate triples; here the smaller grid's points go into a hash table keyed by
// their bit patterns (-0.0 folded onto 0.0, NaN never matches, as with ==), and the larger
// grid is scanned once in index order, so the pairs come out in the reference's order.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "poisson_plan.hpp"

namespace napmo {

namespace {

struct Key {
  uint64_t b[3];
  bool operator==(const Key &o) const { return b[0] == o.b[0] && b[1] == o.b[1] && b[2] == o.b[2]; }
};

struct KeyHash {
  size_t operator()(const Key &k) const {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (int c = 0; c < 3; ++c) {
      h ^= k.b[c] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
      h *= 0xBF58476D1CE4E5B9ull;
    }
    return (size_t)(h ^ (h >> 31));
  }
};

// false for a NaN coordinate (never equal to anything)
bool point_key(const double *p, Key &k) {
  for (int c = 0; c < 3; ++c) {
    double v = p[c];
    if (v != v) return false;
    if (v == 0.0) v = 0.0;  // -0.0 == 0.0
    std::memcpy(&k.b[c], &v, sizeof(v));
  }
  return true;
}

}  // namespace

struct MultiGridHandle {
  int nspecies = 0;
  std::vector<DeviceGrid *> grids;
  std::map<std::pair<int, int>, std::vector<int>> pairs;  // (min id, max id) -> rows
};

static std::vector<int> common_index(const DeviceGrid &ga, const DeviceGrid &gb) {
  const DeviceGrid *A = &ga, *B = &gb;
  if (B->N > A->N) std::swap(A, B);  // rows index the larger grid first (a on a tie)
  std::unordered_map<Key, std::vector<int>, KeyHash> table;
  table.reserve((size_t)B->N * 2);
  for (long j = 0; j < B->N; ++j) {
    Key k;
    if (point_key(&B->points[3 * j], k)) table[k].push_back((int)j);
  }
  std::vector<int> rows;
  for (long i = 0; i < A->N; ++i) {
    Key k;
    if (!point_key(&A->points[3 * i], k)) continue;
    auto it = table.find(k);
    if (it == table.end()) continue;
    for (int j : it->second) rows.push_back((int)i), rows.push_back(j);
  }
  return rows;
}

}  // namespace napmo

using napmo::MultiGridHandle;

namespace {
template <class F> auto guard_abort(const char *name, F &&f) -> decltype(f()) {
  try {
    napmo::clear_error();
    return f();
  } catch (const std::exception &e) {
    napmo::set_error(e.what());
    std::fprintf(stderr, "napmo_gpu: %s: %s\n", name, e.what());
    std::fflush(stderr);
    std::abort();
  }
}
}  // namespace

extern "C" {

void *MultiGrid_new(int nspecies) {
  auto *m = new MultiGridHandle;
  m->nspecies = nspecies;
  return m;
}

void MultiGrid_del(void *mg) { delete (MultiGridHandle *)mg; }

int MultiGrid_add_grid(void *mg, void *grid) {
  auto *m = (MultiGridHandle *)mg;
  m->grids.push_back(&((napmo::BeckeHandle *)grid)->grid);
  return (int)m->grids.size() - 1;
}

int *MultiGrid_get_common_index(void *mg, int a, int b) {
  return guard_abort(__func__, [&]() -> int * {
    auto *m = (MultiGridHandle *)mg;
    const int n = (int)m->grids.size();
    NAPMO_REQUIRE(a >= 0 && a < n && b >= 0 && b < n, "MultiGrid: grid id out of range");
    const auto key = std::make_pair(std::min(a, b), std::max(a, b));
    auto it = m->pairs.find(key);
    if (it == m->pairs.end()) {
      std::vector<int> rows;
      if (a == b) {
        for (long i = 0; i < m->grids[a]->N; ++i) rows.push_back((int)i), rows.push_back((int)i);
      } else {
        rows = napmo::common_index(*m->grids[a], *m->grids[b]);
      }
      it = m->pairs.emplace(key, std::move(rows)).first;
    }
    return it->second.data();
  });
}

int MultiGrid_get_common_index_size(void *mg, int a, int b) {
  auto *m = (MultiGridHandle *)mg;
  auto it = m->pairs.find(std::make_pair(std::min(a, b), std::max(a, b)));
  return it == m->pairs.end() ? -1 : (int)(it->second.size() / 2);
}

int MultiGrid_get_ngrids(void *mg) { return (int)((MultiGridHandle *)mg)->grids.size(); }

int MultiGrid_get_nspecies(void *mg) { return ((MultiGridHandle *)mg)->nspecies; }

}  // extern "C"
