/*
 * neighborlist.hpp -- the reference's Rust API (src/api.rs) for C++ callers,
 * header-only, over the C ABI of nl_api.h (link libnlmi355).
 *
 * Same names, argument meaning and results as neighborlist_rs::api:
 *   NeighborList               src/api.rs:5-17   edge_index = sources then targets
 *                                                 (2 * num_edges), shifts x, y, z
 *                                                 per edge (3 * num_edges)
 *   build_neighborlists        src/api.rs:19-38  one system, optional (H, pbc)
 *   build_neighborlists_batch  src/api.rs:40-84  systems by runs of equal batch
 *                                                 ids, optional cells per system
 * H is given as in the reference: rows of the 3 x 3 array are the rows of H,
 * whose columns are the lattice vectors (api.rs:93-98).  Where the reference
 * returns Err(String), these functions throw neighborlist_rs::Error with the
 * same text.  `parallel` is accepted and does not change results.
 */
#ifndef NEIGHBORLIST_HPP
#define NEIGHBORLIST_HPP

#include <array>
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "nl_api.h"

namespace neighborlist_rs {

struct NeighborList {
  std::vector<int64_t> edge_index;  // [source_0 .. source_{M-1}, target_0 .. target_{M-1}]
  std::vector<int32_t> shifts;      // [x_0, y_0, z_0, x_1, ...]
};

using Mat3 = std::array<std::array<double, 3>, 3>;
using CellInfo = std::pair<Mat3, std::array<bool, 3>>;  // (H, pbc), api.rs:41

class Error : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

namespace detail {

inline NeighborList take(int rc, nl_result* r) {
  if (rc != 0) throw Error(nl_last_error());
  NeighborList out;
  const int64_t m = nl_result_num_edges(r, 0);
  out.edge_index.resize(2 * (size_t)m);
  out.shifts.resize(3 * (size_t)m);
  const int c = m > 0 ? nl_result_copy(r, 0, out.edge_index.data(), out.edge_index.data() + m,
                                       out.shifts.data())
                      : 0;
  nl_result_free(r);
  if (c != 0) throw Error(nl_last_error());
  return out;
}

inline void flatten(const Mat3& h, double* out) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) out[3 * r + c] = h[r][c];
}

}  // namespace detail

// src/api.rs:19-38
inline NeighborList build_neighborlists(const std::vector<std::array<double, 3>>& positions,
                                        double cutoff, const std::optional<CellInfo>& cell,
                                        bool parallel) {
  double h[9];
  uint8_t pbc[3] = {0, 0, 0};
  if (cell) {
    detail::flatten(cell->first, h);
    for (int k = 0; k < 3; ++k) pbc[k] = cell->second[k] ? 1 : 0;
  }
  nl_result* r = nullptr;
  const int rc = nl_search_single(positions.empty() ? nullptr : positions[0].data(),
                                  positions.size(), cell ? h : nullptr, cell ? pbc : nullptr,
                                  cutoff, parallel ? 1 : 0, &r);
  return detail::take(rc, r);
}

// src/api.rs:40-84: without cells every system is isolated (one None per
// run of equal batch ids, api.rs:50-70)
inline NeighborList build_neighborlists_batch(
    const std::vector<std::array<double, 3>>& positions, const std::vector<int32_t>& batch,
    double cutoff, const std::optional<std::vector<std::optional<CellInfo>>>& cells,
    bool parallel) {
  if (positions.size() != batch.size())
    throw Error("positions and batch must have the same length");  // batch.rs:19
  size_t n_systems = 0;
  if (!batch.empty()) {
    n_systems = 1;
    for (size_t i = 1; i < batch.size(); ++i)
      if (batch[i] != batch[i - 1]) ++n_systems;
  }
  const size_t n_cells = cells ? cells->size() : n_systems;
  std::vector<double> h(9 * n_cells, 0.0);
  std::vector<uint8_t> has(n_cells, 0), pbc(3 * n_cells, 0);
  if (cells) {
    for (size_t s = 0; s < n_cells; ++s) {
      const std::optional<CellInfo>& c = (*cells)[s];
      if (!c) continue;
      has[s] = 1;
      detail::flatten(c->first, h.data() + 9 * s);
      for (int k = 0; k < 3; ++k) pbc[3 * s + k] = c->second[k] ? 1 : 0;
    }
  }
  nl_result* r = nullptr;
  const int rc = nl_search_batch(positions.empty() ? nullptr : positions[0].data(),
                                 positions.size(), batch.empty() ? nullptr : batch.data(),
                                 n_cells, h.data(), has.data(), pbc.data(), cutoff,
                                 parallel ? 1 : 0, &r);
  return detail::take(rc, r);
}

}  // namespace neighborlist_rs

#endif  // NEIGHBORLIST_HPP
