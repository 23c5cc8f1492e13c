// The reference's Rust integration tests (tests/test_single_system.rs,
// tests/test_batch_system.rs, tests/test_api_structure.rs) restated against
// the C++ mirror of its API (include/neighborlist.hpp), plus the error texts
// of src/cell.rs:6, src/batch.rs:19 and src/batch.rs:41-45.  Prints one line
// per test; exits nonzero on the first failure.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "neighborlist.hpp"

using neighborlist_rs::build_neighborlists;
using neighborlist_rs::build_neighborlists_batch;
using neighborlist_rs::CellInfo;
using neighborlist_rs::NeighborList;

#define CHECK(cond)                                                       \
  do {                                                                    \
    if (!(cond)) {                                                        \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      std::exit(1);                                                       \
    }                                                                     \
  } while (0)

static const neighborlist_rs::Mat3 kBox10 = {{{10.0, 0.0, 0.0}, {0.0, 10.0, 0.0}, {0.0, 0.0, 10.0}}};

// tests/test_single_system.rs:3-21
static void test_build_neighborlists_simple() {
  const std::vector<std::array<double, 3>> positions = {{0.0, 0.0, 0.0}, {1.0, 0.0, 0.0}};
  const NeighborList r = build_neighborlists(positions, 1.5, std::nullopt, false);
  CHECK(r.edge_index.size() == 2);  // one edge, i < j
  CHECK(r.edge_index[0] == 0 && r.edge_index[1] == 1);
  CHECK(r.shifts.size() == 3);
  for (int32_t s : r.shifts) CHECK(s == 0);
}

// tests/test_single_system.rs:23-44
static void test_build_neighborlists_pbc() {
  const std::vector<std::array<double, 3>> positions = {{1.0, 1.0, 1.0}, {9.0, 1.0, 1.0}};
  const CellInfo cell{kBox10, {true, true, true}};
  const NeighborList r = build_neighborlists(positions, 2.5, cell, false);
  CHECK(r.edge_index.size() == 2);  // (0, 1) through the periodic image, distance 2
  CHECK(r.edge_index[0] == 0 && r.edge_index[1] == 1);
  CHECK(r.shifts.size() == 3);
  CHECK(r.shifts[0] == -1 && r.shifts[1] == 0 && r.shifts[2] == 0);
}

// tests/test_batch_system.rs:3-50
static void test_build_neighborlists_batch_mixed() {
  const std::vector<std::array<double, 3>> positions = {
      {0.0, 0.0, 0.0}, {1.0, 0.0, 0.0}, {1.0, 1.0, 1.0}, {9.0, 1.0, 1.0}};
  const std::vector<int32_t> batch = {0, 0, 1, 1};
  const std::vector<std::optional<CellInfo>> cells = {std::nullopt,
                                                      CellInfo{kBox10, {true, true, true}}};
  const NeighborList r = build_neighborlists_batch(positions, batch, 2.5, cells, true);
  const size_t m = r.edge_index.size() / 2;
  CHECK(m == 2);
  size_t i0 = m, i1 = m;
  for (size_t e = 0; e < m; ++e) {
    if (r.edge_index[e] == 0) i0 = e;
    if (r.edge_index[e] == 2) i1 = e;
  }
  CHECK(i0 < m && r.edge_index[m + i0] == 1);
  CHECK(r.shifts[3 * i0] == 0 && r.shifts[3 * i0 + 1] == 0 && r.shifts[3 * i0 + 2] == 0);
  CHECK(i1 < m && r.edge_index[m + i1] == 3);
  CHECK(r.shifts[3 * i1] == -1 && r.shifts[3 * i1 + 1] == 0 && r.shifts[3 * i1 + 2] == 0);
}

// tests/test_api_structure.rs:3-9
static void test_neighborlist_struct_exists() {
  std::optional<NeighborList> list;
  CHECK(!list.has_value());
}

// without cells every system of a batch is isolated (src/api.rs:50-70)
static void test_batch_without_cells() {
  const std::vector<std::array<double, 3>> positions = {
      {0.0, 0.0, 0.0}, {1.0, 0.0, 0.0}, {1.0, 1.0, 1.0}, {9.0, 1.0, 1.0}};
  const std::vector<int32_t> batch = {0, 0, 1, 1};
  const NeighborList r = build_neighborlists_batch(positions, batch, 2.5, std::nullopt, false);
  CHECK(r.edge_index.size() == 2);  // only (0, 1): 2 and 3 are 8 apart without images
  CHECK(r.edge_index[0] == 0 && r.edge_index[1] == 1);
}

static void expect_error(const char* what, const std::string& text, void (*f)()) {
  try {
    f();
  } catch (const neighborlist_rs::Error& e) {
    if (std::string(e.what()).find(text) != std::string::npos) return;
    std::fprintf(stderr, "FAILED %s: message '%s'\n", what, e.what());
    std::exit(1);
  }
  std::fprintf(stderr, "FAILED %s: no error\n", what);
  std::exit(1);
}

static void test_errors() {
  expect_error("singular cell", "Cell matrix is not invertible", [] {
    const neighborlist_rs::Mat3 h = {{{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 0.0}}};
    build_neighborlists({{0.0, 0.0, 0.0}}, 2.0, CellInfo{h, {true, true, true}}, false);
  });
  expect_error("batch length", "positions and batch must have the same length", [] {
    build_neighborlists_batch({{0.0, 0.0, 0.0}, {1.0, 0.0, 0.0}}, {0}, 2.0, std::nullopt, false);
  });
  expect_error("too few cells", "Expected at least 2 cells, but got 1", [] {
    const std::vector<std::optional<CellInfo>> cells = {std::nullopt};
    build_neighborlists_batch({{0.0, 0.0, 0.0}, {1.0, 0.0, 0.0}}, {0, 1}, 2.0, cells, false);
  });
}

int main() {
  struct {
    const char* name;
    void (*f)();
  } tests[] = {{"test_build_neighborlists_simple", test_build_neighborlists_simple},
               {"test_build_neighborlists_pbc", test_build_neighborlists_pbc},
               {"test_build_neighborlists_batch_mixed", test_build_neighborlists_batch_mixed},
               {"test_neighborlist_struct_exists", test_neighborlist_struct_exists},
               {"test_batch_without_cells", test_batch_without_cells},
               {"test_errors", test_errors}};
  for (const auto& t : tests) {
    t.f();
    std::printf("ok %s\n", t.name);
  }
  return 0;
}
