// nl_capi.cpp -- extern "C" layer of libnlmi355 (include/nl_api.h,
// include/nl_engine_ext.h).  Validation and segmentation follow the
// reference's dispatch layer; all searching happens in the engine.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/nl_api.h"
#include "../../include/nl_engine_ext.h"
#include "nl_engine.h"
#include "nl_server.h"
#include "nl_skin.h"

using nl::Engine;
using nl::Request;
using nl::Result;

struct nl_result {
  std::shared_ptr<Result> r;  // shared with exported (DLPack) views
};

struct nl_plan {
  nl::Plan p;
  nl::Engine* eng;
  int64_t cap;
};

struct nl_server {
  nl::Server* s;
};

struct nl_skin {
  nl::Skin k;
  nl::Engine* eng;
};

namespace {

int wrap_result(Result* r, nl_result** out) {
  *out = new nl_result{std::shared_ptr<Result>(r)};
  return 0;
}

int run(Request& req, nl_result** out) {
  if (!out) {
    nl::set_error("output pointer is NULL");
    return 1;
  }
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    nl::set_error("no HIP device available (libnlmi355 runs on the GPU only)");
    return 1;
  }
  Engine* e = nl::engine_for_current_device();
  Result* r = nullptr;
  int rc;
  {
    std::lock_guard<std::mutex> g(e->mutex());
    rc = e->run(req, &r);
  }
  if (rc != 0) return rc;
  return wrap_result(r, out);
}

bool valid_cutoff(double c) { return std::isfinite(c) && c > 0.0; }

// single.rs: one system, optional cell
int single_request(const double* pos, size_t n, const double* h, const uint8_t* pbc,
                   Request& req) {
  req.pos = pos;
  req.n = (int64_t)n;
  req.starts = {0, (int64_t)n};
  nl::SysIn s;
  if (h) {
    s.has_cell = 1;
    std::memcpy(s.h, h, sizeof s.h);
    for (int k = 0; k < 3; ++k) s.pbc[k] = pbc ? (pbc[k] != 0) : 1;
  }
  req.sys = {s};
  return 0;
}

// batch.rs:25-46: systems are maximal runs of equal ids; cells checked.
int batch_request(const double* pos, size_t n, const int32_t* batch, size_t n_cells,
                  const double* cells, const uint8_t* has_cell, const uint8_t* pbc,
                  Request& req) {
  req.pos = pos;
  req.n = (int64_t)n;
  req.starts.clear();
  req.starts.push_back(0);
  for (size_t i = 1; i < n; ++i)
    if (batch[i] != batch[i - 1]) req.starts.push_back((int64_t)i);
  req.starts.push_back((int64_t)n);
  const size_t ns = req.starts.size() - 1;
  if (has_cell && n_cells < ns) {
    nl::set_error("Expected at least " + std::to_string(ns) + " cells, but got " +
                  std::to_string(n_cells));
    return 1;
  }
  req.sys.assign(ns, nl::SysIn());
  if (has_cell) {
    for (size_t s = 0; s < ns; ++s) {
      if (!has_cell[s]) continue;
      if (!cells) {
        nl::set_error("cells is NULL while has_cell selects a cell");
        return 1;
      }
      req.sys[s].has_cell = 1;
      std::memcpy(req.sys[s].h, cells + 9 * s, sizeof(double) * 9);
      for (int k = 0; k < 3; ++k) req.sys[s].pbc[k] = pbc ? (pbc[3 * s + k] != 0) : 1;
    }
  }
  return 0;
}

int sizes_request(const double* pos, size_t n, const int64_t* sizes, size_t n_sys,
                  const double* cells, const uint8_t* has_cell, const uint8_t* pbc,
                  Request& req) {
  req.pos = pos;
  req.n = (int64_t)n;
  req.starts.assign(1, 0);
  int64_t acc = 0;
  for (size_t s = 0; s < n_sys; ++s) {
    if (sizes[s] <= 0) {
      nl::set_error("every system must hold at least one atom");
      return 1;
    }
    acc += sizes[s];
    req.starts.push_back(acc);
  }
  if (acc != (int64_t)n) {
    nl::set_error("system sizes must add up to the number of positions");
    return 1;
  }
  req.sys.assign(n_sys, nl::SysIn());
  if (has_cell) {
    for (size_t s = 0; s < n_sys; ++s) {
      if (!has_cell[s]) continue;
      req.sys[s].has_cell = 1;
      std::memcpy(req.sys[s].h, cells + 9 * s, sizeof(double) * 9);
      for (int k = 0; k < 3; ++k) req.sys[s].pbc[k] = pbc ? (pbc[3 * s + k] != 0) : 1;
    }
  }
  return 0;
}

int empty_result(int k, nl_result** out) {
  Result* r = new Result();
  r->eng = nullptr;
  r->k = k;
  r->len.assign(k, 0);
  r->base.assign(k, 0);
  return wrap_result(r, out);
}

int check_cutoffs(const double* c, size_t k, const char* msg) {
  for (size_t q = 0; q < k; ++q)
    if (!valid_cutoff(c[q])) {
      nl::set_error(msg);
      return 1;
    }
  return 0;
}

}  // namespace

namespace nl {
// the request builder and cutoff check of the entry points above, for the
// allocation-free entry point (nl_noalloc.cpp)
int capi_sizes_request(const double* pos, size_t n, const int64_t* sizes, size_t n_sys,
                       const double* cells, const uint8_t* has_cell, const uint8_t* pbc,
                       Request& req) {
  return sizes_request(pos, n, sizes, n_sys, cells, has_cell, pbc, req);
}
int capi_check_cutoffs(const double* c, size_t k, const char* msg) {
  return check_cutoffs(c, k, msg);
}
}  // namespace nl

extern "C" {

int nl_search_single(const double* positions, size_t n, const double* h, const uint8_t* pbc,
                     double cutoff, int parallel, nl_result** out) {
  (void)parallel;
  if (check_cutoffs(&cutoff, 1, "Cutoff must be positive and finite")) return 1;
  if (n == 0) return empty_result(1, out);  // single.rs:14-17
  Request req;
  single_request(positions, n, h, pbc, req);
  req.cutoffs = {cutoff};
  return run(req, out);
}

int nl_search_single_multi(const double* positions, size_t n, const double* h,
                           const uint8_t* pbc, const double* cutoffs, size_t n_cutoffs,
                           int disjoint, nl_result** out) {
  if (check_cutoffs(cutoffs, n_cutoffs, "All cutoffs must be positive and finite")) return 1;
  if (n == 0) return empty_result((int)n_cutoffs, out);  // single.rs:99-101
  if (n_cutoffs == 0) return empty_result(0, out);       // single.rs:102-104
  Request req;
  single_request(positions, n, h, pbc, req);
  req.cutoffs.assign(cutoffs, cutoffs + n_cutoffs);
  req.multi = true;
  req.disjoint = disjoint != 0;
  return run(req, out);
}

int nl_search_batch(const double* positions, size_t n, const int32_t* batch, size_t n_cells,
                    const double* cells, const uint8_t* has_cell, const uint8_t* pbc,
                    double cutoff, int parallel, nl_result** out) {
  (void)parallel;
  if (check_cutoffs(&cutoff, 1, "Cutoff must be positive and finite")) return 1;
  if (n == 0) return empty_result(1, out);  // batch.rs:21-23
  Request req;
  if (batch_request(positions, n, batch, n_cells, cells, has_cell, pbc, req)) return 1;
  req.cutoffs = {cutoff};
  return run(req, out);
}

int nl_search_batch_multi(const double* positions, size_t n, const int32_t* batch,
                          size_t n_cells, const double* cells, const uint8_t* has_cell,
                          const uint8_t* pbc, const double* cutoffs, size_t n_cutoffs,
                          int disjoint, nl_result** out) {
  if (check_cutoffs(cutoffs, n_cutoffs, "All cutoffs must be positive and finite")) return 1;
  if (n == 0 || n_cutoffs == 0) return empty_result((int)n_cutoffs, out);  // batch.rs:160-166
  Request req;
  if (batch_request(positions, n, batch, n_cells, cells, has_cell, pbc, req)) return 1;
  req.cutoffs.assign(cutoffs, cutoffs + n_cutoffs);
  req.multi = true;
  req.disjoint = disjoint != 0;
  return run(req, out);
}

static int search_opts(const double* positions, size_t n, const int64_t* sys_sizes, size_t n_sys,
                       const double* cells, const uint8_t* has_cell, const uint8_t* pbc,
                       const double* cutoffs, size_t n_cutoffs, int disjoint, int multi,
                       unsigned options, bool has_stream, void* stream, nl_result** out) {
  if (check_cutoffs(cutoffs, n_cutoffs, "All cutoffs must be positive and finite")) return 1;
  if (n == 0 || n_cutoffs == 0) return empty_result((int)n_cutoffs, out);
  Request req;
  if (sizes_request(positions, n, sys_sizes, n_sys, cells, has_cell, pbc, req)) return 1;
  req.pos_device = (options & NL_DEVICE_POSITIONS) != 0;
  req.cutoffs.assign(cutoffs, cutoffs + n_cutoffs);
  req.multi = multi != 0;
  req.disjoint = disjoint != 0;
  req.global_index = (options & NL_LOCAL_INDEX) == 0;
  req.want_ranges = (options & NL_RANGES) != 0;
  req.want_dist = (options & NL_DIST) != 0;
  req.want_disp = (options & NL_DISP) != 0;
  req.full = (options & NL_FULL) != 0;
  req.has_ustream = has_stream;
  req.ustream = (hipStream_t)stream;
  return run(req, out);
}

int nl_search_ex(const double* positions, size_t n, const int64_t* sys_sizes, size_t n_sys,
                 const double* cells, const uint8_t* has_cell, const uint8_t* pbc,
                 const double* cutoffs, size_t n_cutoffs, int disjoint, int multi,
                 unsigned options, void* stream, nl_result** out) {
  return search_opts(positions, n, sys_sizes, n_sys, cells, has_cell, pbc, cutoffs, n_cutoffs,
                     disjoint, multi, options, true, stream, out);
}

static unsigned legacy_options(bool device, int global_index, int want_ranges, int want_dist) {
  return (device ? NL_DEVICE_POSITIONS : 0u) | (global_index ? 0u : NL_LOCAL_INDEX) |
         (want_ranges ? NL_RANGES : 0u) | (want_dist ? NL_DIST : 0u);
}

int nl_search_batch_device(const double* d_positions, size_t n, const int64_t* sys_sizes,
                           size_t n_sys, const double* cells, const uint8_t* has_cell,
                           const uint8_t* pbc, const double* cutoffs, size_t n_cutoffs,
                           int disjoint, int multi, int global_index, int want_ranges,
                           int want_dist, nl_result** out) {
  return search_opts(d_positions, n, sys_sizes, n_sys, cells, has_cell, pbc, cutoffs, n_cutoffs,
                     disjoint, multi, legacy_options(true, global_index, want_ranges, want_dist),
                     false, nullptr, out);
}

int nl_search_systems(const double* positions, size_t n, const int64_t* sys_sizes,
                      size_t n_sys, const double* cells, const uint8_t* has_cell,
                      const uint8_t* pbc, const double* cutoffs, size_t n_cutoffs, int disjoint,
                      int multi, int global_index, int want_ranges, int want_dist,
                      nl_result** out) {
  return search_opts(positions, n, sys_sizes, n_sys, cells, has_cell, pbc, cutoffs, n_cutoffs,
                     disjoint, multi, legacy_options(false, global_index, want_ranges, want_dist),
                     false, nullptr, out);
}

size_t nl_result_num_lists(const nl_result* r) { return r ? (size_t)r->r->k : 0; }

int64_t nl_result_num_edges(const nl_result* r, size_t list) {
  if (!r || list >= (size_t)r->r->k) return -1;
  return r->r->len[list];
}

int nl_result_copy(const nl_result* r, size_t list, int64_t* edge_i, int64_t* edge_j,
                   int32_t* shifts) {
  if (!r || list >= (size_t)r->r->k) {
    nl::set_error("invalid result or list index");
    return 1;
  }
  const Result* R = r->r.get();
  const int64_t m = R->len[list];
  if (m == 0) return 0;
  const int64_t b = R->base[list];
  if (R->mirror) {
    // page-locked copy made with the search (same layout as the device arrays)
    const int64_t* e = (const int64_t*)R->mirror;
    const int32_t* sh = (const int32_t*)((const char*)R->mirror + 16 * R->mirror_cap);
    if (edge_i) std::memcpy(edge_i, e + 2 * b, 8 * m);
    if (edge_j) std::memcpy(edge_j, e + 2 * b + m, 8 * m);
    if (shifts) std::memcpy(shifts, sh + 3 * b, 12 * m);
    return 0;
  }
  hipStream_t st = R->eng->stream();
  const int64_t* e = (const int64_t*)R->edges;
  hipError_t err = hipSuccess;
  if (edge_i && edge_j == edge_i + m) {
    // the (2, M) layout on both sides: one copy
    err = hipMemcpyAsync(edge_i, e + 2 * b, 16 * m, hipMemcpyDeviceToHost, st);
    edge_i = edge_j = nullptr;
  }
  if (edge_i && err == hipSuccess)
    err = hipMemcpyAsync(edge_i, e + 2 * b, 8 * m, hipMemcpyDeviceToHost, st);
  if (edge_j && err == hipSuccess)
    err = hipMemcpyAsync(edge_j, e + 2 * b + m, 8 * m, hipMemcpyDeviceToHost, st);
  if (shifts && err == hipSuccess)
    err = hipMemcpyAsync(shifts, (const int32_t*)R->shifts + 3 * b, 12 * m,
                         hipMemcpyDeviceToHost, st);
  if (err == hipSuccess) err = hipStreamSynchronize(st);
  if (err != hipSuccess) {
    nl::set_error(std::string("HIP error in nl_result_copy: ") + hipGetErrorString(err));
    return 1;
  }
  return 0;
}

int nl_result_copy_dist(const nl_result* r, size_t list, double* dist) {
  if (!r || list >= (size_t)r->r->k || !r->r->dist) {
    nl::set_error("no distances recorded for this result");
    return 1;
  }
  const Result* R = r->r.get();
  const int64_t m = R->len[list];
  if (m == 0) return 0;
  hipStream_t st = R->eng->stream();
  hipError_t err = hipMemcpyAsync(dist, (const double*)R->dist + R->base[list], 8 * m,
                                  hipMemcpyDeviceToHost, st);
  if (err == hipSuccess) err = hipStreamSynchronize(st);
  if (err != hipSuccess) {
    nl::set_error(std::string("HIP error in nl_result_copy_dist: ") + hipGetErrorString(err));
    return 1;
  }
  return 0;
}

int nl_result_copy_disp(const nl_result* r, size_t list, double* disp) {
  if (!r || list >= (size_t)r->r->k || !r->r->disp) {
    nl::set_error("no displacements recorded for this result");
    return 1;
  }
  const Result* R = r->r.get();
  const int64_t m = R->len[list];
  if (m == 0) return 0;
  hipStream_t st = R->eng->stream();
  hipError_t err = hipMemcpyAsync(disp, (const double*)R->disp + 3 * R->base[list], 24 * m,
                                  hipMemcpyDeviceToHost, st);
  if (err == hipSuccess) err = hipStreamSynchronize(st);
  if (err != hipSuccess) {
    nl::set_error(std::string("HIP error in nl_result_copy_disp: ") + hipGetErrorString(err));
    return 1;
  }
  return 0;
}

int nl_result_device(const nl_result* r, size_t list, const int64_t** edge_index,
                     const int32_t** shifts) {
  if (!r || list >= (size_t)r->r->k) {
    nl::set_error("invalid result or list index");
    return 1;
  }
  const Result* R = r->r.get();
  const int64_t b = R->base[list];
  if (edge_index) *edge_index = R->edges ? (const int64_t*)R->edges + 2 * b : nullptr;
  if (shifts) *shifts = R->shifts ? (const int32_t*)R->shifts + 3 * b : nullptr;
  return 0;
}

int nl_result_system_ranges(const nl_result* r, int64_t* ranges, int32_t* modes) {
  if (!r) {
    nl::set_error("invalid result");
    return 1;
  }
  const Result* R = r->r.get();
  if (R->ranges.empty() && R->total > 0) {
    nl::set_error("system ranges were not requested for this result");
    return 1;
  }
  if (ranges && !R->ranges.empty())
    std::memcpy(ranges, R->ranges.data(), 8 * R->ranges.size());
  if (modes && !R->modes.empty()) std::memcpy(modes, R->modes.data(), 4 * R->modes.size());
  return 0;
}

void nl_result_free(nl_result* r) { delete r; }

// ---- DLPack export ---------------------------------------------------------
// The DLPack ABI (DLManagedTensor of dlpack.h, v0.8): laid out here field for
// field, as every producer does, since consumers bind to exactly this layout.
namespace {
struct DLDevice_ {
  int32_t device_type;  // kDLROCM = 10
  int32_t device_id;
};
struct DLDataType_ {
  uint8_t code;  // kDLInt = 0, kDLFloat = 2
  uint8_t bits;
  uint16_t lanes;
};
struct DLTensor_ {
  void* data;
  DLDevice_ device;
  int32_t ndim;
  DLDataType_ dtype;
  int64_t* shape;
  int64_t* strides;
  uint64_t byte_offset;
};
struct DLManagedTensor_ {
  DLTensor_ dl_tensor;
  void* manager_ctx;
  void (*deleter)(DLManagedTensor_*);
};
struct DlView {
  std::shared_ptr<Result> r;
  int64_t shape[2];
  int64_t strides[2];
  DLManagedTensor_ m;
};
void dl_release(DLManagedTensor_* m) {
  DlView* v = (DlView*)m->manager_ctx;
  v->r->fence_on_consumer();
  delete v;
}
}  // namespace

int nl_result_dlpack(const nl_result* r, size_t list, int what, void** out) {
  if (!out) {
    nl::set_error("output pointer is NULL");
    return 1;
  }
  *out = nullptr;
  if (!r || list >= (size_t)r->r->k || what < 0 || what > 3) {
    nl::set_error("invalid result, list index or array selector");
    return 1;
  }
  const Result* R = r->r.get();
  if (!R->eng) {
    nl::set_error("an empty result has no device arrays");
    return 1;
  }
  if (what == 2 && !R->dist) {
    nl::set_error("no distances recorded for this result");
    return 1;
  }
  if (what == 3 && !R->disp) {
    nl::set_error("no displacements recorded for this result");
    return 1;
  }
  const int64_t m = R->len[list], b = R->base[list];
  DlView* v = new DlView();
  v->r = r->r;
  DLTensor_& t = v->m.dl_tensor;
  t.device = DLDevice_{10, R->eng->device()};
  t.byte_offset = 0;
  t.shape = v->shape;
  t.strides = v->strides;
  if (what == 0) {  // edge_index (2, M) int64: sources, then targets
    t.data = (int64_t*)R->edges + 2 * b;
    t.ndim = 2;
    t.dtype = DLDataType_{0, 64, 1};
    v->shape[0] = 2;
    v->shape[1] = m;
    v->strides[0] = m;
    v->strides[1] = 1;
  } else if (what == 1) {  // shift (M, 3) int32
    t.data = (int32_t*)R->shifts + 3 * b;
    t.ndim = 2;
    t.dtype = DLDataType_{0, 32, 1};
    v->shape[0] = m;
    v->shape[1] = 3;
    v->strides[0] = 3;
    v->strides[1] = 1;
  } else if (what == 2) {  // distance (M,) float64
    t.data = (double*)R->dist + b;
    t.ndim = 1;
    t.dtype = DLDataType_{2, 64, 1};
    v->shape[0] = m;
    v->strides[0] = 1;
  } else {  // displacement (M, 3) float64
    t.data = (double*)R->disp + 3 * b;
    t.ndim = 2;
    t.dtype = DLDataType_{2, 64, 1};
    v->shape[0] = m;
    v->shape[1] = 3;
    v->strides[0] = 3;
    v->strides[1] = 1;
  }
  v->m.manager_ctx = v;
  v->m.deleter = dl_release;
  *out = &v->m;
  return 0;
}

size_t nl_get_brute_force_threshold(void) { return nl::config().bf_threshold.load(); }
void nl_set_brute_force_threshold(size_t v) { nl::config().bf_threshold.store(v); }
size_t nl_get_parallel_threshold(void) { return nl::config().par_threshold.load(); }
void nl_set_parallel_threshold(size_t v) { nl::config().par_threshold.store(v); }
size_t nl_get_stack_threshold(void) { return nl::config().stack_threshold.load(); }
void nl_set_stack_threshold(size_t v) {
  nl::config().stack_threshold.store(v < 1024 ? v : 1024);  // MAX_STACK_SIZE (config.rs:4)
}

const char* nl_last_error(void) { return nl::last_error(); }
const char* nl_version(void) { return "nlmi355 0.1.0 (gfx950)"; }

void nl_set_timing(int on) { nl::engine_for_current_device()->set_timing(on != 0); }

int nl_last_timing(float* build_ms, float* count_ms, float* fill_ms, float* total_ms,
                   int64_t* edges, int64_t* bins) {
  nl::StageTimes t = nl::engine_for_current_device()->last_times();
  if (build_ms) *build_ms = t.build;
  if (count_ms) *count_ms = t.count;
  if (fill_ms) *fill_ms = t.fill;
  if (total_ms) *total_ms = t.total;
  if (edges) *edges = t.edges;
  if (bins) *bins = t.bins;
  return 0;
}

void nl_set_cell_div(int d) { nl::config().cell_div.store(d < 0 ? 0 : (d > 4 ? 4 : d)); }
int nl_get_cell_div(void) { return nl::config().cell_div.load(); }
void nl_set_small_path(int mode) {
  nl::config().small_path.store(mode < 0 ? 0 : (mode > 3 ? 3 : mode));
}
int nl_get_small_path(void) { return nl::config().small_path.load(); }

void nl_set_bin_order(int morton) { nl::config().bin_order = morton ? 1 : 0; }
void nl_set_fused_build(int on) { nl::config().fused_build = on == 2 ? 2 : (on ? 1 : 0); }
int nl_get_fused_build(void) { return nl::config().fused_build.load(); }
int nl_get_bin_order(void) { return nl::config().bin_order.load(); }
void nl_set_fp32_filter(int on) { nl::config().use_fp32.store(on != 0); }
int nl_get_fp32_filter(void) { return nl::config().use_fp32.load(); }
void nl_set_work_split(int group, int parts) {
  nl::config().group.store(group < 1 ? 1 : (group > 64 ? 64 : group));
  nl::config().parts.store(parts < 1 ? 1 : (parts > 16 ? 16 : parts));
}
int nl_get_work_split(int* group, int* parts) {
  if (group) *group = nl::config().group.load();
  if (parts) *parts = nl::config().parts.load();
  return 0;
}

void nl_set_autotune(int on) { nl::config().autotune.store(on != 0); }
int nl_get_autotune(void) { return nl::config().autotune.load(); }
void nl_clear_tuning(void) { nl::engine_for_current_device()->clear_tuning(); }
size_t nl_tune_report(char* buf, size_t len) {
  const std::string r = nl::engine_for_current_device()->tune_report();
  if (buf && len) {
    const size_t m = r.size() < len - 1 ? r.size() : len - 1;
    std::memcpy(buf, r.data(), m);
    buf[m] = 0;
  }
  return r.size();
}

// ---- skin lists for molecular dynamics --------------------------------------
int nl_skin_create(const double* d_positions, size_t n, const int64_t* sys_sizes, size_t n_sys,
                   const double* cells, const uint8_t* has_cell, const uint8_t* pbc,
                   const double* cutoffs, const double* skins, size_t n_cutoffs, int disjoint,
                   int multi, unsigned options, void* stream, nl_skin** out) {
  if (!out) {
    nl::set_error("output pointer is NULL");
    return 1;
  }
  *out = nullptr;
  if (check_cutoffs(cutoffs, n_cutoffs, "All cutoffs must be positive and finite")) return 1;
  if (n == 0 || n_cutoffs == 0) {
    nl::set_error("skin lists need at least one atom and one cutoff");
    return 1;
  }
  if (!skins) {
    nl::set_error("skins is NULL");
    return 1;
  }
  if (options & ~(unsigned)(NL_DEVICE_POSITIONS | NL_LOCAL_INDEX | NL_FULL)) {
    nl::set_error("skin lists take NL_LOCAL_INDEX and NL_FULL only");
    return 1;
  }
  if (disjoint) {
    // a list at rc_q + skin_q holds every pair now within rc_q only when it
    // is cumulative: a disjoint shell misses pairs that were inside the
    // previous shell at the build and have moved out since
    nl::set_error("skin lists are cumulative (disjoint shells cannot keep every pair within "
                  "rc in its shell between rebuilds); split the cumulative lists by distance "
                  "instead");
    return 1;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    nl::set_error("no HIP device available (libnlmi355 runs on the GPU only)");
    return 1;
  }
  Request req;
  if (sizes_request(d_positions, n, sys_sizes, n_sys, cells, has_cell, pbc, req)) return 1;
  req.pos_device = true;
  req.cutoffs.assign(cutoffs, cutoffs + n_cutoffs);
  req.multi = multi != 0;
  req.disjoint = disjoint != 0;
  req.global_index = (options & NL_LOCAL_INDEX) == 0;
  req.full = (options & NL_FULL) != 0;
  nl_skin* k = new nl_skin();
  k->eng = nl::engine_for_current_device();
  int rc;
  {
    std::lock_guard<std::mutex> g(k->eng->mutex());
    rc = k->k.create(k->eng, req, std::vector<double>(skins, skins + n_cutoffs),
                     (hipStream_t)stream);
  }
  if (rc != 0) {
    delete k;
    return rc;
  }
  *out = k;
  return 0;
}

int nl_skin_update(nl_skin* s, const double* d_positions, const double* cells, void* stream,
                   int* rebuilt) {
  if (!s || !d_positions) {
    nl::set_error("invalid skin lists or positions");
    return 1;
  }
  bool r = false;
  int rc;
  {
    std::lock_guard<std::mutex> g(s->eng->mutex());
    rc = s->k.update(d_positions, cells, (hipStream_t)stream, &r);
  }
  if (rebuilt) *rebuilt = r ? 1 : 0;
  return rc;
}

int nl_skin_check(nl_skin* s, const double* d_positions, void* stream, double* max_disp,
                  int* needs_rebuild) {
  if (!s || !d_positions) {
    nl::set_error("invalid skin lists or positions");
    return 1;
  }
  bool needs = false;
  const int rc = s->k.check(d_positions, (hipStream_t)stream, max_disp, &needs);
  if (needs_rebuild) *needs_rebuild = needs ? 1 : 0;
  return rc;
}

int64_t nl_skin_num_edges(const nl_skin* s, size_t list) {
  if (!s || list >= (size_t)s->k.lists()) return -1;
  return s->k.len((int)list);
}

int nl_skin_device(const nl_skin* s, size_t list, const int64_t** edge_index,
                   const int32_t** shifts) {
  if (!s || list >= (size_t)s->k.lists()) {
    nl::set_error("invalid skin lists or list index");
    return 1;
  }
  if (edge_index) *edge_index = s->k.edges((int)list);
  if (shifts) *shifts = s->k.shifts((int)list);
  return 0;
}

int nl_skin_stats(const nl_skin* s, int64_t* updates, int64_t* rebuilds, double* last_disp) {
  if (!s) {
    nl::set_error("invalid skin lists");
    return 1;
  }
  if (updates) *updates = s->k.updates;
  if (rebuilds) *rebuilds = s->k.rebuilds;
  if (last_disp) *last_disp = s->k.last_disp;
  return 0;
}

void nl_skin_free(nl_skin* s) { delete s; }

static int plan_create_opts(const double* d_positions, size_t n, const int64_t* sys_sizes,
                            size_t n_sys, const double* cells, const uint8_t* has_cell,
                            const uint8_t* pbc, const double* cutoffs, size_t n_cutoffs,
                            int disjoint, int multi, unsigned options, int64_t capacity,
                            nl_plan** out) {
  if (!out) {
    nl::set_error("output pointer is NULL");
    return 1;
  }
  *out = nullptr;
  if (check_cutoffs(cutoffs, n_cutoffs, "All cutoffs must be positive and finite")) return 1;
  if (n == 0 || n_cutoffs == 0) {
    nl::set_error("a plan needs at least one atom and one cutoff");
    return 1;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    nl::set_error("no HIP device available (libnlmi355 runs on the GPU only)");
    return 1;
  }
  Request req;
  if (sizes_request(d_positions, n, sys_sizes, n_sys, cells, has_cell, pbc, req)) return 1;
  req.pos_device = true;
  req.cutoffs.assign(cutoffs, cutoffs + n_cutoffs);
  req.multi = multi != 0;
  req.disjoint = disjoint != 0;
  req.global_index = (options & NL_LOCAL_INDEX) == 0;
  req.want_ranges = (options & NL_RANGES) != 0;
  req.want_dist = (options & NL_DIST) != 0;
  req.want_disp = (options & NL_DISP) != 0;
  req.full = (options & NL_FULL) != 0;
  nl_plan* p = new nl_plan();
  p->eng = nl::engine_for_current_device();
  int rc;
  {
    std::lock_guard<std::mutex> g(p->eng->mutex());
    rc = p->p.create(p->eng, req, capacity);
  }
  if (rc != 0) {
    delete p;
    return rc;
  }
  *out = p;
  return 0;
}

int nl_plan_create(const double* d_positions, size_t n, const int64_t* sys_sizes, size_t n_sys,
                   const double* cells, const uint8_t* has_cell, const uint8_t* pbc,
                   const double* cutoffs, size_t n_cutoffs, int disjoint, int multi,
                   int global_index, int64_t capacity, nl_plan** out) {
  return plan_create_opts(d_positions, n, sys_sizes, n_sys, cells, has_cell, pbc, cutoffs,
                          n_cutoffs, disjoint, multi, global_index ? 0u : NL_LOCAL_INDEX,
                          capacity, out);
}

int nl_plan_create_ex(const double* d_positions, size_t n, const int64_t* sys_sizes,
                      size_t n_sys, const double* cells, const uint8_t* has_cell,
                      const uint8_t* pbc, const double* cutoffs, size_t n_cutoffs, int disjoint,
                      int multi, unsigned options, int64_t capacity, nl_plan** out) {
  return plan_create_opts(d_positions, n, sys_sizes, n_sys, cells, has_cell, pbc, cutoffs,
                          n_cutoffs, disjoint, multi, options, capacity, out);
}

int nl_plan_array(const nl_plan* p, size_t list, int what, const void** ptr) {
  if (!p || list >= (size_t)p->p.pipe().K || !ptr) {
    nl::set_error("invalid plan, list index or output pointer");
    return 1;
  }
  const int64_t b = p->p.base((int)list);
  switch (what) {
    case 0:
      *ptr = (const int64_t*)p->p.edges() + 2 * b;
      return 0;
    case 1:
      *ptr = (const int32_t*)p->p.shifts() + 3 * b;
      return 0;
    case 2:
      if (!p->p.dist()) {
        nl::set_error("distances were not requested (NL_DIST)");
        return 1;
      }
      *ptr = (const double*)p->p.dist() + b;
      return 0;
    case 3:
      if (!p->p.disp()) {
        nl::set_error("displacements were not requested (NL_DISP)");
        return 1;
      }
      *ptr = (const double*)p->p.disp() + 3 * b;
      return 0;
    default:
      nl::set_error("unknown array (0 edge_index, 1 shift, 2 distance, 3 displacement)");
      return 1;
  }
}

int nl_plan_system_ranges(const nl_plan* p, int64_t* ranges, int32_t* modes) {
  if (!p) {
    nl::set_error("invalid plan");
    return 1;
  }
  if (!p->p.has_ranges()) {
    nl::set_error("per-system ranges were not requested (NL_RANGES)");
    return 1;
  }
  if (p->p.ranges().empty()) {
    nl::set_error("the plan has not run yet");
    return 1;
  }
  if (ranges) std::copy(p->p.ranges().begin(), p->p.ranges().end(), ranges);
  if (modes) std::copy(p->p.modes().begin(), p->p.modes().end(), modes);
  return 0;
}

int nl_plan_run(nl_plan* p) {
  if (!p) {
    nl::set_error("invalid plan");
    return 1;
  }
  return p->p.run();
}

int64_t nl_plan_num_edges(const nl_plan* p, size_t list) {
  if (!p || list >= (size_t)p->p.pipe().K) return -1;
  return p->p.len((int)list);
}

int nl_plan_device(const nl_plan* p, size_t list, const int64_t** edge_index,
                   const int32_t** shifts) {
  if (!p || list >= (size_t)p->p.pipe().K) {
    nl::set_error("invalid plan or list index");
    return 1;
  }
  const int64_t b = p->p.base((int)list);
  if (edge_index) *edge_index = (const int64_t*)p->p.edges() + 2 * b;
  if (shifts) *shifts = (const int32_t*)p->p.shifts() + 3 * b;
  return 0;
}

int nl_plan_copy(const nl_plan* p, size_t list, int64_t* edge_i, int64_t* edge_j,
                 int32_t* shifts) {
  if (!p || list >= (size_t)p->p.pipe().K) {
    nl::set_error("invalid plan or list index");
    return 1;
  }
  const int64_t m = p->p.len((int)list), b = p->p.base((int)list);
  if (m == 0) return 0;
  hipStream_t st = p->p.stream();
  const int64_t* e = (const int64_t*)p->p.edges();
  hipError_t err = hipSuccess;
  if (edge_i) err = hipMemcpyAsync(edge_i, e + 2 * b, 8 * m, hipMemcpyDeviceToHost, st);
  if (edge_j && err == hipSuccess)
    err = hipMemcpyAsync(edge_j, e + 2 * b + m, 8 * m, hipMemcpyDeviceToHost, st);
  if (shifts && err == hipSuccess)
    err = hipMemcpyAsync(shifts, (const int32_t*)p->p.shifts() + 3 * b, 12 * m,
                         hipMemcpyDeviceToHost, st);
  if (err == hipSuccess) err = hipStreamSynchronize(st);
  if (err != hipSuccess) {
    nl::set_error(std::string("HIP error in nl_plan_copy: ") + hipGetErrorString(err));
    return 1;
  }
  return 0;
}

int nl_plan_stats(const nl_plan* p, int64_t* replays, int64_t* recaptures, int64_t* capacity) {
  if (!p) {
    nl::set_error("invalid plan");
    return 1;
  }
  if (replays) *replays = p->p.replays;
  if (recaptures) *recaptures = p->p.recaptures;
  if (capacity) *capacity = p->p.capacity();
  return 0;
}

void nl_plan_free(nl_plan* p) { delete p; }

int nl_cell_inverse(const double* h, double* hinv) {
  if (!nl::inverse3(h, hinv)) {
    nl::set_error("Cell matrix is not invertible");  // cell.rs:6
    return 1;
  }
  return 0;
}

int nl_cell_wrap(const double* h, const uint8_t* pbc, const double* pos, double* out) {
  double inv[9];
  if (nl_cell_inverse(h, inv)) return 1;
  double f[3];
  nl::matvec3(inv, pos[0], pos[1], pos[2], f);
  for (int k = 0; k < 3; ++k)
    if (!pbc || pbc[k]) f[k] = f[k] - std::floor(f[k]);
  nl::matvec3(h, f[0], f[1], f[2], out);
  return 0;
}

void* nl_device_malloc(size_t bytes) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    nl::set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return nullptr;
  }
  return p;
}

static int hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  nl::set_error(std::string(what) + ": " + hipGetErrorString(e));
  return 1;
}

int nl_device_free(void* p) { return hip_status(hipFree(p), "hipFree"); }
int nl_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}
int nl_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}
int nl_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
}
int nl_device_synchronize(void) { return hip_status(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

// Pinned host blocks for lists delivered to host arrays: a device-to-host
// copy into page-locked memory runs at the link's DMA rate, into pageable
// memory through the runtime's staging buffer at a fraction of it.  Blocks
// come in power-of-two sizes from 64 KiB; freed blocks are kept for reuse up
// to kPinnedCache bytes, and at most kPinnedLive bytes are handed out at once
// (beyond it nl_host_alloc returns NULL and callers use pageable memory).
namespace {
constexpr size_t kPinnedCache = size_t(1) << 30;
constexpr size_t kPinnedLive = size_t(8) << 30;
struct PinnedPool {
  std::mutex mu;
  std::multimap<size_t, void*> idle;           // size -> block
  std::unordered_map<void*, size_t> size_of;  // every block handed out
  size_t cached = 0, live = 0;
};
PinnedPool& pinned_pool() {
  static PinnedPool* pool = new PinnedPool();  // lives until the process ends
  return *pool;
}
}  // namespace

void* nl_host_alloc(size_t bytes) {
  if (bytes == 0) return nullptr;
  size_t cls = size_t(1) << 16;
  while (cls < bytes) cls <<= 1;
  PinnedPool& P = pinned_pool();
  std::lock_guard<std::mutex> g(P.mu);
  // the most recently freed block of the class first
  auto it = P.idle.upper_bound(cls);
  void* p = nullptr;
  if (it != P.idle.begin() && std::prev(it)->first == cls) {
    it = std::prev(it);
    p = it->second;
    P.idle.erase(it);
    P.cached -= cls;
  } else {
    if (P.live + cls > kPinnedLive) return nullptr;
    if (hipHostMalloc(&p, cls, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  P.size_of[p] = cls;
  P.live += cls;
  return p;
}

void nl_host_free(void* p) {
  if (!p) return;
  PinnedPool& P = pinned_pool();
  std::lock_guard<std::mutex> g(P.mu);
  auto it = P.size_of.find(p);
  if (it == P.size_of.end()) return;  // not a block of this pool
  const size_t cls = it->second;
  P.size_of.erase(it);
  P.live -= cls;
  if (P.cached + cls <= kPinnedCache) {
    P.idle.emplace(cls, p);
    P.cached += cls;
  } else {
    (void)hipHostFree(p);
  }
}

nl_server* nl_server_create(int slots, int64_t max_atoms, int64_t max_systems,
                            int64_t edge_budget) {
  return nl_server_create_ex(slots, max_atoms, max_systems, edge_budget, 0u);
}

nl_server* nl_server_create_ex(int slots, int64_t max_atoms, int64_t max_systems,
                               int64_t edge_budget, unsigned options) {
  if (options & ~(NL_FULL | NL_DIST | NL_DISP)) {
    nl::set_error("server options: NL_FULL, NL_DIST and NL_DISP only");
    return nullptr;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    nl::set_error("no HIP device available (libnlmi355 runs on the GPU only)");
    return nullptr;
  }
  if (slots <= 0 || max_atoms <= 0 || max_systems <= 0 || edge_budget <= 0) {
    nl::set_error("server limits must be positive");
    return nullptr;
  }
  return new nl_server{new nl::Server(nl::engine_for_current_device(), slots, max_atoms,
                                      max_systems, edge_budget, (options & NL_FULL) != 0,
                                      (options & NL_DIST) != 0, (options & NL_DISP) != 0)};
}

int64_t nl_server_submit(nl_server* s, const double* positions, size_t n, const double* cell,
                         const uint8_t* pbc, double cutoff) {
  return nl_server_submit_ex(s, positions, n, cell, pbc, cutoff, 0u);
}

int64_t nl_server_submit_ex(nl_server* s, const double* positions, size_t n, const double* cell,
                            const uint8_t* pbc, double cutoff, unsigned flags) {
  if (!s || !positions) {
    nl::set_error("server or positions pointer is NULL");
    return -1;
  }
  if (flags & ~NL_SUBMIT_BORROW) {
    nl::set_error("submit flags: NL_SUBMIT_BORROW only");
    return -1;
  }
  return s->s->submit(positions, (int64_t)n, cell, pbc, cutoff, (flags & NL_SUBMIT_BORROW) != 0);
}

int nl_server_wait(nl_server* s, int64_t id, double timeout_s) {
  if (!s) {
    nl::set_error("server pointer is NULL");
    return -1;
  }
  return s->s->wait(id, timeout_s);
}

int nl_server_result_info(nl_server* s, int64_t id, int64_t* n_edges, double* latency_s,
                          int* batch_size) {
  if (!s) {
    nl::set_error("server pointer is NULL");
    return -1;
  }
  return s->s->info(id, n_edges, latency_s, batch_size);
}

int nl_server_take(nl_server* s, int64_t id, int64_t* edge_i, int64_t* edge_j,
                   int32_t* shifts) {
  if (!s) {
    nl::set_error("server pointer is NULL");
    return -1;
  }
  return s->s->take(id, edge_i, edge_j, shifts);
}

int nl_server_result_device(nl_server* s, int64_t id, const int64_t** edge_i,
                            const int64_t** edge_j, const int32_t** shifts) {
  if (!s) {
    nl::set_error("server pointer is NULL");
    return -1;
  }
  return s->s->device_view(id, edge_i, edge_j, shifts);
}

int nl_server_take_ex(nl_server* s, int64_t id, int64_t* edge_i, int64_t* edge_j,
                      int32_t* shifts, double* dist, double* disp) {
  if (!s) {
    nl::set_error("server pointer is NULL");
    return -1;
  }
  if ((dist && !s->s->has_dist()) || (disp && !s->s->has_disp())) {
    nl::set_error("the server was created without NL_DIST / NL_DISP");
    return -1;
  }
  return s->s->take(id, edge_i, edge_j, shifts, dist, disp);
}

int nl_server_result_arrays(nl_server* s, int64_t id, const int64_t** edge_i,
                            const int64_t** edge_j, const int32_t** shifts, const double** dist,
                            const double** disp) {
  if (!s) {
    nl::set_error("server pointer is NULL");
    return -1;
  }
  if ((dist && !s->s->has_dist()) || (disp && !s->s->has_disp())) {
    nl::set_error("the server was created without NL_DIST / NL_DISP");
    return -1;
  }
  return s->s->device_view(id, edge_i, edge_j, shifts, dist, disp);
}

int nl_server_release(nl_server* s, int64_t id) {
  if (!s) {
    nl::set_error("server pointer is NULL");
    return -1;
  }
  return s->s->release(id);
}

int nl_server_stats(nl_server* s, int64_t* completed, int64_t* batches, int64_t* max_inflight,
                    int64_t* max_batch, int64_t* refills) {
  if (!s) {
    nl::set_error("server pointer is NULL");
    return -1;
  }
  s->s->stats(completed, batches, max_inflight, max_batch, refills);
  return 0;
}

int nl_server_allocations(nl_server* s, int64_t* n) {
  if (!s || !n) {
    nl::set_error("server or output pointer is NULL");
    return -1;
  }
  *n = s->s->allocations();
  return 0;
}

void nl_server_free(nl_server* s) {
  if (!s) return;
  delete s->s;
  delete s;
}

int nl_set_device(int device) {
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    nl::set_error(std::string("hipSetDevice: ") + hipGetErrorString(e));
    return 1;
  }
  return 0;
}

int nl_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

}  // extern "C"
