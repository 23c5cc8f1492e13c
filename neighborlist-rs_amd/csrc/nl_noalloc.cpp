// nl_noalloc.cpp -- nl_search_noalloc, the search entry point that allocates
// nothing: the pipeline of nl_search_ex (device-resident positions) with its
// device workspace in a static array of this file, its host staging in
// ordinary (pageable) memory, and the output buffers and the stream given by
// the caller.  For callers that cannot allocate device memory: code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nl_api.h"
#include "../../include/nl_engine_ext.h"
#include "nl_engine.h"

namespace nl {
// Largest workspace of one call (the pipeline's buffers for the request:
// about 300 B per atom with three lists and the default work split).
constexpr size_t kNoAllocBytes = size_t(64) << 20;
}  // namespace nl

// The workspace: one copy on every GPU the library runs on, held until the
// process ends.  In host code this name is not a device address; the
// address on each device comes from hipGetSymbolAddress (workspace()).
static __device__ __align__(256) unsigned char scratch[nl::kNoAllocBytes];

namespace {

std::mutex g_mu;                           // one call at a time per process
std::map<int, unsigned char*> g_scratch;   // device -> address of `scratch` there

// Address of `scratch` on device `dev` (current), looked up on the first
// call on that device.
unsigned char* workspace(int dev) {
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) return it->second;
  unsigned char* p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    nl::set_error(std::string("hipGetSymbolAddress of the workspace failed: ") +
                  hipGetErrorString(e));
    return nullptr;
  }
  g_scratch[dev] = p;
  return p;
}

size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" int nl_search_noalloc(const double* d_positions, size_t n, const int64_t* sys_sizes,
                                 size_t n_sys, const double* cells, const uint8_t* has_cell,
                                 const uint8_t* pbc, const double* cutoffs, size_t n_cutoffs,
                                 int disjoint, int multi, unsigned options, int64_t* d_edges,
                                 int32_t* d_shifts, int64_t capacity, int64_t* list_base,
                                 int64_t* list_len, void* stream) {
  using nl::set_error;
  if (options & ~(unsigned)(NL_DEVICE_POSITIONS | NL_LOCAL_INDEX)) {
    set_error("nl_search_noalloc takes NL_LOCAL_INDEX only (positions are on the device)");
    return 1;
  }
  if (nl::capi_check_cutoffs(cutoffs, n_cutoffs, "All cutoffs must be positive and finite"))
    return 1;
  if (n_cutoffs > 0 && (!list_base || !list_len)) {
    set_error("list_base and list_len are NULL");
    return 1;
  }
  if (capacity < 0 || (capacity > 0 && (!d_edges || !d_shifts))) {
    set_error("the output buffers are NULL or the capacity is negative");
    return 1;
  }
  for (size_t q = 0; q < n_cutoffs; ++q) list_base[q] = list_len[q] = 0;
  if (n == 0 || n_cutoffs == 0) return 0;
  nl::Request req;
  if (nl::capi_sizes_request(d_positions, n, sys_sizes, n_sys, cells, has_cell, pbc, req))
    return 1;
  req.pos_device = true;
  req.cutoffs.assign(cutoffs, cutoffs + n_cutoffs);
  req.multi = multi != 0;
  req.disjoint = disjoint != 0;
  req.global_index = (options & NL_LOCAL_INDEX) == 0;
  std::lock_guard<std::mutex> g(g_mu);
  // the tunables of the workload class if an engine tuned it already; no
  // engine is created and nothing is timed here
  nl::Engine* eng = nl::existing_engine_for_current_device();
  const nl::Tuning t = eng ? eng->tuned_or_untimed(req) : nl::untimed_tuning(req);
  nl::Pipeline P;
  if (P.prepare(req, t)) return 1;
  // carve the pipeline's buffers out of the workspace (positions are the
  // caller's)
  const std::vector<size_t> sz = P.buffer_sizes();
  size_t need = 0;
  for (int i = 1; i < nl::kNumBuffers; ++i) need += round256(std::max<size_t>(sz[i], 16));
  if (need > nl::kNoAllocBytes) {
    set_error("nl_search_noalloc: the request needs a workspace of " + std::to_string(need) +
              " bytes, more than its " + std::to_string(nl::kNoAllocBytes) +
              " (use nl_search_ex for requests of this size)");
    return 1;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    set_error("no HIP device available (libnlmi355 runs on the GPU only)");
    return 1;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  unsigned char* base = workspace(dev);
  if (!base) return 1;
  std::vector<void*> ptr(nl::kNumBuffers, nullptr);
  ptr[0] = const_cast<double*>(d_positions);
  size_t at = 0;
  for (int i = 1; i < nl::kNumBuffers; ++i) {
    ptr[i] = base + at;
    at += round256(std::max<size_t>(sz[i], 16));
  }
  nl::Buffers B;
  P.assign(B, ptr);
  // host staging: descriptors up, list bounds down (pageable memory; the
  // stream is synchronised before the block goes out of scope)
  std::vector<char> stage(P.pinned_bytes());
  P.stage_host(stage.data());
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = P.enqueue_upload(B, stage.data(), st);
  // the count pass, then the fill pass into the caller's buffers: it writes
  // nothing when the lists exceed the capacity
  if (e == hipSuccess) e = P.enqueue_build_count(B, st);
  if (e == hipSuccess) e = P.enqueue_fill(B, d_edges, d_shifts, nullptr, capacity, st);
  if (e == hipSuccess) e = P.enqueue_bounds(B, stage.data(), st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    set_error(std::string("HIP error in nl_search_noalloc: ") + hipGetErrorString(e));
    return 1;
  }
  const int64_t* hb = P.host_bounds(stage.data());
  for (size_t q = 0; q < n_cutoffs; ++q) {
    const int r = P.rank_of((int)q);
    list_base[q] = hb[r];
    list_len[q] = hb[r + 1] - hb[r];
  }
  if (hb[P.K] > capacity) {
    set_error("nl_search_noalloc: the lists hold " + std::to_string(hb[P.K]) +
              " edges, more than the capacity of " + std::to_string(capacity) +
              "; nothing was written (list_len holds the sizes)");
    return 2;
  }
  return 0;
}
