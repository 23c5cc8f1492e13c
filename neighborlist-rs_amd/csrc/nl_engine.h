// nl_engine.h -- host side of the engine: request preparation, the launch
// sequence (Pipeline), per-device state and scratch (Engine), graph-captured
// steady-state plans (Plan), result objects and the autotuner.
#pragma once
#include <atomic>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "nl_common.h"
#include "nl_kernels.h"

namespace nl {

// Process-wide thresholds (config.rs:4-39 semantics) and engine defaults.
struct Config {
  std::atomic<size_t> bf_threshold{1000};
  std::atomic<size_t> par_threshold{300};
  std::atomic<size_t> stack_threshold{1000};
  std::atomic<int> cell_div{0};  // GPU bin edge = rc_max / cell_div; 0: per system (density)
  std::atomic<int> use_fp32{0};  // FP32 filter + FP64 band recheck
  std::atomic<int> group{4};     // home atoms per cell-list work item
  std::atomic<int> parts{1};     // half-stencil slices per home group
  std::atomic<int> autotune{1};  // pick the four above per workload class
  // single small brute-force systems without the pipeline: 0 off, 1 on (one
  // block up to 128 atoms, two launches up to 1024), 2 one block and 3 two
  // launches for every single system up to 1024 atoms (the plan decides)
  std::atomic<int> small_path{1};
  // pair kernels: XCD runs of this many work-list items in the count pass,
  // four times as many in the fill (0: dispatch order);
  // NLMI355_XCD_CHUNK overrides the default
  std::atomic<int> xcd_chunk{16};
  // cell-list bins keyed by their linear index (0) or their Morton rank (1);
  // linear measured about 1% faster on the ragged batch (the rank lookups
  // cost more than the fewer cache misses save)
  std::atomic<int> bin_order{0};
  // batches of systems of <= kFusedMax atoms: the build stages in one launch
  std::atomic<int> fused_build{1};
  Config();
};
Config& config();

void set_error(const std::string& msg);
const char* last_error();

struct SysIn {
  int has_cell = 0;
  double h[9] = {0};
  uint8_t pbc[3] = {1, 1, 1};
};

struct Request {
  const double* pos = nullptr;  // n x 3
  int64_t n = 0;
  bool pos_device = false;      // pos is a device pointer
  std::vector<int64_t> starts;  // n_sys + 1
  std::vector<SysIn> sys;
  std::vector<double> cutoffs;  // input order
  std::vector<double> sys_cutoff;  // optional: one cutoff per system (then cutoffs = {max})
  bool multi = false;
  bool disjoint = false;
  bool global_index = true;
  bool want_ranges = false;     // per-system edge ranges
  bool want_dist = false;       // FP64 |r_ij| per edge
  bool want_disp = false;       // FP64 r_ij = r_j + H S - r_i per edge
  bool full = false;            // full lists: (i, j, S) and (j, i, -S) for every pair
  // the caller's stream: device-resident positions are read after the work
  // already queued on it, and released outputs are reused only after it
  bool has_ustream = false;
  hipStream_t ustream = nullptr;
  int cell_div = 0;             // 0: tuned / engine default
  int use_fp32 = -1;            // -1: tuned / engine default
  int group = 0;                // 0: tuned / engine default
  int parts = 0;                // 0: tuned / engine default
};

// Requests below this many atoms are not autotuned (see Engine::tuning_for).
constexpr int64_t kTuneMinAtoms = 2048;

struct Tuning {
  int cell_div = 0, use_fp32 = 0, group = 4, parts = 1;
};

// Device buffers of one launch sequence (owned by an Engine or a Plan).
struct Buffers {
  double* pos = nullptr;
  SysDesc* sys = nullptr;
  int64_t* start = nullptr;
  int64_t* items = nullptr;
  double* rc2 = nullptr;
  float2* band = nullptr;
  uint32_t *keysA = nullptr, *valsA = nullptr, *keysB = nullptr, *valsB = nullptr;
  uint32_t* zero = nullptr;
  int32_t* cell_start = nullptr;
  Pos64* spos64 = nullptr;
  float4* spos32 = nullptr;
  int4* sidx = nullptr;
  uint32_t* counts = nullptr;
  int64_t* offsets = nullptr;
  int64_t* bounds = nullptr;
  int64_t* list = nullptr;  // work items that have work (any order)
  uint32_t* rcnt = nullptr;    // radix: digit counts per tile [256][tiles]
  int64_t* roff = nullptr;     // radix: their exclusive offsets [256 * tiles + 1]
  int64_t* rbounds = nullptr;  // radix: scan bounds [2]
  int32_t* bin_rank = nullptr; // Morton order: [bin] -> rank (null: linear order)
  int32_t* rank_bin = nullptr; // Morton order: [bin_base + rank] -> linear bin
};
constexpr int kNumBuffers = 21;

// Host-prepared launch sequence for one request shape.
class Pipeline {
 public:
  // Host work only: validation of cells, sorted cutoffs, descriptors, sizes.
  int prepare(const Request& req, const Tuning& t);

  // sizes of each device buffer (bytes), in Buffers order (pos first)
  std::vector<size_t> buffer_sizes() const;
  void assign(Buffers& b, const std::vector<void*>& ptrs) const;
  size_t pinned_bytes() const;
  void stage_host(char* pinned) const;  // descriptors etc. into pinned memory

  // Stages, all asynchronous on `st`.
  hipError_t enqueue_upload(const Buffers& b, const char* pinned, hipStream_t st) const;
  // restore: a device copy of the upload block (descriptors, system starts,
  // item bases, cutoffs), copied in by the first launch in place of
  // enqueue_upload's host copy (graph-captured plans: one node fewer)
  hipError_t enqueue_build_count(const Buffers& b, hipStream_t st,
                                 const void* restore = nullptr) const;
  hipError_t enqueue_build_stages(const Buffers& b, hipStream_t st) const;
  hipError_t enqueue_fill(const Buffers& b, void* edges, void* shifts, void* dist,
                          int64_t capacity, hipStream_t st, void* disp = nullptr) const;
  hipError_t enqueue_bounds(const Buffers& b, char* pinned, hipStream_t st) const;
  const int64_t* host_bounds(const char* pinned) const;
  int rank_of(int list) const;  // input list -> sorted rank

  int64_t max_work = 0;  // bound on the work items with work (pair-kernel grids)
  int64_t n = 0, n_items = 0, bin_total = 0, scan_len = 0, rtiles = 0, stiles = 0,
          rstiles = 0, z_words = 0;
  int S = 0, K = 0, passes = 0, digit_bits = 8;
  // every system fits one block: one segmented sort launch (seg_passes
  // digit passes in LDS) instead of the global radix passes
  bool seg_sort = false;
  // every system has at most kFusedMax atoms: the whole build in one launch
  // (k_build_systems)
  bool fused_build = false;
  int seg_passes = 0;
  Tuning tune;
  Params prm{};
  bool pos_device = false, global_index = true, disjoint = false, per_sys = false, full = false;
  const double* host_pos = nullptr;
  std::vector<int> order;  // sorted rank -> input list
  std::vector<double> rc2; // sorted
  std::vector<SysDesc> desc;
  std::vector<int64_t> starts, item_base;

  // radix scratch block: counts (u32), then offsets and bounds (i64)
  size_t rdigits() const { return (size_t)1 << digit_bits; }
  size_t roff_at() const { return (4 * rdigits() * (size_t)rtiles + 7) & ~size_t(7); }
  size_t radix_bytes() const { return roff_at() + 8 * (rdigits() * (size_t)rtiles + 3); }
  // layout of the descriptor block (pinned staging and device alike)
  size_t off_start() const { return sizeof(SysDesc) * S; }
  size_t off_items() const { return off_start() + 8 * (size_t)(S + 1); }
  size_t off_rc2() const { return off_items() + 8 * (size_t)(S + 1); }
  size_t off_bounds() const { return off_rc2() + 8 * rc2.size(); }
};

class Engine;

// An event recorded on a consumer's stream when it released device memory:
// the memory is handed out again only once the event has completed.
struct Fence {
  hipEvent_t ev = nullptr;
  ~Fence();
};

// One request's output (the nl_result of the C ABI).
}  // namespace nl
extern "C" {
void* nl_host_alloc(size_t bytes);  // page-locked host pool (nl_capi.cpp)
void nl_host_free(void* p);
}
namespace nl {

struct Result {
  Engine* eng = nullptr;
  int k = 0;
  std::vector<int64_t> len;    // per list, input order
  std::vector<int64_t> base;   // first global edge of each list, input order
  int64_t total = 0;
  void* edges = nullptr;       // int64 [2 * capacity]
  size_t edges_bytes = 0;
  void* shifts = nullptr;      // int32 [3 * capacity]
  size_t shifts_bytes = 0;
  void* dist = nullptr;        // double [capacity] (optional)
  size_t dist_bytes = 0;
  void* disp = nullptr;        // double [3 * capacity] (optional)
  size_t disp_bytes = 0;
  // page-locked copy of edges [2 * mirror_cap] then shifts [3 * mirror_cap],
  // made before the search's one synchronisation (small single systems
  // whose positions came from host memory: the host reads the lists without
  // another device round trip)
  void* mirror = nullptr;
  int64_t mirror_cap = 0;
  std::vector<int64_t> ranges; // [s][list(input order)][2] global edge ids (optional)
  std::vector<int32_t> modes;  // per system (optional, with ranges)
  Tuning tune;
  bool full = false;
  // consumer stream of exported views (DLPack): the output blocks go back
  // to the pool behind an event recorded there when the last view is dropped
  hipStream_t ustream = nullptr;
  std::mutex fence_mu;
  std::shared_ptr<Fence> fence;
  void fence_on_consumer();
  ~Result();
};

struct StageTimes {  // milliseconds, hipEvent-measured on the engine stream
  float build = 0, count = 0, fill = 0, total = 0;
  int64_t units = 0, edges = 0, bins = 0;
};

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();
  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }

  // Runs the whole pipeline; on success *out owns device buffers.
  int run(const Request& req, Result** out);
  // Picks the tunables for the request (autotuning on a new workload class).
  Tuning tuning_for(const Request& req);

  void* dev_alloc(size_t bytes);
  void dev_release(void* p, size_t bytes, std::shared_ptr<Fence> fence = nullptr);
  StageTimes last_times() const { return times_; }
  void set_timing(bool on) { timing_ = on; }
  std::mutex& mutex() { return mu_; }
  int fail(hipError_t e, const char* where);
  int alloc_outputs(Result* res, int64_t edges, bool dist, bool disp = false);
  std::string tune_report();
  void clear_tuning();
  // the tuned tunables of the request's workload class, if it has been
  // tuned already (times nothing); else untimed_tuning(req)
  Tuning tuned_or_untimed(const Request& req);

 private:
  int run_with(const Request& req, const Tuning& t, Result** out);
  bool small_eligible(const Request& req, const Pipeline& P) const;
  int run_small(const Request& req, const Pipeline& P, Result* res);
  bool small_batch_eligible(const Request& req, const Pipeline& P) const;
  int run_small_batch(const Request& req, const Pipeline& P, Result* res);
  int ensure_small_buf();
  void* scratch(int slot, size_t bytes);
  void* pinned(size_t bytes);

  int device_;
  hipStream_t stream_ = nullptr;
  std::mutex mu_;
  std::mutex pool_mu_;
  struct PoolBlock {
    void* p;
    std::shared_ptr<Fence> fence;  // null: free for use at once
  };
  std::multimap<size_t, PoolBlock> pool_;
  struct Scratch {
    void* p = nullptr;
    size_t cap = 0;
  };
  std::vector<Scratch> scratch_;
  hipEvent_t ev_[6] = {};
  bool timing_ = false;
  StageTimes times_;
  void* pinned_ = nullptr;  // host staging for small transfers (hipHostMalloc)
  void* small_buf_ = nullptr;  // device block of the single-small-system paths
  size_t pinned_cap_ = 0;
  // output-size estimate of the last request shape: lets the fill pass be
  // launched without waiting for the counts
  int64_t last_sig_[4] = {-1, -1, -1, -1};
  int64_t last_total_ = -1;
  std::mutex tune_mu_;
  std::map<std::tuple<int, int, int, int, int>, Tuning> tuned_;
};

Engine* engine_for_current_device();
// the current device's engine if one was created, else null (creates nothing)
Engine* existing_engine_for_current_device();

// The tunables a request gets without timing anything: the engine defaults
// (or the request's own), with the small-request work split below
// kTuneMinAtoms atoms.
Tuning untimed_tuning(const Request& req);

// Request builder and cutoff check of the C ABI (nl_capi.cpp).
int capi_sizes_request(const double* pos, size_t n, const int64_t* sizes, size_t n_sys,
                       const double* cells, const uint8_t* has_cell, const uint8_t* pbc,
                       Request& req);
int capi_check_cutoffs(const double* c, size_t k, const char* msg);

// Graph-captured steady state: a fixed batch shape (systems, cells, cutoffs)
// whose positions live at a fixed device address and change between runs
// (a molecular-dynamics loop, a serving slot).  One hipGraph launch per run.
class Plan {
 public:
  ~Plan();
  int create(Engine* eng, const Request& req, int64_t capacity);
  int run();  // replay; grows outputs (re-capture) when too small
  int64_t total() const { return total_; }
  int64_t capacity() const { return cap_; }
  int64_t len(int list) const;
  int64_t base(int list) const;
  const Pipeline& pipe() const { return pipe_; }
  void* edges() const { return edges_; }
  void* shifts() const { return shifts_; }
  void* dist() const { return dist_; }  // FP64 |r_ij| (want_dist), else null
  void* disp() const { return disp_; }  // FP64 r_ij (want_disp), else null
  bool has_ranges() const { return want_ranges_; }
  // per-system edge ranges of the last run, input list order, as
  // nl_result_system_ranges (want_ranges)
  const std::vector<int64_t>& ranges() const { return ranges_; }
  const std::vector<int32_t>& modes() const { return modes_; }
  hipStream_t stream() const { return st_; }
  int replays = 0, recaptures = 0;

 private:
  int capture();
  int alloc_out(int64_t cap);
  void free_out();
  void read_ranges();
  Engine* eng_ = nullptr;
  Pipeline pipe_;
  Buffers buf_;
  std::vector<void*> owned_;
  char* pinned_ = nullptr;
  void* d_static_ = nullptr;  // device copy of the upload block (set at create)
  hipStream_t st_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  void* edges_ = nullptr;
  void* shifts_ = nullptr;
  void* dist_ = nullptr;
  void* disp_ = nullptr;
  bool want_dist_ = false, want_disp_ = false, want_ranges_ = false;
  int64_t* d_ranges_ = nullptr;  // device: [S][K_sorted][2]
  char* pin_ranges_ = nullptr;   // pinned: the ranges, then the S descriptors
  std::vector<int64_t> ranges_;
  std::vector<int32_t> modes_;
  int64_t cap_ = 0, total_ = 0;
  std::vector<int64_t> bounds_;
};

}  // namespace nl
