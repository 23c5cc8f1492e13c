// nl_engine.h -- host side of the engine: per-device state, the launch
// sequence for one (possibly ragged, multi-cutoff) request, result objects.
#pragma once
#include <atomic>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "nl_common.h"
#include "nl_kernels.h"

namespace nl {

// Process-wide thresholds (config.rs:4-39 semantics).
struct Config {
  std::atomic<size_t> bf_threshold{1000};
  std::atomic<size_t> par_threshold{300};
  std::atomic<size_t> stack_threshold{1000};
  std::atomic<int> cell_div{2};  // GPU bin edge = rc_max / cell_div
  std::atomic<int> use_fp32{0};  // FP32 filter + FP64 band recheck
  std::atomic<int> group{8};     // home atoms per cell-list work item
  std::atomic<int> parts{2};     // half-stencil slices per home group
  std::atomic<int> autotune{1};
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
  bool multi = false;
  bool disjoint = false;
  bool global_index = true;
  bool want_ranges = false;     // per-system edge ranges
  bool want_dist = false;       // FP64 |r_ij| per edge
  int cell_div = 0;             // 0: engine default / tuned
  int use_fp32 = -1;            // -1: engine default
  int group = 0;                // 0: engine default
  int parts = 0;                // 0: engine default
};

class Engine;

// One request's output (the nl_result of the C ABI).
struct Result {
  Engine* eng = nullptr;
  int k = 0;
  std::vector<int64_t> len;    // per list, input order
  std::vector<int64_t> base;   // first global edge of each list, input order
  int64_t total = 0;
  void* edges = nullptr;       // int64 [2 * total]
  size_t edges_bytes = 0;
  void* shifts = nullptr;      // int32 [3 * total]
  size_t shifts_bytes = 0;
  void* dist = nullptr;        // double [total] (optional)
  size_t dist_bytes = 0;
  std::vector<int64_t> ranges; // [s][list(input order)][2] global edge ids (optional)
  std::vector<int32_t> modes;  // per system (optional, with ranges)
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

  void* dev_alloc(size_t bytes);
  void dev_release(void* p, size_t bytes);
  StageTimes last_times() const { return times_; }
  void set_timing(bool on) { timing_ = on; }
  std::mutex& mutex() { return mu_; }

 private:
  struct Scratch {
    void* p = nullptr;
    size_t cap = 0;
  };
  void* scratch(int slot, size_t bytes);
  void* pinned(size_t bytes);
  int fail(hipError_t e, const char* where);
  int alloc_outputs(Result* res, int64_t edges, bool dist);

  int device_;
  hipStream_t stream_ = nullptr;
  std::mutex mu_;
  std::mutex pool_mu_;
  std::multimap<size_t, void*> pool_;
  std::vector<Scratch> scratch_;
  hipEvent_t ev_[6] = {};
  bool timing_ = false;
  StageTimes times_;
  void* pinned_ = nullptr;  // host staging for small transfers (hipHostMalloc)
  size_t pinned_cap_ = 0;
  // output-size estimate of the last request shape: lets the fill pass be
  // launched without waiting for the counts
  int64_t last_sig_[4] = {-1, -1, -1, -1};
  int64_t last_total_ = -1;
};

Engine* engine_for_current_device();

}  // namespace nl
