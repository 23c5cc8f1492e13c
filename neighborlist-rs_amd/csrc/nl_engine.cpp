// nl_engine.cpp -- host engine: request planning, scratch buffers, the
// launch sequence and result objects.  One Engine per device; all calls on
// a device are serialised by the engine mutex.
#include "nl_engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

namespace nl {

Config& config() {
  static Config c;
  return c;
}

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
const char* last_error() { return g_error.c_str(); }

Result::~Result() {
  if (!eng) return;
  if (edges) eng->dev_release(edges, edges_bytes);
  if (shifts) eng->dev_release(shifts, shifts_bytes);
  if (dist) eng->dev_release(dist, dist_bytes);
}

// scratch slots
enum Slot {
  kPos, kSys, kStart, kItems, kRc2, kBand, kKeysA, kValsA, kKeysB, kValsB, kZero, kCellStart, kSpos64,
  kSpos32, kSidx, kCounts, kOffsets, kBounds, kRanges, kSlots
};

Engine::Engine(int device) : device_(device), scratch_(kSlots) {
  (void)hipSetDevice(device_);
  (void)hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking);
  for (auto& e : ev_) (void)hipEventCreate(&e);
}

Engine::~Engine() {
  // process-lifetime object; device memory is reclaimed at exit
}

int Engine::fail(hipError_t e, const char* where) {
  set_error(std::string("HIP error in ") + where + ": " + hipGetErrorString(e));
  return 1;
}

void* Engine::scratch(int slot, size_t bytes) {
  Scratch& s = scratch_[slot];
  if (bytes == 0) bytes = 16;
  if (s.cap < bytes) {
    if (s.p) (void)hipFree(s.p);
    size_t cap = std::max(bytes, s.cap + s.cap / 2);
    cap = (cap + 255) & ~size_t(255);
    if (hipMalloc(&s.p, cap) != hipSuccess) {
      s.p = nullptr;
      s.cap = 0;
      return nullptr;
    }
    s.cap = cap;
  }
  return s.p;
}

void* Engine::pinned(size_t bytes) {
  if (pinned_cap_ < bytes) {
    if (pinned_) (void)hipHostFree(pinned_);
    size_t cap = std::max(bytes, pinned_cap_ * 2);
    if (hipHostMalloc(&pinned_, cap, hipHostMallocDefault) != hipSuccess) {
      pinned_ = nullptr;
      pinned_cap_ = 0;
      return nullptr;
    }
    pinned_cap_ = cap;
  }
  return pinned_;
}

int Engine::alloc_outputs(Result* res, int64_t edges, bool dist) {
  if (res->edges) dev_release(res->edges, res->edges_bytes);
  if (res->shifts) dev_release(res->shifts, res->shifts_bytes);
  if (res->dist) dev_release(res->dist, res->dist_bytes);
  res->edges = res->shifts = res->dist = nullptr;
  const int64_t m = std::max<int64_t>(edges, 1);
  res->edges_bytes = sizeof(int64_t) * 2 * m;
  res->shifts_bytes = sizeof(int32_t) * 3 * m;
  res->edges = dev_alloc(res->edges_bytes);
  res->shifts = dev_alloc(res->shifts_bytes);
  if (dist) {
    res->dist_bytes = sizeof(double) * m;
    res->dist = dev_alloc(res->dist_bytes);
  }
  if (!res->edges || !res->shifts || (dist && !res->dist)) {
    set_error("out of device memory for the edge lists");
    return 1;
  }
  return 0;
}

void* Engine::dev_alloc(size_t bytes) {
  bytes = (std::max<size_t>(bytes, 256) + 255) & ~size_t(255);
  {
    std::lock_guard<std::mutex> g(pool_mu_);
    auto it = pool_.lower_bound(bytes);
    if (it != pool_.end() && it->first <= 2 * bytes) {
      void* p = it->second;
      pool_.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    // release cached blocks and retry once
    std::lock_guard<std::mutex> g(pool_mu_);
    for (auto& kv : pool_) (void)hipFree(kv.second);
    pool_.clear();
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  }
  return p;
}

void Engine::dev_release(void* p, size_t bytes) {
  if (!p) return;
  bytes = (std::max<size_t>(bytes, 256) + 255) & ~size_t(255);
  std::lock_guard<std::mutex> g(pool_mu_);
  pool_.emplace(bytes, p);
}

static int bits_for(int64_t v) {
  int b = 1;
  while (b < 32 && (int64_t(1) << b) <= v) ++b;
  return b;
}

#define NL_CHECK(call)                           \
  do {                                           \
    hipError_t e_ = (call);                      \
    if (e_ != hipSuccess) return fail(e_, #call); \
  } while (0)

int Engine::run(const Request& req, Result** out) {
  *out = nullptr;
  const int64_t n = req.n;
  const int S = (int)req.sys.size();
  const int K = (int)req.cutoffs.size();
  if (K <= 0 || K > kMaxCutoffs) {
    set_error("number of cutoffs must be between 1 and 64");
    return 1;
  }
  std::unique_ptr<Result> res(new Result());
  res->eng = this;
  res->k = K;
  res->len.assign(K, 0);
  res->base.assign(K, 0);
  if (n == 0 || S == 0) {
    *out = res.release();
    return 0;
  }
  (void)hipSetDevice(device_);

  // cutoffs: stable ascending order (Rust sort_by is stable), squares in f64
  std::vector<int> order(K);
  for (int q = 0; q < K; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return req.cutoffs[a] < req.cutoffs[b]; });
  std::vector<double> rc2(K);
  for (int q = 0; q < K; ++q) rc2[q] = req.cutoffs[order[q]] * req.cutoffs[order[q]];
  double rc_max = NAN;  // NaN-ignoring fold (single.rs:106)
  for (double c : req.cutoffs) rc_max = std::isnan(rc_max) ? c : std::fmax(rc_max, c);

  // per-system descriptors (static part), bin reservation
  std::vector<SysDesc> desc(S);
  int64_t bin_total = 0;
  for (int s = 0; s < S; ++s) {
    SysDesc& d = desc[s];
    std::memset(&d, 0, sizeof d);
    d.atom_start = req.starts[s];
    d.n_atoms = (int32_t)(req.starts[s + 1] - req.starts[s]);
    d.has_cell = req.sys[s].has_cell;
    for (int k = 0; k < 3; ++k) d.pbc[k] = req.sys[s].has_cell ? (req.sys[s].pbc[k] != 0) : 0;
    if (d.has_cell) {
      std::memcpy(d.h, req.sys[s].h, sizeof d.h);
      double inv[9];
      if (!inverse3(d.h, inv)) {
        set_error("Cell matrix is not invertible");  // cell.rs:6
        return 1;
      }
    }
    const int64_t cap = std::min<int64_t>(int64_t(1) << 22, 2 * (int64_t)d.n_atoms + 64);
    d.bin_cap = (int32_t)cap;
    d.bin_base = bin_total;
    bin_total += cap;
  }
  if (bin_total >= (int64_t(1) << 31) || n >= (int64_t(1) << 31)) {
    set_error("request too large for one launch (more than 2^31 atoms or bins)");
    return 1;
  }
  const int passes = (bits_for(bin_total - 1) + kRadixBits - 1) / kRadixBits;
  const int64_t rtiles = radix_tiles(n);
  // work items: per system its atoms (brute-force rows), then
  // (slot, stencil slice) pairs (cell-list home groups)
  const int group = std::max(1, std::min(64, req.group > 0 ? req.group : config().group.load()));
  const int parts = std::max(1, std::min(16, req.parts > 0 ? req.parts : config().parts.load()));
  const int64_t n_items = n * (1 + parts);
  std::vector<int64_t> item_base(S + 1);
  for (int s = 0; s <= S; ++s) item_base[s] = req.starts[s] * (1 + parts);
  const int64_t scan_len = (int64_t)K * n_items;
  const int64_t stiles = radix_tiles(scan_len);

  Params prm{};
  prm.n_atoms = n;
  prm.n_sys = S;
  prm.k = K;
  prm.disjoint = req.disjoint;
  prm.multi = req.multi;
  const uint64_t thr = config().bf_threshold.load();
  prm.bf_threshold_lo = (int32_t)(uint32_t)(thr & 0xffffffffu);
  prm.bf_threshold_hi = (int32_t)(uint32_t)(thr >> 32);
  prm.cell_div = req.cell_div > 0 ? req.cell_div : config().cell_div.load();
  prm.use_fp32 = req.use_fp32 >= 0 ? req.use_fp32 : config().use_fp32.load();
  prm.global_index = req.global_index;
  prm.rc_max = rc_max;

  // scratch
  double* d_pos = req.pos_device ? const_cast<double*>(req.pos)
                                 : (double*)scratch(kPos, sizeof(double) * 3 * n);
  SysDesc* d_sys = (SysDesc*)scratch(kSys, sizeof(SysDesc) * S);
  int64_t* d_start = (int64_t*)scratch(kStart, sizeof(int64_t) * (S + 1));
  int64_t* d_items = (int64_t*)scratch(kItems, sizeof(int64_t) * (S + 1));
  double* d_rc2 = (double*)scratch(kRc2, sizeof(double) * K);
  float2* d_band = (float2*)scratch(kBand, sizeof(float2) * S * K);
  uint32_t* keysA = (uint32_t*)scratch(kKeysA, 4 * n);
  uint32_t* valsA = (uint32_t*)scratch(kValsA, 4 * n);
  uint32_t* keysB = (uint32_t*)scratch(kKeysB, 4 * n);
  uint32_t* valsB = (uint32_t*)scratch(kValsB, 4 * n);
  // one zero-initialised block: radix histograms, radix look-back status,
  // tile tickets, scan look-back status (64-bit aligned first)
  const int64_t z_sstatus = 2 * stiles;             // u32 words
  const int64_t z_hist = 256 * passes;
  const int64_t z_rstatus = 256 * rtiles * passes;
  const int64_t z_tickets = 8;
  const int64_t z_words = z_sstatus + z_hist + z_rstatus + z_tickets;
  uint32_t* zero = (uint32_t*)scratch(kZero, 4 * z_words);
  int32_t* cell_start = (int32_t*)scratch(kCellStart, 4 * (bin_total + 1));
  Pos64* spos64 = (Pos64*)scratch(kSpos64, sizeof(Pos64) * n);
  float4* spos32 = (float4*)scratch(kSpos32, sizeof(float4) * n);
  int4* sidx = (int4*)scratch(kSidx, sizeof(int4) * n);
  uint32_t* counts = (uint32_t*)scratch(kCounts, 4 * scan_len);
  int64_t* offsets = (int64_t*)scratch(kOffsets, 8 * (scan_len + 1));
  int64_t* bounds = (int64_t*)scratch(kBounds, 8 * (K + 1));
  const size_t pin_desc = sizeof(SysDesc) * S, pin_start = 8 * (S + 1), pin_rc2 = 8 * K;
  const size_t pin_bounds = 8 * (K + 1);
  char* pin = (char*)pinned(pin_desc + 2 * pin_start + pin_rc2 + pin_bounds + 64);
  if (!d_pos || !d_sys || !d_start || !d_items || !d_rc2 || !d_band || !keysA || !valsA || !keysB || !valsB ||
      !zero || !cell_start || !spos64 || !spos32 || !sidx || !counts || !offsets || !bounds ||
      !pin) {
    set_error("out of device memory");
    return 1;
  }
  uint64_t* sstatus = (uint64_t*)zero;
  uint32_t* hist = zero + z_sstatus;
  uint32_t* rstatus = hist + z_hist;
  uint32_t* tickets = rstatus + z_rstatus;
  char* h_desc = pin;
  char* h_start = h_desc + pin_desc;
  char* h_items = h_start + pin_start;
  char* h_rc2 = h_items + pin_start;
  int64_t* h_bounds = (int64_t*)(h_rc2 + pin_rc2);
  std::memcpy(h_desc, desc.data(), pin_desc);
  std::memcpy(h_start, req.starts.data(), pin_start);
  std::memcpy(h_items, item_base.data(), pin_start);
  std::memcpy(h_rc2, rc2.data(), pin_rc2);

  hipStream_t st = stream_;
  if (!req.pos_device)
    NL_CHECK(hipMemcpyAsync(d_pos, req.pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  NL_CHECK(hipMemcpyAsync(d_sys, h_desc, pin_desc, hipMemcpyHostToDevice, st));
  NL_CHECK(hipMemcpyAsync(d_start, h_start, pin_start, hipMemcpyHostToDevice, st));
  NL_CHECK(hipMemcpyAsync(d_items, h_items, pin_start, hipMemcpyHostToDevice, st));
  NL_CHECK(hipMemcpyAsync(d_rc2, h_rc2, pin_rc2, hipMemcpyHostToDevice, st));
  if (timing_) NL_CHECK(hipEventRecord(ev_[0], st));

  // build: plan, bin, radix sort, finalize
  NL_CHECK(launch_fill_u32(zero, z_words, 0, st));
  NL_CHECK(launch_plan(d_pos, d_sys, d_band, d_rc2, prm, st));
  NL_CHECK(launch_bin(d_pos, d_sys, d_start, S, n, keysA, valsA, hist, passes, st));
  uint32_t *kin = keysA, *vin = valsA, *kout = keysB, *vout = valsB;
  for (int ps = 0; ps < passes; ++ps) {
    NL_CHECK(launch_radix_pass(kin, vin, kout, vout, n, kRadixBits * ps, hist + 256 * ps,
                               rstatus + (int64_t)256 * rtiles * ps, tickets + ps, st));
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  NL_CHECK(launch_finalize(kin, vin, d_pos, d_sys, d_start, S, n, bin_total, cell_start, spos64,
                           spos32, sidx, st));
  if (timing_) NL_CHECK(hipEventRecord(ev_[1], st));

  PairArgs A{};
  A.pos = d_pos;
  A.sys = d_sys;
  A.sys_start = d_start;
  A.item_base = d_items;
  A.keys = kin;
  A.cell_start = cell_start;
  A.spos64 = spos64;
  A.spos32 = spos32;
  A.sidx = sidx;
  A.rc2 = d_rc2;
  A.band = d_band;
  A.counts = counts;
  A.offsets = offsets;
  A.bounds = bounds;
  A.n = n_items;
  A.n_sys = S;
  A.k = K;
  A.disjoint = req.disjoint;
  A.use_fp32 = prm.use_fp32;
  A.global_index = req.global_index;
  A.group = group;
  A.parts = parts;
  for (int q0 = 0; q0 < K; q0 += kGroup) {
    A.q_lo = q0;
    A.q_cnt = std::min(kGroup, K - q0);
    NL_CHECK(launch_pairs(false, A, st));
  }
  NL_CHECK(launch_scan(counts, offsets, scan_len, n_items, K, bounds, sstatus, tickets + 4, st));
  if (timing_) NL_CHECK(hipEventRecord(ev_[2], st));

  // Fill straight away when an estimate of the output size exists (same
  // request shape as last time); the kernel writes nothing if it is short.
  const int64_t sig[4] = {n, S, K, (int64_t)req.multi * 2 + req.disjoint};
  const bool have_est = last_total_ >= 0 && std::equal(sig, sig + 4, last_sig_);
  int64_t cap = have_est ? last_total_ + last_total_ / 8 + 4096 : -1;
  auto launch_fill = [&](int64_t capacity) -> int {
    A.edges = (int64_t*)res->edges;
    A.shifts = (int32_t*)res->shifts;
    A.dist = (double*)res->dist;
    A.capacity = capacity;
    if (timing_) NL_CHECK(hipEventRecord(ev_[3], st));
    for (int q0 = 0; q0 < K; q0 += kGroup) {
      A.q_lo = q0;
      A.q_cnt = std::min(kGroup, K - q0);
      NL_CHECK(launch_pairs(true, A, st));
    }
    if (timing_) NL_CHECK(hipEventRecord(ev_[4], st));
    return 0;
  };
  if (have_est) {
    if (alloc_outputs(res.get(), cap, req.want_dist)) return 1;
    if (launch_fill(cap)) return 1;
  }
  NL_CHECK(hipMemcpyAsync(h_bounds, bounds, pin_bounds, hipMemcpyDeviceToHost, st));
  NL_CHECK(hipStreamSynchronize(st));
  std::vector<int64_t> hb(h_bounds, h_bounds + K + 1);
  const int64_t total = hb[K];
  res->total = total;
  if (!have_est || total > cap) {
    if (alloc_outputs(res.get(), total, req.want_dist)) return 1;
    if (total > 0 && launch_fill(total)) return 1;
  }
  for (int q = 0; q < 4; ++q) last_sig_[q] = sig[q];
  last_total_ = total;
  for (int q = 0; q < K; ++q) {
    const int r = (int)(std::find(order.begin(), order.end(), q) - order.begin());
    res->base[q] = hb[r];
    res->len[q] = hb[r + 1] - hb[r];
  }
  if (req.want_ranges) {
    int64_t* d_ranges = (int64_t*)scratch(kRanges, 8 * 2 * (int64_t)S * K);
    if (!d_ranges) {
      set_error("out of device memory");
      return 1;
    }
    NL_CHECK(launch_sys_ranges(offsets, d_items, S, K, n_items, d_ranges, st));
    std::vector<int64_t> rr((size_t)2 * S * K);
    NL_CHECK(hipMemcpyAsync(rr.data(), d_ranges, 8 * rr.size(), hipMemcpyDeviceToHost, st));
    std::vector<SysDesc> dd(S);
    NL_CHECK(hipMemcpyAsync(dd.data(), d_sys, sizeof(SysDesc) * S, hipMemcpyDeviceToHost, st));
    NL_CHECK(hipStreamSynchronize(st));
    res->ranges.assign((size_t)2 * S * K, 0);
    res->modes.resize(S);
    for (int s = 0; s < S; ++s) {
      res->modes[s] = dd[s].mode;
      for (int q = 0; q < K; ++q) {
        const int r = (int)(std::find(order.begin(), order.end(), q) - order.begin());
        res->ranges[2 * ((size_t)s * K + q)] = rr[2 * ((size_t)s * K + r)];
        res->ranges[2 * ((size_t)s * K + q) + 1] = rr[2 * ((size_t)s * K + r) + 1];
      }
    }
  }
  NL_CHECK(hipStreamSynchronize(st));
  if (timing_) {
    float t = 0;
    (void)hipEventElapsedTime(&t, ev_[0], ev_[1]);
    times_.build = t;
    (void)hipEventElapsedTime(&t, ev_[1], ev_[2]);
    times_.count = t;
    times_.fill = 0;
    if (total > 0) {
      (void)hipEventElapsedTime(&t, ev_[3], ev_[4]);
      times_.fill = t;
    }
    (void)hipEventElapsedTime(&t, ev_[0], total > 0 ? ev_[4] : ev_[2]);
    times_.total = t;
    times_.units = n;
    times_.edges = total;
    times_.bins = bin_total;
  }
  *out = res.release();
  return 0;
}

Engine* engine_for_current_device() {
  static std::mutex mu;
  static std::map<int, Engine*> engines;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  std::lock_guard<std::mutex> g(mu);
  auto it = engines.find(dev);
  if (it != engines.end()) return it->second;
  Engine* e = new Engine(dev);
  engines[dev] = e;
  return e;
}

}  // namespace nl
