// nl_engine.cpp -- host engine: request preparation (Pipeline), scratch and
// output pools, the eager launch path, the autotuner and graph-captured
// plans.  One Engine per device; all calls on a device are serialised by the
// engine mutex.
#include "nl_engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <sstream>

namespace nl {

Config::Config() {
  if (const char* v = std::getenv("NLMI355_XCD_CHUNK")) xcd_chunk = std::max(0, std::atoi(v));
}

Config& config() {
  static Config c;
  return c;
}

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
const char* last_error() { return g_error.c_str(); }

Fence::~Fence() {
  if (ev) (void)hipEventDestroy(ev);
}

// Called when an exported view of this result is dropped: (re)records the
// fence on the consumer's stream, so the last drop covers every use queued
// there before it.
void Result::fence_on_consumer() {
  if (!eng) return;
  int old = -1;
  (void)hipGetDevice(&old);
  if (old != eng->device()) (void)hipSetDevice(eng->device());
  {
    std::lock_guard<std::mutex> g(fence_mu);
    if (!fence) {
      fence = std::make_shared<Fence>();
      if (hipEventCreateWithFlags(&fence->ev, hipEventDisableTiming) != hipSuccess)
        fence->ev = nullptr;
    }
    if (fence->ev && hipEventRecord(fence->ev, ustream) != hipSuccess) {
      // no fence: wait here instead
      (void)hipStreamSynchronize(ustream);
    }
  }
  if (old >= 0 && old != eng->device()) (void)hipSetDevice(old);
}

Result::~Result() {
  if (mirror) nl_host_free(mirror);
  if (!eng) return;
  if (edges) eng->dev_release(edges, edges_bytes, fence);
  if (shifts) eng->dev_release(shifts, shifts_bytes, fence);
  if (dist) eng->dev_release(dist, dist_bytes, fence);
  if (disp) eng->dev_release(disp, disp_bytes, fence);
}

static int bits_for(int64_t v) {
  int b = 1;
  while (b < 32 && (int64_t(1) << b) <= v) ++b;
  return b;
}

#define NL_TRY(call)                 \
  do {                               \
    hipError_t e_ = (call);          \
    if (e_ != hipSuccess) return e_; \
  } while (0)

// ---------------------------------------------------------------------------
// Pipeline

int Pipeline::prepare(const Request& req, const Tuning& t) {
  n = req.n;
  S = (int)req.sys.size();
  K = (int)req.cutoffs.size();
  tune = t;
  // any number of cutoffs (search.rs:344-361): every per-list array is sized
  // by K and the pair kernels take the lists kGroup at a time
  if (K <= 0) {
    set_error("at least one cutoff is needed");
    return 1;
  }
  pos_device = req.pos_device;
  host_pos = req.pos;
  global_index = req.global_index;
  disjoint = req.disjoint;
  full = req.full;
  // cutoffs: stable ascending order (Rust sort_by is stable, search.rs:354-356),
  // squares in f64 (search.rs:357-360)
  order.resize(K);
  for (int q = 0; q < K; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return req.cutoffs[a] < req.cutoffs[b]; });
  per_sys = !req.sys_cutoff.empty();
  if (per_sys && (K != 1 || (int)req.sys_cutoff.size() != S)) {
    set_error("per-system cutoffs need one list and one cutoff per system");
    return 1;
  }
  rc2.resize(per_sys ? S : K);
  if (per_sys) {
    for (int s = 0; s < S; ++s) rc2[s] = req.sys_cutoff[s] * req.sys_cutoff[s];
  } else {
    for (int q = 0; q < K; ++q) rc2[q] = req.cutoffs[order[q]] * req.cutoffs[order[q]];
  }
  double rc_max = NAN;  // NaN-ignoring fold (single.rs:106)
  for (double c : req.cutoffs) rc_max = std::isnan(rc_max) ? c : std::fmax(rc_max, c);

  tune.group = std::max(1, std::min(64, tune.group));
  tune.parts = std::max(1, std::min(16, tune.parts));
  tune.cell_div = std::max(0, std::min(4, tune.cell_div));  // 0: per system
  // per-system descriptors (static part), bin reservation
  desc.assign(S, SysDesc{});
  bin_total = 0;
  for (int s = 0; s < S; ++s) {
    SysDesc& d = desc[s];
    std::memset(&d, 0, sizeof d);
    d.atom_start = req.starts[s];
    d.n_atoms = (int32_t)(req.starts[s + 1] - req.starts[s]);
    d.has_cell = req.sys[s].has_cell;
    d.rc_max = per_sys ? req.sys_cutoff[s] : rc_max;
    for (int k = 0; k < 3; ++k) d.pbc[k] = req.sys[s].has_cell ? (req.sys[s].pbc[k] != 0) : 0;
    int64_t cap = std::min<int64_t>(int64_t(1) << 22, 2 * (int64_t)d.n_atoms + 64);
    if (d.has_cell) {
      std::memcpy(d.h, req.sys[s].h, sizeof d.h);
      double inv[9];
      if (!inverse3(d.h, inv)) {
        set_error("Cell matrix is not invertible");  // cell.rs:6
        return 1;
      }
      // the grid k_plan will choose for this cell (fewer key bits, often
      // one radix pass less); should it come out larger on the device,
      // k_plan shrinks it to the reservation, which any covering grid allows
      double w[3];
      widths3(inv, w);
      const int cd = tune.cell_div > 0
                         ? tune.cell_div
                         : bins_per_cutoff(d.rc_max, (double)d.n_atoms, std::fabs(det3(d.h)));
      const double edge = d.rc_max / (double)cd;
      int64_t prod = 1;
      for (int k = 0; k < 3; ++k) {
        const double v = std::floor(w[k] / edge);
        prod *= v < 1.0 ? 1 : (v > 4096.0 ? 4096 : (int64_t)v);
        if (prod >= cap) break;
      }
      cap = std::min(cap, prod);
    }
    d.bin_cap = (int32_t)cap;
    d.bin_base = bin_total;
    bin_total += cap;
  }
  if (bin_total >= (int64_t(1) << 31) || n >= (int64_t(1) << 28)) {
    set_error("request too large for one launch (more than 2^28 atoms or 2^31 bins)");
    return 1;
  }
  {
    // fewest passes of at most 12-bit digits, digits as even as possible
    const int kb = bits_for(bin_total - 1);
    passes = (kb + kRadixMaxBits - 1) / kRadixMaxBits;
    digit_bits = (kb + passes - 1) / passes;
    // batches of small systems: one block sorts each system's local keys
    int64_t max_n = 0, max_cap = 1;
    for (int s = 0; s < S; ++s) {
      max_n = std::max<int64_t>(max_n, desc[s].n_atoms);
      max_cap = std::max<int64_t>(max_cap, desc[s].bin_cap);
    }
    seg_sort = max_n <= kSegMax;
    // the one-launch build (Morton-ranked bins take the separate launches,
    // see below).  One block builds a whole system, so above the segmented
    // sort's size it pays only when the batch has other systems to build
    // beside its largest: a single triclinic system of 10,008 atoms built in
    // 0.170 ms this way against 0.146 ms through the separate launches.
    // Mode 2 takes it whenever the systems fit.
    const int fb = config().fused_build.load();
    fused_build = max_n <= kFusedMax &&
                  (fb == 2 || (fb == 1 && (max_n <= kSegMax || n >= 4 * max_n)));
    const int lb = bits_for(max_cap - 1);
    seg_passes = (lb + kSegBits - 1) / kSegBits;
  }
  rtiles = radix_tiles(n);
  // work items: per system its atoms (brute-force rows), then
  // (slot, stencil slice) pairs (cell-list home groups)
  n_items = n * (1 + tune.parts);
  // Items that can hold work, at most (the pair-kernel grids stop there): a
  // system has n_s rows on the brute-force path, or `parts` items per home
  // group on the cell list; groups start every `group` slots of a bin, so
  // there are at most ceil(n_s / group) + min(n_s, bins) of them, and at
  // most n_s (one per slot)
  {
    const int64_t G = std::max(1, tune.group), Pp = std::max(1, tune.parts);
    max_work = 0;
    for (int s = 0; s < S; ++s) {
      const int64_t ns = desc[s].n_atoms;
      const int64_t groups = std::min(ns, (ns + G - 1) / G + std::min<int64_t>(ns, desc[s].bin_cap));
      max_work += std::max(ns, Pp * groups);
    }
  }
  starts = req.starts;
  item_base.resize(S + 1);
  for (int s = 0; s <= S; ++s) item_base[s] = starts[s] * (1 + tune.parts);
  scan_len = (int64_t)K * n_items;
  stiles = radix_tiles(scan_len);
  // one zero-initialised block: count-scan look-back status (u64, first),
  // radix digit-count scan status (u64, one block per pass), tickets
  rstiles = radix_tiles(((int64_t)1 << digit_bits) * rtiles);
  z_words = 2 * stiles + 2 * rstiles * passes + 16;

  prm = Params{};
  prm.n_atoms = n;
  prm.n_sys = S;
  prm.k = K;
  prm.disjoint = req.disjoint;
  prm.multi = req.multi;
  const uint64_t thr = config().bf_threshold.load();
  prm.bf_threshold_lo = (int32_t)(uint32_t)(thr & 0xffffffffu);
  prm.bf_threshold_hi = (int32_t)(uint32_t)(thr >> 32);
  prm.cell_div = tune.cell_div;
  prm.use_fp32 = tune.use_fp32;
  prm.global_index = req.global_index;
  prm.per_sys_rc = per_sys;
  prm.bin_order = config().bin_order.load() ? 1 : 0;
  if (prm.bin_order) fused_build = false;
  prm.rc_max = rc_max;
  return 0;
}

std::vector<size_t> Pipeline::buffer_sizes() const {
  return {
      pos_device ? 0 : sizeof(double) * 3 * (size_t)n,              // pos
      off_bounds(),            // sys | start | items | rc2: one upload
      0, 0, 0,                 // (start, items, rc2 live in the block above)
      sizeof(float2) * (size_t)S * K,                               // band
      4 * (size_t)n, 4 * (size_t)n, 4 * (size_t)n, 4 * (size_t)n,  // keys/vals A/B
      4 * (size_t)z_words,                                          // zero
      4 * (size_t)(bin_total + 1),                                  // cell_start
      sizeof(Pos64) * (size_t)n,                                    // spos64
      sizeof(float4) * (size_t)n,                                   // spos32
      sizeof(int4) * (size_t)n,                                     // sidx
      4 * (size_t)scan_len,                                         // counts
      8 * (size_t)(scan_len + 1),                                   // offsets
      8 * (size_t)(K + 1),                                          // bounds
      8 * (size_t)n_items,                                          // work list
      radix_bytes(),  // radix scratch: digit counts, their offsets, bounds
      prm.bin_order ? 8 * (size_t)bin_total : 0,                    // Morton tables
  };
}

void Pipeline::assign(Buffers& b, const std::vector<void*>& p) const {
  b.pos = (double*)p[0];
  b.sys = (SysDesc*)p[1];
  b.start = (int64_t*)((char*)p[1] + off_start());
  b.items = (int64_t*)((char*)p[1] + off_items());
  b.rc2 = (double*)((char*)p[1] + off_rc2());
  b.band = (float2*)p[5];
  b.keysA = (uint32_t*)p[6];
  b.valsA = (uint32_t*)p[7];
  b.keysB = (uint32_t*)p[8];
  b.valsB = (uint32_t*)p[9];
  b.zero = (uint32_t*)p[10];
  b.cell_start = (int32_t*)p[11];
  b.spos64 = (Pos64*)p[12];
  b.spos32 = (float4*)p[13];
  b.sidx = (int4*)p[14];
  b.counts = (uint32_t*)p[15];
  b.offsets = (int64_t*)p[16];
  b.bounds = (int64_t*)p[17];
  b.list = (int64_t*)p[18];
  b.rcnt = (uint32_t*)p[19];
  b.roff = (int64_t*)((char*)p[19] + roff_at());
  b.rbounds = b.roff + ((int64_t)1 << digit_bits) * rtiles + 1;
  b.bin_rank = prm.bin_order ? (int32_t*)p[20] : nullptr;
  b.rank_bin = prm.bin_order ? (int32_t*)p[20] + bin_total : nullptr;
}

size_t Pipeline::pinned_bytes() const { return off_bounds() + 8 * (size_t)(K + 1) + 64; }

void Pipeline::stage_host(char* pin) const {
  std::memcpy(pin, desc.data(), sizeof(SysDesc) * S);
  std::memcpy(pin + off_start(), starts.data(), 8 * (size_t)(S + 1));
  std::memcpy(pin + off_items(), item_base.data(), 8 * (size_t)(S + 1));
  std::memcpy(pin + off_rc2(), rc2.data(), 8 * rc2.size());
}

const int64_t* Pipeline::host_bounds(const char* pin) const {
  return (const int64_t*)(pin + off_bounds());
}

int Pipeline::rank_of(int list) const {
  return (int)(std::find(order.begin(), order.end(), list) - order.begin());
}

hipError_t Pipeline::enqueue_upload(const Buffers& b, const char* pin, hipStream_t st) const {
  if (!pos_device)
    NL_TRY(hipMemcpyAsync(b.pos, host_pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  // descriptors, system starts, item bases and cutoffs: one contiguous copy
  NL_TRY(hipMemcpyAsync(b.sys, pin, off_bounds(), hipMemcpyHostToDevice, st));
  return hipSuccess;
}

static PairArgs pair_args(const Pipeline& P, const Buffers& b) {
  PairArgs A{};
  A.pos = b.pos;
  A.sys = b.sys;
  A.sys_start = b.start;
  A.item_base = b.items;
  // the sorted keys end in A after an even number of radix passes (in B
  // after the segmented sort or the one-launch build)
  A.keys = (P.seg_sort || P.fused_build) ? b.keysB : ((P.passes % 2 == 0) ? b.keysA : b.keysB);
  A.bin_rank = b.bin_rank;
  A.rank_bin = b.rank_bin;
  A.cell_start = b.cell_start;
  A.spos64 = b.spos64;
  A.spos32 = b.spos32;
  A.sidx = b.sidx;
  A.rc2 = b.rc2;
  A.band = b.band;
  A.counts = b.counts;
  A.offsets = b.offsets;
  A.list = b.list;
  A.n_real = b.zero + 2 * P.stiles + 2 * P.rstiles * P.passes + 6;
  A.bounds = b.bounds;
  A.n = P.n_items;
  // the bound on the items with work: the pair-kernel grids stop there
  A.max_items = P.max_work;
  A.n_sys = P.S;
  A.k = P.K;
  A.disjoint = P.disjoint;
  A.use_fp32 = P.tune.use_fp32;
  A.global_index = P.global_index;
  A.per_sys_rc = P.per_sys;
  A.full = P.full ? 1 : 0;
  A.group = P.tune.group;
  A.parts = P.tune.parts;
  A.xcd_chunk = config().xcd_chunk.load();
  return A;
}

static_assert(sizeof(SysDesc) % 8 == 0, "the upload block is copied in whole words");
hipError_t Pipeline::enqueue_build_count(const Buffers& b, hipStream_t st,
                                         const void* restore) const {
  uint64_t* sstatus = (uint64_t*)b.zero;
  uint32_t* tickets = b.zero + 2 * stiles + 2 * rstiles * passes;
  // the upload block is whole 8-byte words (descriptors, i64 and f64 arrays)
  uint32_t* cdst = restore ? (uint32_t*)b.sys : nullptr;
  const uint32_t* csrc = (const uint32_t*)restore;
  const int64_t cwords = restore ? (int64_t)(off_bounds() / 4) : 0;
  if (fused_build) {
    // one launch builds every system (one block each): plan, sort, bin
    // offsets, sorted positions, work list, zero counts of idle items.
    // With `restore` (a captured plan) it plans from the device copy of the
    // descriptors and clears the scan's words itself: no launch before it.
    // The work list's lengths are zero then already (the scan of the
    // previous replay cleared them, or the plan's set-up did).
    if (!restore) NL_TRY(launch_zero2(b.zero, z_words, nullptr, 0, st));
    BuildArgs ba{};
    ba.pos = b.pos;
    ba.sys = b.sys;
    ba.band = b.band;
    ba.rc2 = b.rc2;
    ba.prm = prm;
    ba.item_base = b.items;
    ba.n_items = n_items;
    ba.group = tune.group;
    ba.parts = tune.parts;
    ba.keys = b.keysB;
    ba.vals = b.valsB;
    ba.cell_start = b.cell_start;
    ba.spos64 = b.spos64;
    ba.spos32 = b.spos32;
    ba.sidx = b.sidx;
    ba.list = b.list;
    ba.n_real = tickets + 6;
    ba.counts = b.counts;
    ba.sys_in = static_cast<const SysDesc*>(restore);
    ba.zero = restore ? b.zero : nullptr;
    ba.n_zero = restore ? (int64_t)(tickets + 6 - b.zero) : 0;
    NL_TRY(launch_build_systems(ba, S, st));
  } else {
    // zero block and counts (items without work keep zero counts)
    NL_TRY(launch_zero2(b.zero, z_words, b.counts, scan_len, st, cdst, csrc, cwords));
    NL_TRY(enqueue_build_stages(b, st));
  }
  PairArgs A = pair_args(*this, b);
  for (int q0 = 0; q0 < K; q0 += kGroup) {
    A.q_lo = q0;
    A.q_cnt = std::min(kGroup, K - q0);
    NL_TRY(launch_pairs(false, A, st));
  }
  // the scan also hands the work list's lengths on to the fill (tickets
  // 8-9) and clears them (tickets 6-7)
  NL_TRY(launch_scan(b.counts, b.offsets, scan_len, n_items, K, b.bounds, sstatus, tickets + 4,
                     st, tickets + 6));
  return hipSuccess;
}

// The build as separate launches (global radix sort, or Morton-ranked bins):
// plan, bins, sort, bin offsets and sorted positions, work list.
hipError_t Pipeline::enqueue_build_stages(const Buffers& b, hipStream_t st) const {
  uint64_t* rstatus = (uint64_t*)(b.zero + 2 * stiles);
  uint32_t* tickets = b.zero + 2 * stiles + 2 * rstiles * passes;
  NL_TRY(launch_plan(b.pos, b.sys, b.band, b.rc2, prm, st));
  if (b.bin_rank) NL_TRY(launch_morton(b.sys, S, bin_total, b.bin_rank, b.rank_bin, st));
  NL_TRY(launch_bin(b.pos, b.sys, b.start, S, n, b.bin_rank, b.keysA, b.valsA, st));
  uint32_t *kin = b.keysA, *vin = b.valsA, *kout = b.keysB, *vout = b.valsB;
  if (seg_sort) {
    NL_TRY(launch_sort_systems(kin, vin, kout, vout, b.sys, b.start, S, seg_passes,
                               b.cell_start, st));
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  for (int ps = 0; ps < (seg_sort ? 0 : passes); ++ps) {
    NL_TRY(launch_radix_pass(kin, vin, kout, vout, n, digit_bits * ps, digit_bits, b.rcnt,
                             b.roff, b.rbounds, rstatus + rstiles * ps, tickets + ps, st));
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  // (the segmented sort has written the bin offsets already)
  NL_TRY(launch_finalize(kin, vin, b.pos, b.sys, b.start, S, n, bin_total, b.cell_start,
                         b.spos64, b.spos32, b.sidx, !seg_sort, st));
  // work list of the items that have work
  NL_TRY(launch_items(kin, b.cell_start, b.sys, b.items, S, n_items, tune.group, tune.parts,
                      b.list, tickets + 6, st));
  return hipSuccess;
}

hipError_t Pipeline::enqueue_fill(const Buffers& b, void* edges, void* shifts, void* dist,
                                  int64_t capacity, hipStream_t st, void* disp) const {
  PairArgs A = pair_args(*this, b);
  A.n_real += 2;  // the lengths as the count scan handed them on
  A.edges = (int64_t*)edges;
  A.shifts = (int32_t*)shifts;
  A.dist = (double*)dist;
  A.disp = (double*)disp;
  A.capacity = capacity;
  for (int q0 = 0; q0 < K; q0 += kGroup) {
    A.q_lo = q0;
    A.q_cnt = std::min(kGroup, K - q0);
    NL_TRY(launch_pairs(true, A, st));
  }
  if (dist) NL_TRY(launch_sqrt_dist((double*)dist, b.bounds + K, capacity, st));
  return hipSuccess;
}

hipError_t Pipeline::enqueue_bounds(const Buffers& b, char* pin, hipStream_t st) const {
  return hipMemcpyAsync(pin + off_bounds(), b.bounds, 8 * (K + 1), hipMemcpyDeviceToHost, st);
}

// ---------------------------------------------------------------------------
// Engine

enum { kRangesSlot = kNumBuffers, kSmallBatchSlot, kSlots };

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

// Pooled device blocks: the smallest one of at least `bytes` (at most
// twice that), preferring blocks whose consumer fence has completed; a
// block still fenced is waited for only when no other fits.
void* Engine::dev_alloc(size_t bytes) {
  bytes = (std::max<size_t>(bytes, 256) + 255) & ~size_t(255);
  std::shared_ptr<Fence> wait;
  void* got = nullptr;
  {
    std::lock_guard<std::mutex> g(pool_mu_);
    auto pick = pool_.end();
    for (auto it = pool_.lower_bound(bytes); it != pool_.end() && it->first <= 2 * bytes; ++it) {
      const Fence* f = it->second.fence.get();
      if (!f || !f->ev || hipEventQuery(f->ev) != hipErrorNotReady) {
        pick = it;
        break;
      }
      if (pick == pool_.end()) pick = it;
    }
    if (pick != pool_.end()) {
      got = pick->second.p;
      wait = pick->second.fence;
      pool_.erase(pick);
    }
  }
  if (got) {
    if (wait && wait->ev) (void)hipEventSynchronize(wait->ev);
    return got;
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    // release cached blocks (hipFree waits for the device) and retry once
    std::lock_guard<std::mutex> g(pool_mu_);
    for (auto& kv : pool_) (void)hipFree(kv.second.p);
    pool_.clear();
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  }
  return p;
}

void Engine::dev_release(void* p, size_t bytes, std::shared_ptr<Fence> fence) {
  if (!p) return;
  bytes = (std::max<size_t>(bytes, 256) + 255) & ~size_t(255);
  std::lock_guard<std::mutex> g(pool_mu_);
  pool_.emplace(bytes, PoolBlock{p, std::move(fence)});
}

int Engine::alloc_outputs(Result* res, int64_t edges, bool dist, bool disp) {
  if (res->edges) dev_release(res->edges, res->edges_bytes);
  if (res->shifts) dev_release(res->shifts, res->shifts_bytes);
  if (res->dist) dev_release(res->dist, res->dist_bytes);
  if (res->disp) dev_release(res->disp, res->disp_bytes);
  res->edges = res->shifts = res->dist = res->disp = nullptr;
  const int64_t m = std::max<int64_t>(edges, 1);
  res->edges_bytes = sizeof(int64_t) * 2 * m;
  res->shifts_bytes = sizeof(int32_t) * 3 * m;
  res->edges = dev_alloc(res->edges_bytes);
  res->shifts = dev_alloc(res->shifts_bytes);
  if (dist) {
    res->dist_bytes = sizeof(double) * m;
    res->dist = dev_alloc(res->dist_bytes);
  }
  if (disp) {
    res->disp_bytes = sizeof(double) * 3 * m;
    res->disp = dev_alloc(res->disp_bytes);
  }
  if (!res->edges || !res->shifts || (dist && !res->dist) || (disp && !res->disp)) {
    set_error("out of device memory for the edge lists");
    return 1;
  }
  return 0;
}

#define NL_CHECK(call)                            \
  do {                                            \
    hipError_t e_ = (call);                       \
    if (e_ != hipSuccess) return fail(e_, #call); \
  } while (0)

int Engine::run(const Request& req, Result** out) {
  *out = nullptr;
  return run_with(req, tuning_for(req), out);
}

int Engine::run_with(const Request& req, const Tuning& t, Result** out) {
  *out = nullptr;
  const int K = (int)req.cutoffs.size();
  std::unique_ptr<Result> res(new Result());
  res->eng = this;
  res->k = K;
  res->len.assign(K, 0);
  res->base.assign(K, 0);
  res->tune = t;
  res->full = req.full;
  res->ustream = req.has_ustream ? req.ustream : nullptr;
  if (req.n == 0 || req.sys.empty()) {
    *out = res.release();
    return 0;
  }
  (void)hipSetDevice(device_);
  Pipeline P;
  if (P.prepare(req, t)) return 1;
  if (small_eligible(req, P)) {
    const int r = run_small(req, P, res.get());
    if (r != 2) {
      if (r == 0) *out = res.release();
      return r;
    }
    // the plan took the cell-list path: the pipeline below decides alike
  } else if (small_batch_eligible(req, P)) {
    const int r = run_small_batch(req, P, res.get());
    if (r != 2) {
      if (r == 0) *out = res.release();
      return r;
    }
  }
  const int S = P.S;
  std::vector<size_t> sz = P.buffer_sizes();
  std::vector<void*> ptr(kNumBuffers);
  for (int i = 0; i < kNumBuffers; ++i) {
    if (i == 0 && P.pos_device) {
      ptr[0] = const_cast<double*>(req.pos);
      continue;
    }
    ptr[i] = scratch(i, sz[i]);
    if (!ptr[i]) {
      set_error("out of device memory");
      return 1;
    }
  }
  Buffers B;
  P.assign(B, ptr);
  char* pin = (char*)pinned(P.pinned_bytes());
  if (!pin) {
    set_error("out of pinned host memory");
    return 1;
  }
  P.stage_host(pin);

  hipStream_t st = stream_;
  if (req.has_ustream) {
    // device-resident inputs are read after the caller's queued work
    NL_CHECK(hipEventRecord(ev_[5], req.ustream));
    NL_CHECK(hipStreamWaitEvent(st, ev_[5], 0));
  }
  NL_CHECK(P.enqueue_upload(B, pin, st));
  if (timing_) NL_CHECK(hipEventRecord(ev_[0], st));
  NL_CHECK(P.enqueue_build_count(B, st));
  if (timing_) NL_CHECK(hipEventRecord(ev_[2], st));

  // Fill straight away when an estimate of the output size exists (same
  // request shape as last time); the kernel writes nothing if it is short.
  const int64_t sig[4] = {P.n, S, K, (int64_t)req.full * 4 + (int64_t)req.multi * 2 + req.disjoint};
  const bool have_est = last_total_ >= 0 && std::equal(sig, sig + 4, last_sig_);
  const int64_t cap = have_est ? last_total_ + last_total_ / 8 + 4096 : -1;
  auto launch_fill = [&](int64_t capacity) -> int {
    if (timing_) NL_CHECK(hipEventRecord(ev_[3], st));
    NL_CHECK(P.enqueue_fill(B, res->edges, res->shifts, res->dist, capacity, st, res->disp));
    if (timing_) NL_CHECK(hipEventRecord(ev_[4], st));
    return 0;
  };
  if (have_est) {
    if (alloc_outputs(res.get(), cap, req.want_dist, req.want_disp)) return 1;
    if (launch_fill(cap)) return 1;
  }
  NL_CHECK(P.enqueue_bounds(B, pin, st));
  NL_CHECK(hipStreamSynchronize(st));
  const int64_t* hb = P.host_bounds(pin);
  const int64_t total = hb[K];
  res->total = total;
  if (!have_est || total > cap) {
    if (alloc_outputs(res.get(), total, req.want_dist, req.want_disp)) return 1;
    if (total > 0 && launch_fill(total)) return 1;
  }
  for (int q = 0; q < 4; ++q) last_sig_[q] = sig[q];
  last_total_ = total;
  for (int q = 0; q < K; ++q) {
    const int r = P.rank_of(q);
    res->base[q] = hb[r];
    res->len[q] = hb[r + 1] - hb[r];
  }
  if (req.want_ranges) {
    int64_t* d_ranges = (int64_t*)scratch(kRangesSlot, 8 * 2 * (int64_t)S * K);
    if (!d_ranges) {
      set_error("out of device memory");
      return 1;
    }
    NL_CHECK(launch_sys_ranges(B.offsets, B.items, S, K, P.n_items, d_ranges, st));
    std::vector<int64_t> rr((size_t)2 * S * K);
    NL_CHECK(hipMemcpyAsync(rr.data(), d_ranges, 8 * rr.size(), hipMemcpyDeviceToHost, st));
    std::vector<SysDesc> dd(S);
    NL_CHECK(hipMemcpyAsync(dd.data(), B.sys, sizeof(SysDesc) * S, hipMemcpyDeviceToHost, st));
    NL_CHECK(hipStreamSynchronize(st));
    res->ranges.assign((size_t)2 * S * K, 0);
    res->modes.resize(S);
    for (int s = 0; s < S; ++s) {
      res->modes[s] = dd[s].mode;
      for (int q = 0; q < K; ++q) {
        // the scan's positions run over all lists: relative to list q's start
        const int r = P.rank_of(q);
        res->ranges[2 * ((size_t)s * K + q)] = rr[2 * ((size_t)s * K + r)] - hb[r];
        res->ranges[2 * ((size_t)s * K + q) + 1] = rr[2 * ((size_t)s * K + r) + 1] - hb[r];
      }
    }
  }
  NL_CHECK(hipStreamSynchronize(st));
  if (timing_) {
    float tm = 0;
    (void)hipEventElapsedTime(&tm, ev_[0], ev_[2]);
    times_.build = 0;
    times_.count = tm;  // build + count pass + scan
    times_.fill = 0;
    if (total > 0) {
      (void)hipEventElapsedTime(&tm, ev_[3], ev_[4]);
      times_.fill = tm;
    }
    (void)hipEventElapsedTime(&tm, ev_[0], total > 0 ? ev_[4] : ev_[2]);
    times_.total = tm;
    times_.units = P.n_items;
    times_.edges = total;
    times_.bins = P.bin_total;
  }
  *out = res.release();
  return 0;
}

// ---------------------------------------------------------------------------
// Single small systems below the brute-force threshold, without the
// pipeline: one launch of k_small (one block does every row) up to
// kSmallAuto atoms, two launches (count over many blocks with the scan in the
// last one, then fill) up to kSmallRows.  The plan inside the kernels is the
// pipeline's, so the host only screens: one system, at most kGroup lists, no
// per-system ranges and, with a cell, widths that leave the cutoff clear of
// the minimum-image bound (otherwise the pipeline runs at once rather than
// after a declined launch).  Measured per call against the pipeline's
// 78-95 us (rc 6, device positions): one block 31 us at 9 atoms and 36 us at
// 100; two launches 40 us at 9 atoms, 44-47 at 100-200, 52-62 at 500, 64-81
// at 999.  One block wins while its pair work is small: up to 128 atoms on
// raw positions, about 74 with the minimum-image branch (about three times
// the work per pair; taken with a cell and under the multi-cutoff rules).
// Modes 2 and 3 force one form for every single system up to kSmallRows
// atoms, skipping the host's screen (tests).
constexpr int64_t kSmallAuto = 128;

bool Engine::small_eligible(const Request& req, const Pipeline& P) const {
  const int on = config().small_path.load();
  if (on == 0 || P.S != 1 || P.n < 1 || P.n > kSmallRows || P.K > kGroup || req.want_ranges)
    return false;
  if ((uint64_t)P.n >= (uint64_t)config().bf_threshold.load()) return false;
  const SysDesc& d = P.desc[0];
  if (!d.has_cell || on >= 2) return true;
  double inv[9], w[3];
  if (!inverse3(d.h, inv)) return false;
  widths3(inv, w);
  return d.rc_max * 2.0 < 0.999 * std::fmin(std::fmin(w[0], w[1]), w[2]);
}

// Device block of the small paths (allocated and zeroed once): bounds,
// band, the count launch's ticket, the planned descriptor, counts, offsets.
enum : size_t {
  kSmBounds = 0,
  kSmBand = 128,
  kSmTicket = 192,
  kSmBatchFlag = 224,  // cell-list flag of the batched form
  kSmPlanned = 256,
  kSmCounts = 768,
  kSmOffs = kSmCounts + 4 * (size_t)kGroup * kSmallRows,
  kSmBytes = kSmOffs + 4 * (size_t)kGroup * kSmallRows,
};
static_assert(sizeof(SysDesc) <= kSmCounts - kSmPlanned, "descriptor slot");

int Engine::ensure_small_buf() {
  if (small_buf_) return 0;
  if (hipMalloc(&small_buf_, kSmBytes) != hipSuccess) {
    small_buf_ = nullptr;
    set_error("out of device memory");
    return 1;
  }
  // the tickets start at zero (on the engine's stream: a blocking hipMemset
  // would put the null stream on one of the runtime's hardware queues)
  NL_CHECK(hipMemsetAsync(small_buf_, 0, kSmBytes, stream_));
  NL_CHECK(hipStreamSynchronize(stream_));
  return 0;
}

// 0: done, 1: error, 2: the plan took the cell-list path (nothing written)
int Engine::run_small(const Request& req, const Pipeline& P, Result* res) {
  const int K = P.K;
  const int on = config().small_path.load();
  const int64_t work = P.n * P.n * ((req.multi || P.desc[0].has_cell) ? 3 : 1);
  const bool one_block = on == 2 || (on != 3 && work <= kSmallAuto * kSmallAuto);
  double* dpos = const_cast<double*>(req.pos);
  if (!P.pos_device) {
    dpos = (double*)scratch(0, sizeof(double) * 3 * (size_t)P.n);
    if (!dpos) {
      set_error("out of device memory");
      return 1;
    }
  }
  if (ensure_small_buf()) return 1;
  // page-locked block: the bounds read back, then the host positions (at
  // most 24 KB: one memcpy and an asynchronous upload instead of the
  // runtime's staged copy of pageable memory)
  constexpr size_t kPinPos = 8 * (kGroup + 1) + 56;
  const size_t pos_bytes = sizeof(double) * 3 * (size_t)P.n;
  char* pin = (char*)pinned(kPinPos + (P.pos_device ? 0 : pos_bytes));
  if (!pin) {
    set_error("out of pinned host memory");
    return 1;
  }
  int64_t* hb = (int64_t*)pin;
  char* sb = (char*)small_buf_;
  SmallArgs a{};
  a.pos = dpos;
  a.desc = P.desc[0];
  for (int q = 0; q < K; ++q) a.rc2[q] = P.rc2[q];
  a.prm = P.prm;
  a.bounds = (int64_t*)(sb + kSmBounds);
  a.band = (float2*)(sb + kSmBand);
  a.ticket = (uint32_t*)(sb + kSmTicket);
  a.planned = (SysDesc*)(sb + kSmPlanned);
  a.counts = (uint32_t*)(sb + kSmCounts);
  a.offs = (uint32_t*)(sb + kSmOffs);
  a.k = K;
  a.disjoint = P.disjoint;
  a.full = P.full;
  hipStream_t st = stream_;
  if (req.has_ustream) {
    NL_CHECK(hipEventRecord(ev_[5], req.ustream));
    NL_CHECK(hipStreamWaitEvent(st, ev_[5], 0));
  }
  if (!P.pos_device) {
    std::memcpy(pin + kPinPos, req.pos, pos_bytes);
    NL_CHECK(hipMemcpyAsync(dpos, pin + kPinPos, pos_bytes, hipMemcpyHostToDevice, st));
  }
  // room from the last request of the same shape, else a small default;
  // a larger total is counted and the fill repeated with exact room
  const int64_t sig[4] = {P.n, 1, K, (int64_t)req.full * 4 + (int64_t)req.multi * 2 + req.disjoint};
  const bool have_est = last_total_ >= 0 && std::equal(sig, sig + 4, last_sig_);
  // (never more than the brute force can find: n (n - 1) / 2 pairs per
  // list, one image each, both directions in full lists)
  const int64_t most = std::max<int64_t>(1, P.n * (P.n - 1) / 2 * K * (req.full ? 2 : 1));
  const int64_t cap = std::min(most, (have_est ? last_total_ + last_total_ / 8 : 0) + 4096);
  // host callers read the lists back: a page-locked copy of small lists is
  // made before the synchronisation, within the same stream.  It is sized
  // by the last total of the same shape and dropped if the lists outgrow
  // it.  It pays for lists of a few hundred edges (measured per numpy call:
  // 9 atoms, 36 edges, 71 -> 43 us); at 2,131 edges (60 KB) it cost 7 us
  // more than copying into the caller's arrays after the synchronisation.
  constexpr int64_t kMirrorMax = int64_t(8) << 10;
  auto launch = [&](int64_t room, bool count) -> int {
    if (alloc_outputs(res, room, req.want_dist, req.want_disp)) return 1;
    a.edges = (int64_t*)res->edges;
    a.shifts = (int32_t*)res->shifts;
    a.dist = (double*)res->dist;
    a.disp = (double*)res->disp;
    a.capacity = room;
    if (timing_) NL_CHECK(hipEventRecord(ev_[0], st));
    if (one_block) {
      NL_CHECK(launch_small(a, st));
    } else {
      if (count) NL_CHECK(launch_brute(false, a, st));
      NL_CHECK(launch_brute(true, a, st));
      if (a.dist) NL_CHECK(launch_sqrt_dist(a.dist, a.bounds + K, room, st));
    }
    if (timing_) NL_CHECK(hipEventRecord(ev_[4], st));
    if (res->mirror) nl_host_free(res->mirror);
    res->mirror = nullptr;
    res->mirror_cap = 0;
    const int64_t want = count && have_est ? std::min(room, last_total_ + 64) : room;
    if (!P.pos_device && 28 * want <= kMirrorMax) {
      char* m = (char*)nl_host_alloc(28 * want);
      if (m) {
        NL_CHECK(hipMemcpyAsync(m, a.edges, 16 * want, hipMemcpyDeviceToHost, st));
        NL_CHECK(hipMemcpyAsync(m + 16 * want, a.shifts, 12 * want, hipMemcpyDeviceToHost, st));
        res->mirror = m;
        res->mirror_cap = want;
      }
    }
    NL_CHECK(hipMemcpyAsync(hb, a.bounds, 8 * (K + 1), hipMemcpyDeviceToHost, st));
    NL_CHECK(hipStreamSynchronize(st));
    return 0;
  };
  if (launch(cap, true)) return 1;
  if (hb[K] < 0) return 2;
  const int64_t total = hb[K];
  if (total > cap && launch(total, false)) return 1;
  if (res->mirror && total > res->mirror_cap) {
    nl_host_free(res->mirror);  // the lists outgrew the copy: read the device
    res->mirror = nullptr;
    res->mirror_cap = 0;
  }
  res->total = total;
  for (int q = 0; q < K; ++q) {
    const int r = P.rank_of(q);
    res->base[q] = hb[r];
    res->len[q] = hb[r + 1] - hb[r];
  }
  for (int q = 0; q < 4; ++q) last_sig_[q] = sig[q];
  last_total_ = total;
  if (timing_) {
    float tm = 0;
    (void)hipEventElapsedTime(&tm, ev_[0], ev_[4]);
    times_ = StageTimes{};
    times_.count = times_.total = tm;
    times_.units = P.n;
    times_.edges = total;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Batches of small systems that all take a brute-force path, without the
// pipeline: three launches (count, a one-block scan, fill) instead of the
// pipeline's dozen, whose binning, sorting and work-list stages do nothing
// for brute-force systems.  The host screens as for single
// systems (every system below the brute-force threshold and at most
// kSmallRows atoms; with a cell, widths that leave the cutoff clear of the
// minimum-image bound); the plan inside the kernels decides, and a batch in
// which some system takes the cell-list path goes to the pipeline.  One block
// scans the [list][atom] counts, so the batch is limited to kBatchSmallRows
// of them.
constexpr int64_t kBatchSmallRows = int64_t(1) << 16;

bool Engine::small_batch_eligible(const Request& req, const Pipeline& P) const {
  const int on = config().small_path.load();
  if (on == 0 || P.S < 2 || P.n < 1 || P.K > kGroup || req.want_ranges || P.per_sys ||
      (int64_t)P.K * P.n > kBatchSmallRows)
    return false;
  const uint64_t thr = config().bf_threshold.load();
  int64_t most = 0;
  for (const SysDesc& d : P.desc) {
    if (d.n_atoms > kSmallRows || (uint64_t)d.n_atoms >= thr) return false;
    if (d.has_cell && on < 2) {  // modes 2 and 3 leave the screen to the plan (tests)
      double inv[9], w[3];
      if (!inverse3(d.h, inv)) return false;
      widths3(inv, w);
      if (!(d.rc_max * 2.0 < 0.999 * std::fmin(std::fmin(w[0], w[1]), w[2]))) return false;
    }
    most += (int64_t)d.n_atoms * (d.n_atoms - 1) / 2;
  }
  // offsets are 32-bit: never more than the brute force can find
  return most * P.K * (req.full ? 2 : 1) < (int64_t(1) << 31);
}

// 0: done, 1: error, 2: some system took the cell-list path (nothing written)
int Engine::run_small_batch(const Request& req, const Pipeline& P, Result* res) {
  const int K = P.K, S = P.S;
  const int64_t n = P.n;
  // block table: first block of each system, then the system of each block
  int64_t most = 0, blocks64 = 0;
  for (int s = 0; s < S; ++s) {
    const int64_t ns = P.desc[s].n_atoms;
    blocks64 += (ns + kBatchRows - 1) / kBatchRows;
    most += ns * (ns - 1) / 2;
  }
  const int blocks = (int)blocks64;
  std::vector<int32_t> blk((size_t)S + 1 + (size_t)blocks, 0);
  for (int s = 0; s < S; ++s) {
    const int32_t nb = (int32_t)((P.desc[s].n_atoms + kBatchRows - 1) / kBatchRows);
    blk[s + 1] = blk[s] + nb;
    for (int32_t b = blk[s]; b < blk[s + 1]; ++b) blk[(size_t)S + 1 + b] = s;
  }
  most = std::max<int64_t>(1, most * K * (req.full ? 2 : 1));
  // device block: descriptors and block table (one upload), planned
  // descriptors, bands, counts, offsets, bounds
  auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
  const size_t o_blk = sizeof(SysDesc) * S;
  const size_t blk_bytes = 4 * blk.size();
  const size_t o_planned = up(o_blk + blk_bytes);
  const size_t o_band = up(o_planned + sizeof(SysDesc) * S);
  const size_t o_counts = up(o_band + sizeof(float2) * (size_t)S * K);
  const size_t o_offs = up(o_counts + 4 * (size_t)K * n);
  const size_t o_bounds = up(o_offs + 4 * (size_t)K * n);
  const size_t dev_bytes = o_bounds + 8 * ((size_t)K + 1);
  char* dv = (char*)scratch(kSmallBatchSlot, dev_bytes);
  double* dpos = const_cast<double*>(req.pos);
  if (!P.pos_device) dpos = (double*)scratch(0, sizeof(double) * 3 * (size_t)n);
  if (!dv || !dpos) {
    set_error("out of device memory");
    return 1;
  }
  if (ensure_small_buf()) return 1;
  // page-locked block: the bounds read back, the upload, the host positions
  constexpr size_t kPinUp = 8 * (kGroup + 1) + 56;
  const size_t up_bytes = o_blk + blk_bytes;
  const size_t o_pinpos = up(kPinUp + up_bytes);
  const size_t pos_bytes = sizeof(double) * 3 * (size_t)n;
  char* pin = (char*)pinned(o_pinpos + (P.pos_device ? 0 : pos_bytes));
  if (!pin) {
    set_error("out of pinned host memory");
    return 1;
  }
  int64_t* hb = (int64_t*)pin;
  std::memcpy(pin + kPinUp, P.desc.data(), sizeof(SysDesc) * S);
  std::memcpy(pin + kPinUp + o_blk, blk.data(), blk_bytes);

  BatchSmallArgs a{};
  a.pos = dpos;
  a.sys = (const SysDesc*)dv;
  a.blk_start = (const int32_t*)(dv + o_blk);
  a.blk_sys = a.blk_start + S + 1;
  a.n_sys = S;
  a.n = n;
  for (int q = 0; q < K; ++q) a.rc2[q] = P.rc2[q];
  a.prm = P.prm;
  a.band = (float2*)(dv + o_band);
  a.bounds = (int64_t*)(dv + o_bounds);
  a.k = K;
  a.disjoint = P.disjoint;
  a.full = P.full;
  a.counts = (uint32_t*)(dv + o_counts);
  a.offs = (uint32_t*)(dv + o_offs);
  a.planned = (SysDesc*)(dv + o_planned);
  a.flag = (uint32_t*)((char*)small_buf_ + kSmBatchFlag);

  hipStream_t st = stream_;
  if (req.has_ustream) {
    NL_CHECK(hipEventRecord(ev_[5], req.ustream));
    NL_CHECK(hipStreamWaitEvent(st, ev_[5], 0));
  }
  if (!P.pos_device) {
    std::memcpy(pin + o_pinpos, req.pos, pos_bytes);
    NL_CHECK(hipMemcpyAsync(dpos, pin + o_pinpos, pos_bytes, hipMemcpyHostToDevice, st));
  }
  NL_CHECK(hipMemcpyAsync(dv, pin + kPinUp, up_bytes, hipMemcpyHostToDevice, st));
  // room from the last request of the same shape (the pipeline's signature:
  // both paths give the same lists), else a small default
  const int64_t sig[4] = {P.n, S, K, (int64_t)req.full * 4 + (int64_t)req.multi * 2 + req.disjoint};
  const bool have_est = last_total_ >= 0 && std::equal(sig, sig + 4, last_sig_);
  const int64_t cap = std::min(most, (have_est ? last_total_ + last_total_ / 8 : 0) + 4096);
  auto launch = [&](int64_t room, bool count) -> int {
    if (alloc_outputs(res, room, req.want_dist, req.want_disp)) return 1;
    a.edges = (int64_t*)res->edges;
    a.shifts = (int32_t*)res->shifts;
    a.dist = (double*)res->dist;
    a.disp = (double*)res->disp;
    a.capacity = room;
    if (timing_) NL_CHECK(hipEventRecord(ev_[0], st));
    if (count) NL_CHECK(launch_batch_brute(false, a, blocks, st));
    NL_CHECK(launch_batch_brute(true, a, blocks, st));
    if (a.dist) NL_CHECK(launch_sqrt_dist(a.dist, a.bounds + K, room, st));
    if (timing_) NL_CHECK(hipEventRecord(ev_[4], st));
    NL_CHECK(hipMemcpyAsync(hb, a.bounds, 8 * (K + 1), hipMemcpyDeviceToHost, st));
    NL_CHECK(hipStreamSynchronize(st));
    return 0;
  };
  if (launch(cap, true)) return 1;
  if (hb[K] < 0) return 2;
  const int64_t total = hb[K];
  if (total > cap && launch(total, false)) return 1;
  res->total = total;
  for (int q = 0; q < K; ++q) {
    const int r = P.rank_of(q);
    res->base[q] = hb[r];
    res->len[q] = hb[r + 1] - hb[r];
  }
  for (int q = 0; q < 4; ++q) last_sig_[q] = sig[q];
  last_total_ = total;
  if (timing_) {
    float tm = 0;
    (void)hipEventElapsedTime(&tm, ev_[0], ev_[4]);
    times_ = StageTimes{};
    times_.count = times_.total = tm;
    times_.units = n;
    times_.edges = total;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Autotuner: per workload class (lists, multi, log2 atoms, log2 systems,
// log2 max cutoff) a staged search over bin size, home-group size, stencil
// slices and the FP32 filter, timing full launches.  None of the knobs can
// change a result, only the time it takes.

static int lg2(double x) { return x <= 1.0 ? 0 : (int)std::floor(std::log2(x)); }

static std::tuple<int, int, int, int, int> tune_key(const Request& req) {
  double rc_max = 0;
  for (double x : req.cutoffs) rc_max = std::max(rc_max, x);
  return std::make_tuple((int)req.cutoffs.size(), (int)req.multi, lg2((double)req.n),
                         lg2((double)req.sys.size()), lg2(rc_max));
}

static Tuning explicit_tuning(const Request& req) {
  Tuning t;
  Config& c = config();
  t.cell_div = req.cell_div > 0 ? req.cell_div : c.cell_div.load();
  t.use_fp32 = req.use_fp32 >= 0 ? req.use_fp32 : c.use_fp32.load();
  t.group = req.group > 0 ? req.group : c.group.load();
  t.parts = req.parts > 0 ? req.parts : c.parts.load();
  return t;
}

static bool is_explicit(const Request& req) {
  return req.cell_div > 0 || req.use_fp32 >= 0 || req.group > 0 || req.parts > 0;
}

Tuning untimed_tuning(const Request& req) {
  Tuning t = explicit_tuning(req);
  if (config().autotune.load() && !is_explicit(req) && req.n < kTuneMinAtoms) {
    // too small to fill the GPU (and to be worth timing): the launch's
    // latency is its longest work item, so more, smaller items -- two home
    // atoms per group and four stencil slices (measured 1.5x faster than
    // 4 x 1 per single-system call on systems of 8 to 2048 atoms)
    t.group = 2;
    t.parts = 4;
  }
  return t;
}

Tuning Engine::tuning_for(const Request& req) {
  Tuning t = explicit_tuning(req);
  Config& c = config();
  if (!c.autotune.load() || is_explicit(req)) return t;
  if (req.n < kTuneMinAtoms) return untimed_tuning(req);
  const auto key = tune_key(req);
  {
    std::lock_guard<std::mutex> g(tune_mu_);
    auto it = tuned_.find(key);
    if (it != tuned_.end()) return it->second;
  }
  Request probe = req;
  probe.want_ranges = false;
  probe.want_dist = false;
  probe.want_disp = false;
  // device time of the search stages (hipEvents around the launches), best
  // of `reps` runs
  auto time_one = [&](const Tuning& x, int reps) -> double {
    double best = 1e30;
    const bool was = timing_;
    timing_ = true;
    for (int rep = 0; rep < reps; ++rep) {
      Result* r = nullptr;
      if (run_with(probe, x, &r) != 0) {
        timing_ = was;
        return 1e30;
      }
      delete r;
      best = std::min(best, (double)times_.total);
    }
    timing_ = was;
    return best;
  };
  // bin size first (a candidate must beat the incumbent by 2%, so timing
  // noise does not walk the choice away from the default; 0 = each system
  // by its density) ...
  Tuning best = t;
  double tb = time_one(best, 3);
  for (int dv : {0, 1, 2, 3, 4}) {
    if (dv == t.cell_div) continue;
    Tuning x = best;
    x.cell_div = dv;
    const double tx = time_one(x, 3);
    if (tx < 0.98 * tb) {
      tb = tx;
      best = x;
    }
  }
  // ... then home-group size and stencil slices jointly (they interact: more
  // slices suit larger groups), FP32 filter on and off for the best pair ...
  std::vector<std::pair<double, Tuning>> cand;
  for (int g : {2, 4, 8, 16, 32})
    for (int p : {1, 2, 4, 8}) {
      Tuning x = best;
      x.group = g;
      x.parts = p;
      cand.emplace_back(time_one(x, 3), x);
    }
  auto by_time = [](const std::pair<double, Tuning>& a, const std::pair<double, Tuning>& b) {
    return a.first < b.first;
  };
  std::sort(cand.begin(), cand.end(), by_time);
  // ... the bin size again under the two best work splits (the first pass
  // chose it under the default split; the best bin size depends on the
  // split) ...
  for (int c = 0; c < 2; ++c) {
    const Tuning base = cand[c].second;
    for (int dv : {0, 1, 2, 3, 4}) {
      if (dv == base.cell_div) continue;
      Tuning x = base;
      x.cell_div = dv;
      cand.emplace_back(time_one(x, 3), x);
    }
  }
  std::sort(cand.begin(), cand.end(), by_time);
  {
    Tuning x = cand[0].second;
    x.use_fp32 = 1 - x.use_fp32;
    cand.insert(cand.begin() + 1, {time_one(x, 3), x});
  }
  // ... and the leading three re-timed in turn, best of seven each, three
  // rounds (they are often within a few per cent of each other)
  const int top = std::min<int>(3, (int)cand.size());
  std::vector<double> tt(top, 1e30);
  for (int round = 0; round < 3; ++round)
    for (int i = 0; i < top; ++i) tt[i] = std::min(tt[i], time_one(cand[i].second, 7));
  int bi = 0;
  for (int i = 1; i < top; ++i)
    if (tt[i] < tt[bi]) bi = i;
  best = cand[bi].second;
  std::lock_guard<std::mutex> g(tune_mu_);
  tuned_[key] = best;
  return best;
}

Tuning Engine::tuned_or_untimed(const Request& req) {
  if (config().autotune.load() && !is_explicit(req) && req.n >= kTuneMinAtoms) {
    std::lock_guard<std::mutex> g(tune_mu_);
    auto it = tuned_.find(tune_key(req));
    if (it != tuned_.end()) return it->second;
  }
  return untimed_tuning(req);
}

std::string Engine::tune_report() {
  std::lock_guard<std::mutex> g(tune_mu_);
  std::ostringstream o;
  for (auto& kv : tuned_) {
    const auto& k = kv.first;
    const Tuning& t = kv.second;
    o << "k=" << std::get<0>(k) << " multi=" << std::get<1>(k) << " log2n=" << std::get<2>(k)
      << " log2sys=" << std::get<3>(k) << " log2rc=" << std::get<4>(k)
      << " -> cell_div=" << t.cell_div << " group=" << t.group << " parts=" << t.parts
      << " fp32=" << t.use_fp32 << "\n";
  }
  return o.str();
}

void Engine::clear_tuning() {
  std::lock_guard<std::mutex> g(tune_mu_);
  tuned_.clear();
}

static std::mutex g_engines_mu;
static std::map<int, Engine*> g_engines;

Engine* engine_for_current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  std::lock_guard<std::mutex> g(g_engines_mu);
  auto it = g_engines.find(dev);
  if (it != g_engines.end()) return it->second;
  Engine* e = new Engine(dev);
  g_engines[dev] = e;
  return e;
}

Engine* existing_engine_for_current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  std::lock_guard<std::mutex> g(g_engines_mu);
  auto it = g_engines.find(dev);
  return it == g_engines.end() ? nullptr : it->second;
}

// ---------------------------------------------------------------------------
// Plan: graph-captured steady state

Plan::~Plan() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  for (void* p : owned_)
    if (p) (void)hipFree(p);
  free_out();
  if (pinned_) (void)hipHostFree(pinned_);
  if (pin_ranges_) (void)hipHostFree(pin_ranges_);
  if (st_) (void)hipStreamDestroy(st_);
}

void Plan::free_out() {
  for (void** p : {&edges_, &shifts_, &dist_, &disp_}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
}

int Plan::alloc_out(int64_t cap) {
  free_out();
  cap_ = std::max<int64_t>(cap, 1);
  if (hipMalloc(&edges_, 16 * (size_t)cap_) != hipSuccess ||
      hipMalloc(&shifts_, 12 * (size_t)cap_) != hipSuccess ||
      (want_dist_ && hipMalloc(&dist_, 8 * (size_t)cap_) != hipSuccess) ||
      (want_disp_ && hipMalloc(&disp_, 24 * (size_t)cap_) != hipSuccess)) {
    set_error("out of device memory for the plan's edge lists");
    return 1;
  }
  return 0;
}

int Plan::capture() {
  if (exec_) {
    (void)hipGraphExecDestroy(exec_);
    exec_ = nullptr;
  }
  if (graph_) {
    (void)hipGraphDestroy(graph_);
    graph_ = nullptr;
  }
  hipError_t e = hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return eng_->fail(e, "hipStreamBeginCapture");
  // positions are device-resident and the upload block never changes: the
  // first kernel restores the descriptors (which the build overwrites with
  // its plans) from the plan's device copy
  e = pipe_.enqueue_build_count(buf_, st_, d_static_);
  if (e == hipSuccess) e = pipe_.enqueue_fill(buf_, edges_, shifts_, dist_, cap_, st_, disp_);
  if (e == hipSuccess) e = pipe_.enqueue_bounds(buf_, pinned_, st_);
  if (e == hipSuccess && want_ranges_) {
    const size_t rb = 16 * (size_t)pipe_.S * pipe_.K;
    e = launch_sys_ranges(buf_.offsets, buf_.items, pipe_.S, pipe_.K, pipe_.n_items, d_ranges_,
                          st_);
    if (e == hipSuccess)
      e = hipMemcpyAsync(pin_ranges_, d_ranges_, rb, hipMemcpyDeviceToHost, st_);
    if (e == hipSuccess)
      e = hipMemcpyAsync(pin_ranges_ + rb, buf_.sys, sizeof(SysDesc) * pipe_.S,
                         hipMemcpyDeviceToHost, st_);
  }
  hipError_t e2 = hipStreamEndCapture(st_, &graph_);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0);
  if (e != hipSuccess) return eng_->fail(e, "graph capture");
  ++recaptures;
  return 0;
}

void Plan::read_ranges() {
  if (!want_ranges_) return;
  const int S = pipe_.S, K = pipe_.K;
  const int64_t* rr = (const int64_t*)pin_ranges_;
  const SysDesc* dd = (const SysDesc*)(pin_ranges_ + 16 * (size_t)S * K);
  ranges_.assign((size_t)2 * S * K, 0);
  modes_.resize(S);
  for (int s = 0; s < S; ++s) {
    modes_[s] = dd[s].mode;
    for (int q = 0; q < K; ++q) {
      // the scan's positions run over all lists: relative to list q's start
      const int r = pipe_.rank_of(q);
      ranges_[2 * ((size_t)s * K + q)] = rr[2 * ((size_t)s * K + r)] - bounds_[r];
      ranges_[2 * ((size_t)s * K + q) + 1] = rr[2 * ((size_t)s * K + r) + 1] - bounds_[r];
    }
  }
}

int Plan::create(Engine* eng, const Request& req, int64_t capacity) {
  eng_ = eng;
  (void)hipSetDevice(eng->device());
  if (!req.pos_device) {
    set_error("a plan needs device-resident positions");
    return 1;
  }
  const Tuning t = eng->tuning_for(req);
  if (pipe_.prepare(req, t)) return 1;
  want_dist_ = req.want_dist;
  want_disp_ = req.want_disp;
  want_ranges_ = req.want_ranges;
  std::vector<size_t> sz = pipe_.buffer_sizes();
  std::vector<void*> ptr(kNumBuffers, nullptr);
  ptr[0] = const_cast<double*>(req.pos);
  for (int i = 1; i < kNumBuffers; ++i) {
    if (hipMalloc(&ptr[i], std::max<size_t>(sz[i], 16)) != hipSuccess) {
      set_error("out of device memory for the plan");
      return 1;
    }
    owned_.push_back(ptr[i]);
  }
  pipe_.assign(buf_, ptr);
  if (hipHostMalloc((void**)&pinned_, pipe_.pinned_bytes(), hipHostMallocDefault) !=
      hipSuccess) {
    pinned_ = nullptr;
    set_error("out of pinned host memory");
    return 1;
  }
  pipe_.stage_host(pinned_);
  if (hipMalloc(&d_static_, std::max<size_t>(pipe_.off_bounds(), 16)) != hipSuccess) {
    set_error("out of device memory for the plan");
    return 1;
  }
  owned_.push_back(d_static_);
  if (want_ranges_) {
    const size_t rb = 16 * (size_t)pipe_.S * pipe_.K;
    if (hipMalloc((void**)&d_ranges_, std::max<size_t>(rb, 16)) != hipSuccess) {
      set_error("out of device memory for the plan");
      return 1;
    }
    owned_.push_back(d_ranges_);
    if (hipHostMalloc((void**)&pin_ranges_, rb + sizeof(SysDesc) * pipe_.S,
                      hipHostMallocDefault) != hipSuccess) {
      pin_ranges_ = nullptr;
      set_error("out of pinned host memory");
      return 1;
    }
  }
  if (hipStreamCreateWithFlags(&st_, hipStreamNonBlocking) != hipSuccess) {
    set_error("stream creation failed");
    return 1;
  }
  {
    // the device copy, the block in place (system starts, item bases and
    // cutoffs are only read: a replay restores the descriptors alone) and a
    // cleared zero block (a replay's one-launch build expects the work
    // list's lengths at zero)
    hipError_t e = hipMemcpyAsync(d_static_, pinned_, pipe_.off_bounds(),
                                  hipMemcpyHostToDevice, st_);
    if (e == hipSuccess) e = pipe_.enqueue_upload(buf_, pinned_, st_);
    if (e == hipSuccess) e = hipMemsetAsync(buf_.zero, 0, 4 * (size_t)pipe_.z_words, st_);
    if (e == hipSuccess) e = hipStreamSynchronize(st_);
    if (e != hipSuccess) return eng_->fail(e, "plan upload");
  }
  if (capacity <= 0) {
    // size the lists from one eager count
    hipError_t e = pipe_.enqueue_upload(buf_, pinned_, st_);
    if (e == hipSuccess) e = pipe_.enqueue_build_count(buf_, st_);
    if (e == hipSuccess) e = pipe_.enqueue_bounds(buf_, pinned_, st_);
    if (e == hipSuccess) e = hipStreamSynchronize(st_);
    if (e != hipSuccess) return eng_->fail(e, "plan sizing");
    const int64_t tot = pipe_.host_bounds(pinned_)[pipe_.K];
    capacity = tot + tot / 8 + 4096;
  }
  if (alloc_out(capacity)) return 1;
  return capture();
}

int Plan::run() {
  (void)hipSetDevice(eng_->device());
  hipError_t e = hipGraphLaunch(exec_, st_);
  if (e == hipSuccess) e = hipStreamSynchronize(st_);
  if (e != hipSuccess) return eng_->fail(e, "graph launch");
  ++replays;
  const int64_t* hb = pipe_.host_bounds(pinned_);
  total_ = hb[pipe_.K];
  if (total_ > cap_) {
    // overflow: the fill wrote nothing; grow, re-capture, replay once more
    if (alloc_out(total_ + total_ / 4 + 4096)) return 1;
    if (capture()) return 1;
    e = hipGraphLaunch(exec_, st_);
    if (e == hipSuccess) e = hipStreamSynchronize(st_);
    if (e != hipSuccess) return eng_->fail(e, "graph launch");
    ++replays;
    total_ = hb[pipe_.K];
  }
  bounds_.assign(hb, hb + pipe_.K + 1);
  read_ranges();
  return 0;
}

int64_t Plan::len(int list) const {
  const int r = pipe_.rank_of(list);
  return bounds_.empty() ? 0 : bounds_[r + 1] - bounds_[r];
}

int64_t Plan::base(int list) const {
  const int r = pipe_.rank_of(list);
  return bounds_.empty() ? 0 : bounds_[r];
}

}  // namespace nl
