// nl_server.cpp -- serving: continuous batching of single-system requests.
//
// Requests (positions, optional cell, one cutoff each) arrive at any time.
// Up to `slots` batches are in flight, each on its own HIP stream with its
// own device buffers.  Whenever a slot is free it admits the
// queued requests in arrival order into one ragged batch (per-system cutoffs,
// local indices), bounded by atoms and systems per batch and by an edge
// budget: every request carries an estimate of its edge count, and the sum
// over the batches on the device may not exceed the budget (a request alone
// is always admitted, so nothing starves).  Each slot runs the same launch
// sequence as the plain path (plan, bin, radix sort, finalize, count, scan,
// fill), so a result is the same whether a request ran alone or in a batch.
//
// Results stay on the device: a batch writes into an output block (edges
// and shifts of all its requests), and each request of the batch holds a
// reference to the block until the caller takes (copies out) or releases
// it.  Blocks are recycled through a pool, so the steady state allocates
// nothing.  Each slot has a host thread of its own, so the host work of one
// batch (staging, preparation, launches) overlaps the device work and host
// work of the others:
//   admit (FIFO, budget) -> launch -> counts known -> [refill if the
//   estimate was short] -> deliver
#include "nl_server.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "nl_engine.h"

namespace nl {

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Expected half-list size: half of N times the atoms within rc of an atom
// (number density times the sphere volume; images included for periodic
// cells, at most N-1 partners for an isolated cluster).  An isolated
// cluster's volume is its bounding box widened by rc (mn / mx: the box).
static int64_t estimate_from_box(int64_t n, const double* h, const double* mn, const double* mx,
                                 double rc) {
  double vol;
  if (h) {
    const double det = (h[0] * (h[4] * h[8] - h[7] * h[5]) - h[1] * (h[3] * h[8] - h[6] * h[5]) +
                        h[2] * (h[3] * h[7] - h[6] * h[4]));
    vol = std::fabs(det);
  } else {
    vol = 1.0;
    for (int k = 0; k < 3; ++k) vol *= (mx[k] - mn[k]) + rc;
  }
  const double rho = (double)n / std::max(vol, 1e-9);
  double per_atom = rho * 4.18879020478639 * rc * rc * rc;
  if (!h) per_atom = std::min(per_atom, (double)(n - 1));
  return (int64_t)(0.5 * (double)n * per_atom) + 16;
}

int64_t estimate_edges(const double* pos, int64_t n, const double* h, double rc) {
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  if (!h)
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        mn[k] = std::min(mn[k], pos[3 * i + k]);
        mx[k] = std::max(mx[k], pos[3 * i + k]);
      }
  return estimate_from_box(n, h, mn, mx, rc);
}

// A request's positions, copied at submit into a page-locked block of the
// host pool (nl_host_alloc): reused blocks take no page faults, and the slot
// thread's copy into the batch's staging buffer runs at memory speed.
// Ordinary memory when the pool has no block.
struct PosBuf {
  double* p = nullptr;
  int64_t n = 0;
  bool pinned = false;
  void reset() {
    if (p) {
      if (pinned)
        nl_host_free(p);
      else
        std::free(p);
    }
    p = nullptr;
  }
  ~PosBuf() { reset(); }
};

// Device output of one batch: int64 edges [2][cap], int32 shifts [cap][3],
// then (when the server emits them) FP64 distances [cap] and displacements
// [cap][3].  Every block of a server has the same layout.
struct OutBlock {
  char* dev = nullptr;
  int64_t cap = 0;    // edges the block can hold
  int64_t total = 0;  // edges of the batch written into it
  int refs = 0;
  bool dist = false, disp = false;
  static size_t bytes_per_edge(bool dist, bool disp) { return 28 + (dist ? 8 : 0) + (disp ? 24 : 0); }
  int64_t* edges() const { return (int64_t*)dev; }
  int32_t* shifts() const { return (int32_t*)(dev + 16 * (size_t)cap); }
  double* distances() const { return dist ? (double*)(dev + 28 * (size_t)cap) : nullptr; }
  double* displacements() const {
    return disp ? (double*)(dev + (28 + (dist ? 8 : 0)) * (size_t)cap) : nullptr;
  }
};

struct SrvReq {
  int64_t id = 0;
  PosBuf pos;  // freed once the batch holding the request has run
  const double* ext = nullptr;  // the caller's positions (borrowed), else pos.p
  int has_cell = 0;
  double h[9] = {0};
  uint8_t pbc[3] = {1, 1, 1};
  double cutoff = 0;
  double t_submit = 0, t_done = 0;
  int64_t est = 0;
  int state = 0;  // 0 queued, 1 running, 2 done, 3 failed, 4 being copied out
  std::string err;
  OutBlock* blk = nullptr;
  int64_t first = 0, len = 0;  // edge range within the block
  int batch_size = 0;
};

struct Server::Slot {
  int index = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  int state = 0;  // 0 free, 1 launched, 2 refilling
  Pipeline P;
  Buffers B;
  std::vector<void*> bufs;
  std::vector<size_t> caps;
  char* pin = nullptr;
  size_t pin_cap = 0;
  double* hpos = nullptr;
  size_t hpos_cap = 0;
  int64_t* hranges = nullptr;
  size_t hranges_cap = 0;
  int64_t* dranges = nullptr;
  size_t dranges_cap = 0;
  OutBlock* blk = nullptr;
  int64_t est_used = 0;
  std::vector<SrvReq*> reqs;
  std::thread th;
};

// Slot buffers grow to `target` (at least `bytes`, the batch's need):
// launch() asks for the need scaled towards the server's largest batch, so a
// slot reaches its working size in a few allocations, not one per new size.
static bool grow_pinned(void** p, size_t* cap, size_t bytes, size_t target,
                        std::atomic<int64_t>* n) {
  if (*cap >= bytes) return true;
  ++*n;
  if (*p) (void)hipHostFree(*p);
  *p = nullptr;
  const size_t c = std::max({bytes, target, *cap * 2});
  if (hipHostMalloc(p, c, hipHostMallocDefault) != hipSuccess) {
    *cap = 0;
    return false;
  }
  *cap = c;
  return true;
}

static bool grow_dev(void** p, size_t* cap, size_t bytes, size_t target,
                     std::atomic<int64_t>* n) {
  if (*cap >= bytes) return true;
  ++*n;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  const size_t c = std::max({bytes, target, *cap + *cap / 2});
  if (hipMalloc(p, c) != hipSuccess) {
    *cap = 0;
    return false;
  }
  *cap = c;
  return true;
}

Server::Server(Engine* eng, int slots, int64_t max_atoms, int64_t max_systems,
               int64_t edge_budget, bool full, bool dist, bool disp)
    : eng_(eng),
      max_atoms_(max_atoms),
      max_systems_(max_systems),
      budget_(edge_budget),
      full_(full),
      dist_(dist),
      disp_(disp) {
  (void)hipSetDevice(eng->device());
  slots = std::max(1, std::min(8, slots));
  for (int i = 0; i < slots; ++i) {
    Slot* s = new Slot();
    s->index = i;
    (void)hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
    (void)hipEventCreateWithFlags(&s->ev, hipEventDisableTiming);
    s->bufs.assign(kNumBuffers, nullptr);
    s->caps.assign(kNumBuffers, 0);
    slots_.push_back(s);
  }
  for (Slot* s : slots_) s->th = std::thread([this, s] { slot_loop(*s); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  for (Slot* s : slots_)
    if (s->th.joinable()) s->th.join();
  (void)hipSetDevice(eng_->device());
  for (Slot* s : slots_) (void)hipStreamSynchronize(s->st);
  for (Slot* s : slots_) {
    for (void* p : s->bufs)
      if (p) (void)hipFree(p);
    if (s->dranges) (void)hipFree(s->dranges);
    if (s->pin) (void)hipHostFree(s->pin);
    if (s->hpos) (void)hipHostFree(s->hpos);
    if (s->hranges) (void)hipHostFree(s->hranges);
    (void)hipEventDestroy(s->ev);
    (void)hipStreamDestroy(s->st);
    delete s;
  }
  std::vector<OutBlock*> blocks = pool_;
  for (auto& kv : reqs_) {
    if (kv.second->blk && --kv.second->blk->refs == 0) blocks.push_back(kv.second->blk);
    delete kv.second;
  }
  for (OutBlock* b : blocks) {
    (void)hipFree(b->dev);
    delete b;
  }
}

int64_t Server::submit(const double* pos, int64_t n, const double* h, const uint8_t* pbc,
                       double cutoff, bool borrow) {
  if (!(std::isfinite(cutoff) && cutoff > 0.0)) {
    set_error("Cutoff must be positive and finite");
    return -1;
  }
  if (n <= 0) {
    set_error("a request needs at least one atom");
    return -1;
  }
  if (h) {
    double inv[9];
    if (!inverse3(h, inv)) {
      set_error("Cell matrix is not invertible");
      return -1;
    }
  }
  SrvReq* r = new SrvReq();
  r->pos.n = n;
  double* dst = nullptr;
  if (borrow) {
    r->ext = pos;
  } else {
    const size_t bytes = sizeof(double) * 3 * (size_t)n;
    r->pos.p = (double*)nl_host_alloc(bytes);
    r->pos.pinned = r->pos.p != nullptr;
    if (!r->pos.p) r->pos.p = (double*)std::malloc(bytes);
    if (!r->pos.p) {
      delete r;
      set_error("out of host memory for the request's positions");
      return -1;
    }
    r->ext = r->pos.p;
    dst = r->pos.p;
  }
  // copy (unless borrowed), with the bounding box of an isolated cluster in
  // the same pass
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  if (h) {
    if (dst) std::memcpy(dst, pos, sizeof(double) * 3 * (size_t)n);
  } else {
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        const double v = pos[3 * i + k];
        if (dst) dst[3 * i + k] = v;
        mn[k] = std::min(mn[k], v);
        mx[k] = std::max(mx[k], v);
      }
  }
  if (h) {
    r->has_cell = 1;
    std::memcpy(r->h, h, sizeof r->h);
    for (int k = 0; k < 3; ++k) r->pbc[k] = pbc ? (pbc[k] != 0) : 1;
  }
  r->cutoff = cutoff;
  r->est = estimate_from_box(n, h, mn, mx, cutoff) * (full_ ? 2 : 1);
  std::lock_guard<std::mutex> g(mu_);
  r->id = next_id_++;
  r->t_submit = now_s();
  reqs_[r->id] = r;
  queue_.push_back(r);
  cv_.notify_all();
  return r->id;
}

SrvReq* Server::find_done(int64_t id) {
  auto it = reqs_.find(id);
  if (it == reqs_.end() || it->second->state != 2) {
    set_error("request is not complete");
    return nullptr;
  }
  return it->second;
}

int Server::wait(int64_t id, double timeout_s) {
  std::unique_lock<std::mutex> g(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) {
    set_error("unknown request id");
    return -1;
  }
  SrvReq* r = it->second;
  auto pred = [&] { return r->state >= 2; };
  if (timeout_s < 0)
    done_cv_.wait(g, pred);
  else if (!done_cv_.wait_for(g, std::chrono::duration<double>(timeout_s), pred))
    return 1;
  if (r->state == 3) {
    set_error(r->err);
    reqs_.erase(id);
    delete r;
    return -1;
  }
  return 0;
}

int Server::info(int64_t id, int64_t* n_edges, double* latency_s, int* batch_size) {
  std::lock_guard<std::mutex> g(mu_);
  SrvReq* r = find_done(id);
  if (!r) return -1;
  if (n_edges) *n_edges = r->len;
  if (latency_s) *latency_s = r->t_done - r->t_submit;
  if (batch_size) *batch_size = r->batch_size;
  return 0;
}

int Server::device_view(int64_t id, const int64_t** ei, const int64_t** ej, const int32_t** sh,
                        const double** dist, const double** disp) {
  std::lock_guard<std::mutex> g(mu_);
  SrvReq* r = find_done(id);
  if (!r) return -1;
  const OutBlock* b = r->blk;
  if (ei) *ei = b ? b->edges() + r->first : nullptr;
  if (ej) *ej = b ? b->edges() + b->total + r->first : nullptr;
  if (sh) *sh = b ? b->shifts() + 3 * r->first : nullptr;
  if (dist) *dist = b && b->dist ? b->distances() + r->first : nullptr;
  if (disp) *disp = b && b->disp ? b->displacements() + 3 * r->first : nullptr;
  return 0;
}

void Server::drop_locked(SrvReq* r) {
  reqs_.erase(r->id);
  if (r->blk && --r->blk->refs == 0) pool_.push_back(r->blk);
  delete r;
}

int Server::release(int64_t id) {
  std::lock_guard<std::mutex> g(mu_);
  SrvReq* r = find_done(id);
  if (!r) return -1;
  drop_locked(r);
  return 0;
}

int Server::take(int64_t id, int64_t* ei, int64_t* ej, int32_t* sh, double* dist,
                 double* disp) {
  const int64_t *di = nullptr, *dj = nullptr;
  const int32_t* ds = nullptr;
  const double *dd = nullptr, *dv = nullptr;
  int64_t m = 0;
  SrvReq* r;
  {
    std::lock_guard<std::mutex> g(mu_);
    r = find_done(id);
    if (!r) return -1;
    m = r->len;
    r->state = 4;  // being copied out: no second take or release meanwhile
    if (r->blk) {
      di = r->blk->edges() + r->first;
      dj = r->blk->edges() + r->blk->total + r->first;
      ds = r->blk->shifts() + 3 * r->first;
      if (r->blk->dist) dd = r->blk->distances() + r->first;
      if (r->blk->disp) dv = r->blk->displacements() + 3 * r->first;
    }
  }
  hipError_t e = hipSuccess;
  if (m > 0) {
    (void)hipSetDevice(eng_->device());
    if (ei && e == hipSuccess) e = hipMemcpy(ei, di, 8 * (size_t)m, hipMemcpyDeviceToHost);
    if (ej && e == hipSuccess) e = hipMemcpy(ej, dj, 8 * (size_t)m, hipMemcpyDeviceToHost);
    if (sh && e == hipSuccess) e = hipMemcpy(sh, ds, 12 * (size_t)m, hipMemcpyDeviceToHost);
    if (dist && dd && e == hipSuccess)
      e = hipMemcpy(dist, dd, 8 * (size_t)m, hipMemcpyDeviceToHost);
    if (disp && dv && e == hipSuccess)
      e = hipMemcpy(disp, dv, 24 * (size_t)m, hipMemcpyDeviceToHost);
  }
  std::lock_guard<std::mutex> g(mu_);
  drop_locked(r);
  if (e != hipSuccess) {
    set_error(std::string("HIP error copying a served result: ") + hipGetErrorString(e));
    return -1;
  }
  return 0;
}

void Server::stats(int64_t* completed, int64_t* batches, int64_t* max_inflight,
                   int64_t* max_batch, int64_t* refills) {
  std::lock_guard<std::mutex> g(mu_);
  if (completed) *completed = completed_;
  if (batches) *batches = batches_;
  if (max_inflight) *max_inflight = max_inflight_;
  if (max_batch) *max_batch = max_batch_;
  if (refills) *refills = refills_;
}

void Server::fail_slot(Slot& s, const std::string& msg) {
  (void)hipStreamSynchronize(s.st);  // the block may still be written to
  std::lock_guard<std::mutex> g(mu_);
  for (SrvReq* r : s.reqs) {
    r->state = 3;
    r->err = msg;
    r->t_done = now_s();
    r->pos.reset();
    r->ext = nullptr;
  }
  if (s.blk) pool_.push_back(s.blk);
  s.blk = nullptr;
  inflight_est_ -= s.est_used;
  s.est_used = 0;
  s.reqs.clear();
  s.state = 0;
  --busy_;
  done_cv_.notify_all();
  cv_.notify_all();
}

// An output block for at least `cap` edges: the smallest pooled block that
// fits, else a new allocation (when the pool already holds more than a few
// blocks, its smallest one is freed first).
// Output blocks come in capacity classes (2^20 edges, then steps of 1.5x):
// batches of similar size share blocks, so a warmed-up server allocates and
// frees none (hipFree synchronises the whole device).
static int64_t block_class(int64_t cap) {
  int64_t c = int64_t(1) << 20;
  while (c < cap) c += c / 2;
  return c;
}

OutBlock* Server::get_block(int64_t cap) {
  cap = block_class(cap);
  OutBlock* victim = nullptr;
  {
    std::lock_guard<std::mutex> g(mu_);
    size_t bi = pool_.size();
    for (size_t i = 0; i < pool_.size(); ++i)
      if (pool_[i]->cap >= cap && (bi == pool_.size() || pool_[i]->cap < pool_[bi]->cap)) bi = i;
    if (bi < pool_.size()) {
      OutBlock* b = pool_[bi];
      pool_.erase(pool_.begin() + bi);
      return b;
    }
    if (pool_.size() > 2 * slots_.size() + 2) {
      auto it = std::min_element(pool_.begin(), pool_.end(),
                                 [](OutBlock* a, OutBlock* b) { return a->cap < b->cap; });
      victim = *it;
      pool_.erase(it);
    }
  }
  if (victim) {
    (void)hipFree(victim->dev);
    delete victim;
  }
  ++allocs_;
  OutBlock* b = new OutBlock();
  b->cap = cap;
  b->dist = dist_;
  b->disp = disp_;
  if (hipMalloc((void**)&b->dev, OutBlock::bytes_per_edge(dist_, disp_) * (size_t)cap) !=
      hipSuccess) {
    delete b;
    return nullptr;
  }
  return b;
}

// Launch one batch on a slot (called by the slot's thread, queue lock not held).
bool Server::launch(Slot& s, std::vector<SrvReq*>& batch) {
  Request req;
  int64_t n = 0;
  req.starts.push_back(0);
  double rcmax = 0;
  for (SrvReq* r : batch) {
    n += r->pos.n;
    req.starts.push_back(n);
    SysIn si;
    si.has_cell = r->has_cell;
    std::memcpy(si.h, r->h, sizeof si.h);
    std::memcpy(si.pbc, r->pbc, 3);
    req.sys.push_back(si);
    req.sys_cutoff.push_back(r->cutoff);
    rcmax = std::max(rcmax, r->cutoff);
  }
  req.cutoffs = {rcmax};
  req.n = n;
  req.pos_device = true;
  req.global_index = false;
  req.full = full_;
  // Serving batches (many systems, each cutoff its own) are not autotuned:
  // bins per system by density (or the configured bin size), and sixteen
  // home atoms per group over the whole half stencil, measured 1.25x faster
  // than four on batches from the serving trace
  Tuning t;
  Config& c = config();
  t.cell_div = c.cell_div.load();
  t.use_fp32 = c.use_fp32.load();
  t.group = 16;
  t.parts = 1;
  s.reqs = batch;
  if (s.P.prepare(req, t)) {
    fail_slot(s, last_error());
    return false;
  }
  std::vector<size_t> sz = s.P.buffer_sizes();
  sz[0] = sizeof(double) * 3 * (size_t)n;
  const size_t S = batch.size();
  // every buffer grows with the batch's atoms or systems: scale a need
  // towards the largest batch the server admits (at most 16x at a time)
  const double scale = std::min(
      16.0, std::max({1.0, (double)max_atoms_ / (double)std::max<int64_t>(n, 1),
                      (double)max_systems_ / (double)std::max<size_t>(S, 1)}));
  auto target = [scale](size_t need) { return (size_t)((double)need * scale); };
  for (int i = 0; i < kNumBuffers; ++i) {
    const size_t need = std::max<size_t>(sz[i], 16);
    if (!grow_dev(&s.bufs[i], &s.caps[i], need, target(need), &allocs_)) {
      fail_slot(s, "out of device memory in the server");
      return false;
    }
  }
  s.P.assign(s.B, s.bufs);
  const size_t pin_need = s.P.pinned_bytes(), hpos_need = sizeof(double) * 3 * (size_t)n;
  if (!grow_pinned((void**)&s.pin, &s.pin_cap, pin_need, target(pin_need), &allocs_) ||
      !grow_pinned((void**)&s.hpos, &s.hpos_cap, hpos_need, target(hpos_need), &allocs_) ||
      !grow_pinned((void**)&s.hranges, &s.hranges_cap, 16 * S, target(16 * S), &allocs_) ||
      !grow_dev((void**)&s.dranges, &s.dranges_cap, 16 * S, target(16 * S), &allocs_)) {
    fail_slot(s, "out of memory in the server");
    return false;
  }
  size_t off = 0;
  for (SrvReq* r : batch) {
    std::memcpy(s.hpos + off, r->ext, 24 * (size_t)r->pos.n);
    off += 3 * (size_t)r->pos.n;
  }
  s.P.stage_host(s.pin);
  int64_t est = 0;
  for (SrvReq* r : batch) est += r->est;
  s.blk = get_block(est + est / 4 + 4096);
  if (!s.blk) {
    fail_slot(s, "out of device memory for the batch's edge lists");
    return false;
  }
  hipError_t e = hipMemcpyAsync(s.B.pos, s.hpos, sizeof(double) * 3 * (size_t)n,
                                hipMemcpyHostToDevice, s.st);
  if (e == hipSuccess) e = s.P.enqueue_upload(s.B, s.pin, s.st);
  if (e == hipSuccess) e = s.P.enqueue_build_count(s.B, s.st);
  if (e == hipSuccess)
    e = s.P.enqueue_fill(s.B, s.blk->edges(), s.blk->shifts(), s.blk->distances(), s.blk->cap,
                         s.st, s.blk->displacements());
  if (e == hipSuccess)
    e = launch_sys_ranges(s.B.offsets, s.B.items, (int)S, 1, s.P.n_items, s.dranges, s.st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s.hranges, s.dranges, 16 * S, hipMemcpyDeviceToHost, s.st);
  if (e == hipSuccess) e = s.P.enqueue_bounds(s.B, s.pin, s.st);
  if (e == hipSuccess) e = hipEventRecord(s.ev, s.st);
  if (e != hipSuccess) {
    fail_slot(s, std::string("HIP error in the server: ") + hipGetErrorString(e));
    return false;
  }
  s.state = 1;
  return true;
}

// Counts are known: deliver, or first refill into a larger block if the
// estimate was short (the first fill pass then wrote nothing).  Returns 0
// when delivered, 1 when a refill is on the stream, -1 on failure.
int Server::on_counted(Slot& s) {
  const int64_t total = s.P.host_bounds(s.pin)[1];
  if (total <= s.blk->cap) {
    deliver(s, total);
    return 0;
  }
  {
    std::lock_guard<std::mutex> g(mu_);
    pool_.push_back(s.blk);
    ++refills_;
  }
  s.blk = get_block(total + total / 8 + 4096);
  if (!s.blk) {
    fail_slot(s, "out of device memory for the batch's edge lists");
    return -1;
  }
  hipError_t e = s.P.enqueue_fill(s.B, s.blk->edges(), s.blk->shifts(), s.blk->distances(),
                                  s.blk->cap, s.st, s.blk->displacements());
  if (e == hipSuccess) e = hipEventRecord(s.ev, s.st);
  if (e != hipSuccess) {
    fail_slot(s, std::string("HIP error in the server: ") + hipGetErrorString(e));
    return -1;
  }
  s.state = 2;
  return 1;
}

void Server::deliver(Slot& s, int64_t total) {
  const double t = now_s();
  std::lock_guard<std::mutex> g(mu_);
  s.blk->total = total;
  s.blk->refs = (int)s.reqs.size();
  for (size_t k = 0; k < s.reqs.size(); ++k) {
    SrvReq* r = s.reqs[k];
    r->pos.reset();  // the batch's upload completed before its counts did
    r->ext = nullptr;
    r->blk = s.blk;
    r->first = s.hranges[2 * k];
    r->len = s.hranges[2 * k + 1] - r->first;
    r->t_done = t;
    r->state = 2;
    ++completed_;
  }
  s.blk = nullptr;
  inflight_est_ -= s.est_used;
  s.est_used = 0;
  s.reqs.clear();
  s.state = 0;
  --busy_;
  done_cv_.notify_all();
  cv_.notify_all();
}

// The head of the queue may start now (mu_ held): it fits the edge budget
// next to the batches in flight, or nothing is in flight at all.
bool Server::admissible_locked() const {
  return !queue_.empty() && (busy_ == 0 || inflight_est_ + queue_.front()->est <= budget_);
}

void Server::slot_loop(Slot& s) {
  (void)hipSetDevice(eng_->device());
  while (true) {
    std::vector<SrvReq*> batch;
    int64_t est = 0;
    {
      std::unique_lock<std::mutex> g(mu_);
      cv_.wait(g, [&] { return admissible_locked() || (stop_ && queue_.empty()); });
      if (queue_.empty()) break;  // stopping, nothing left
      int64_t atoms = 0;
      while (!queue_.empty()) {
        SrvReq* r = queue_.front();
        const int64_t na = r->pos.n;
        const bool first = batch.empty();
        if (!first && (atoms + na > max_atoms_ || (int64_t)batch.size() >= max_systems_))
          break;
        // edge-budget admission; a lone request is admitted when nothing
        // else is on the device, so oversized requests cannot starve
        if (inflight_est_ + est + r->est > budget_ && !(first && busy_ == 0)) break;
        batch.push_back(r);
        atoms += na;
        est += r->est;
        queue_.pop_front();
      }
      inflight_est_ += est;
      ++busy_;
      max_inflight_ = std::max<int64_t>(max_inflight_, busy_);
      for (SrvReq* r : batch) {
        r->state = 1;
        r->batch_size = (int)batch.size();
      }
      ++batches_;
      max_batch_ = std::max<int64_t>(max_batch_, (int64_t)batch.size());
      if (!queue_.empty()) cv_.notify_all();  // another slot may take the rest
    }
    s.est_used = est;
    if (!launch(s, batch)) continue;  // failed and accounted for
    hipError_t e = hipEventSynchronize(s.ev);
    if (e != hipSuccess) {
      fail_slot(s, std::string("HIP error in the server: ") + hipGetErrorString(e));
      continue;
    }
    if (on_counted(s) != 1) continue;  // delivered (or failed)
    e = hipEventSynchronize(s.ev);
    if (e != hipSuccess) {
      fail_slot(s, std::string("HIP error in the server: ") + hipGetErrorString(e));
      continue;
    }
    deliver(s, s.P.host_bounds(s.pin)[1]);
  }
}

}  // namespace nl
