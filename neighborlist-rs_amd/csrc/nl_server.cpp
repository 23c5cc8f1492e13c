// nl_server.cpp -- serving: continuous batching of single-system requests.
//
// Requests (positions, optional cell, one cutoff each) arrive at any time.
// A worker thread keeps up to `slots` batches in flight, each on its own HIP
// stream with its own device buffers.  Whenever a slot is free it admits the
// queued requests in arrival order into one ragged batch (per-system cutoffs,
// local indices), bounded by atoms per batch and by an edge-memory budget:
// every request carries an estimate of its edge count, and the sum over the
// batches on the device may not exceed the budget (a request alone always
// fits, so nothing starves).  Each slot runs the same launch sequence as the
// plain path (plan, bin, radix sort, finalize, count, scan, fill), so a
// result is identical whether a request ran alone or inside a batch.
//
// Per slot the worker drives a small state machine on events, never
// blocking on one slot while others run:
//   launched -> (counts known) -> [refill if the estimate was short] ->
//   copying (lists to pinned host memory) -> delivered
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "nl_engine.h"
#include "nl_server.h"

namespace nl {

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Expected half-list size: half of N times the atoms within rc of an atom
// (number density times the sphere volume; images included for periodic
// cells, at most N-1 partners for an isolated cluster).
int64_t estimate_edges(const double* pos, int64_t n, const double* h, double rc) {
  double vol;
  if (h) {
    const double det = (h[0] * (h[4] * h[8] - h[7] * h[5]) - h[1] * (h[3] * h[8] - h[6] * h[5]) +
                        h[2] * (h[3] * h[7] - h[6] * h[4]));
    vol = std::fabs(det);
  } else {
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        mn[k] = std::min(mn[k], pos[3 * i + k]);
        mx[k] = std::max(mx[k], pos[3 * i + k]);
      }
    vol = 1.0;
    for (int k = 0; k < 3; ++k) vol *= (mx[k] - mn[k]) + rc;
  }
  const double rho = (double)n / std::max(vol, 1e-9);
  double per_atom = rho * 4.18879020478639 * rc * rc * rc;
  if (!h) per_atom = std::min(per_atom, (double)(n - 1));
  return (int64_t)(0.5 * (double)n * per_atom) + 16;
}

struct SrvReq {
  int64_t id = 0;
  std::vector<double> pos;
  int has_cell = 0;
  double h[9] = {0};
  uint8_t pbc[3] = {1, 1, 1};
  double cutoff = 0;
  double t_submit = 0, t_done = 0;
  int64_t est = 0;
  int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
  std::string err;
  std::vector<int64_t> ei, ej;
  std::vector<int32_t> sh;
  int batch = -1, batch_size = 0, slot = -1;
};

struct Server::Slot {
  int index = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  int state = 0;  // 0 free, 1 launched, 2 refilled, 3 copying
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
  void* edges = nullptr;
  void* shifts = nullptr;
  int64_t out_cap = 0;
  char* hout = nullptr;
  size_t hout_cap = 0;
  int64_t cap_used = 0, total = 0;
  std::vector<SrvReq*> reqs;
  int batch = -1;
};

template <typename T>
static bool grow_pinned(T** p, size_t* cap, size_t bytes) {
  if (*cap >= bytes) return true;
  if (*p) (void)hipHostFree(*p);
  *p = nullptr;
  size_t c = std::max(bytes, *cap * 2);
  if (hipHostMalloc((void**)p, c, hipHostMallocDefault) != hipSuccess) {
    *cap = 0;
    return false;
  }
  *cap = c;
  return true;
}

template <typename T>
static bool grow_dev(T** p, size_t* cap, size_t bytes) {
  if (*cap >= bytes) return true;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  size_t c = std::max(bytes, *cap + *cap / 2);
  if (hipMalloc((void**)p, c) != hipSuccess) {
    *cap = 0;
    return false;
  }
  *cap = c;
  return true;
}

Server::Server(Engine* eng, int slots, int64_t max_atoms, int64_t max_systems,
               int64_t edge_budget)
    : eng_(eng), max_atoms_(max_atoms), max_systems_(max_systems), budget_(edge_budget) {
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
  th_ = std::thread([this] { loop(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (th_.joinable()) th_.join();
  for (Slot* s : slots_) {
    (void)hipStreamSynchronize(s->st);
    for (void* p : s->bufs)
      if (p) (void)hipFree(p);
    if (s->dranges) (void)hipFree(s->dranges);
    if (s->edges) (void)hipFree(s->edges);
    if (s->shifts) (void)hipFree(s->shifts);
    if (s->pin) (void)hipHostFree(s->pin);
    if (s->hpos) (void)hipHostFree(s->hpos);
    if (s->hranges) (void)hipHostFree(s->hranges);
    if (s->hout) (void)hipHostFree(s->hout);
    (void)hipEventDestroy(s->ev);
    (void)hipStreamDestroy(s->st);
    delete s;
  }
  for (auto& kv : reqs_) delete kv.second;
}

int64_t Server::submit(const double* pos, int64_t n, const double* h, const uint8_t* pbc,
                       double cutoff) {
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
  r->pos.assign(pos, pos + 3 * n);
  if (h) {
    r->has_cell = 1;
    std::memcpy(r->h, h, sizeof r->h);
    for (int k = 0; k < 3; ++k) r->pbc[k] = pbc ? (pbc[k] != 0) : 1;
  }
  r->cutoff = cutoff;
  r->est = estimate_edges(pos, n, h, cutoff);
  std::lock_guard<std::mutex> g(mu_);
  r->id = next_id_++;
  r->t_submit = now_s();
  reqs_[r->id] = r;
  queue_.push_back(r);
  ++submitted_;
  cv_.notify_all();
  return r->id;
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
    return -1;
  }
  return 0;
}

int Server::info(int64_t id, int64_t* n_edges, double* latency_s, int* batch_size) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end() || it->second->state != 2) {
    set_error("request is not complete");
    return -1;
  }
  SrvReq* r = it->second;
  if (n_edges) *n_edges = (int64_t)r->ei.size();
  if (latency_s) *latency_s = r->t_done - r->t_submit;
  if (batch_size) *batch_size = r->batch_size;
  return 0;
}

int Server::take(int64_t id, int64_t* ei, int64_t* ej, int32_t* sh) {
  SrvReq* r;
  {
    std::lock_guard<std::mutex> g(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end() || it->second->state != 2) {
      set_error("request is not complete");
      return -1;
    }
    r = it->second;
    reqs_.erase(it);
  }
  const size_t m = r->ei.size();
  if (ei && m) std::memcpy(ei, r->ei.data(), 8 * m);
  if (ej && m) std::memcpy(ej, r->ej.data(), 8 * m);
  if (sh && m) std::memcpy(sh, r->sh.data(), 12 * m);
  delete r;
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
  std::lock_guard<std::mutex> g(mu_);
  for (SrvReq* r : s.reqs) {
    r->state = 3;
    r->err = msg;
    r->t_done = now_s();
  }
  inflight_est_ -= s.cap_used;
  s.reqs.clear();
  s.state = 0;
  done_cv_.notify_all();
}

// Launch one batch on a free slot (called by the worker, queue lock not held).
bool Server::launch(Slot& s, std::vector<SrvReq*>& batch) {
  (void)hipSetDevice(eng_->device());
  Request req;
  int64_t n = 0;
  req.starts.push_back(0);
  for (SrvReq* r : batch) {
    n += (int64_t)(r->pos.size() / 3);
    req.starts.push_back(n);
    SysIn si;
    si.has_cell = r->has_cell;
    std::memcpy(si.h, r->h, sizeof si.h);
    std::memcpy(si.pbc, r->pbc, 3);
    req.sys.push_back(si);
    req.sys_cutoff.push_back(r->cutoff);
  }
  double rcmax = 0;
  for (SrvReq* r : batch) rcmax = std::max(rcmax, r->cutoff);
  req.cutoffs = {rcmax};
  req.n = n;
  req.pos_device = true;
  req.global_index = false;
  Tuning t;
  Config& c = config();
  t.cell_div = c.cell_div.load();
  t.use_fp32 = c.use_fp32.load();
  t.group = c.group.load();
  t.parts = c.parts.load();
  s.reqs = batch;
  if (s.P.prepare(req, t)) {
    fail_slot(s, last_error());
    return false;
  }
  std::vector<size_t> sz = s.P.buffer_sizes();
  sz[0] = sizeof(double) * 3 * (size_t)n;
  for (int i = 0; i < kNumBuffers; ++i)
    if (!grow_dev(&s.bufs[i], &s.caps[i], std::max<size_t>(sz[i], 16))) {
      fail_slot(s, "out of device memory in the server");
      return false;
    }
  Pipeline::assign(s.B, s.bufs);
  const size_t S = batch.size();
  if (!grow_pinned(&s.pin, &s.pin_cap, s.P.pinned_bytes()) ||
      !grow_pinned(&s.hpos, &s.hpos_cap, sizeof(double) * 3 * (size_t)n) ||
      !grow_pinned(&s.hranges, &s.hranges_cap, 16 * S) ||
      !grow_dev(&s.dranges, &s.dranges_cap, 16 * S)) {
    fail_slot(s, "out of memory in the server");
    return false;
  }
  size_t off = 0;
  for (SrvReq* r : batch) {
    std::memcpy(s.hpos + off, r->pos.data(), 8 * r->pos.size());
    off += r->pos.size();
  }
  s.P.stage_host(s.pin);
  int64_t est = 0;
  for (SrvReq* r : batch) est += r->est;
  s.cap_used = est + est / 4 + 4096;
  if (s.out_cap < s.cap_used) {
    if (s.edges) (void)hipFree(s.edges);
    if (s.shifts) (void)hipFree(s.shifts);
    s.edges = s.shifts = nullptr;
    if (hipMalloc(&s.edges, 16 * (size_t)s.cap_used) != hipSuccess ||
        hipMalloc(&s.shifts, 12 * (size_t)s.cap_used) != hipSuccess) {
      s.out_cap = 0;
      fail_slot(s, "out of device memory for the batch's edge lists");
      return false;
    }
    s.out_cap = s.cap_used;
  }
  hipError_t e = hipMemcpyAsync(s.B.pos, s.hpos, sizeof(double) * 3 * (size_t)n,
                                hipMemcpyHostToDevice, s.st);
  if (e == hipSuccess) e = s.P.enqueue_upload(s.B, s.pin, s.st);
  if (e == hipSuccess) e = s.P.enqueue_build_count(s.B, s.st);
  if (e == hipSuccess) e = s.P.enqueue_fill(s.B, s.edges, s.shifts, nullptr, s.out_cap, s.st);
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

// Counts are known: refill if the estimate was short, else copy the lists.
bool Server::on_counted(Slot& s) {
  const int64_t total = s.P.host_bounds(s.pin)[1];
  s.total = total;
  hipError_t e = hipSuccess;
  if (total > s.out_cap) {
    (void)hipFree(s.edges);
    (void)hipFree(s.shifts);
    s.edges = s.shifts = nullptr;
    const int64_t cap = total + total / 8 + 4096;
    if (hipMalloc(&s.edges, 16 * (size_t)cap) != hipSuccess ||
        hipMalloc(&s.shifts, 12 * (size_t)cap) != hipSuccess) {
      s.out_cap = 0;
      fail_slot(s, "out of device memory for the batch's edge lists");
      return false;
    }
    s.out_cap = cap;
    e = s.P.enqueue_fill(s.B, s.edges, s.shifts, nullptr, s.out_cap, s.st);
    std::lock_guard<std::mutex> g(mu_);
    ++refills_;
  }
  if (!grow_pinned(&s.hout, &s.hout_cap, 28 * (size_t)std::max<int64_t>(total, 1))) {
    fail_slot(s, "out of pinned memory for the results");
    return false;
  }
  if (total > 0) {
    const int64_t* ed = (const int64_t*)s.edges;
    if (e == hipSuccess)
      e = hipMemcpyAsync(s.hout, ed, 16 * (size_t)total, hipMemcpyDeviceToHost, s.st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s.hout + 16 * (size_t)total, s.shifts, 12 * (size_t)total,
                         hipMemcpyDeviceToHost, s.st);
  }
  if (e == hipSuccess) e = hipEventRecord(s.ev, s.st);
  if (e != hipSuccess) {
    fail_slot(s, std::string("HIP error in the server: ") + hipGetErrorString(e));
    return false;
  }
  s.state = 3;
  return true;
}

void Server::deliver(Slot& s) {
  const int64_t total = s.total;
  const int64_t* ei = (const int64_t*)s.hout;
  const int64_t* ej = ei + total;
  const int32_t* sh = (const int32_t*)(s.hout + 16 * (size_t)total);
  const double t = now_s();
  for (size_t k = 0; k < s.reqs.size(); ++k) {
    SrvReq* r = s.reqs[k];
    const int64_t a = s.hranges[2 * k], b = s.hranges[2 * k + 1];
    r->ei.assign(ei + a, ei + b);
    r->ej.assign(ej + a, ej + b);
    r->sh.assign(sh + 3 * a, sh + 3 * b);
  }
  std::lock_guard<std::mutex> g(mu_);
  for (SrvReq* r : s.reqs) {
    r->t_done = t;
    r->state = 2;
    ++completed_;
  }
  inflight_est_ -= s.cap_used;
  s.reqs.clear();
  s.state = 0;
  done_cv_.notify_all();
}

void Server::loop() {
  (void)hipSetDevice(eng_->device());
  int batch_id = 0;
  while (true) {
    bool progress = false;
    // 1. advance slots whose device work finished
    for (Slot* s : slots_) {
      if (s->state == 0) continue;
      if (hipEventQuery(s->ev) != hipSuccess) continue;
      progress = true;
      if (s->state == 1)
        on_counted(*s);
      else if (s->state == 3)
        deliver(*s);
    }
    // 2. admit queued requests into free slots
    for (Slot* s : slots_) {
      if (s->state != 0) continue;
      std::vector<SrvReq*> batch;
      {
        std::lock_guard<std::mutex> g(mu_);
        if (queue_.empty()) break;
        int64_t atoms = 0, est = 0;
        int busy = 0;
        for (Slot* o : slots_) busy += o->state != 0;
        while (!queue_.empty()) {
          SrvReq* r = queue_.front();
          const int64_t na = (int64_t)(r->pos.size() / 3);
          const int64_t ce = r->est + r->est / 4 + 4096;
          const bool first = batch.empty();
          if (!first && (atoms + na > max_atoms_ || (int64_t)batch.size() >= max_systems_))
            break;
          // edge-memory admission: a lone request is admitted when nothing
          // else is on the device, so oversized requests cannot starve
          if (inflight_est_ + est + ce > budget_ && !(first && busy == 0 && inflight_est_ == 0))
            break;
          batch.push_back(r);
          atoms += na;
          est += ce;
          queue_.pop_front();
        }
        if (batch.empty()) break;
        inflight_est_ += est;
        for (SrvReq* r : batch) {
          r->state = 1;
          r->batch = batch_id;
          r->batch_size = (int)batch.size();
          r->slot = s->index;
        }
        ++batches_;
        max_batch_ = std::max<int64_t>(max_batch_, (int64_t)batch.size());
      }
      s->batch = batch_id++;
      if (launch(*s, batch)) progress = true;
      int busy = 0;
      for (Slot* o : slots_) busy += o->state != 0;
      std::lock_guard<std::mutex> g(mu_);
      max_inflight_ = std::max<int64_t>(max_inflight_, busy);
    }
    // 3. idle: wait for new requests, or poll the device briefly
    bool any_busy = false;
    for (Slot* s : slots_) any_busy |= s->state != 0;
    std::unique_lock<std::mutex> g(mu_);
    if (stop_ && !any_busy && queue_.empty()) break;
    if (!progress) {
      if (any_busy)
        cv_.wait_for(g, std::chrono::microseconds(20));
      else
        cv_.wait_for(g, std::chrono::milliseconds(5));
    }
  }
}

}  // namespace nl
