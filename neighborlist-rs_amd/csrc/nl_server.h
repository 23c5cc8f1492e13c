// nl_server.h -- continuous-batching request server (see nl_server.cpp).
#pragma once
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace nl {

class Engine;
struct SrvReq;
struct OutBlock;

int64_t estimate_edges(const double* pos, int64_t n, const double* h, double rc);

class Server {
 public:
  // full: full (directed) lists; dist / disp: FP64 |r_ij| / r_ij per edge
  // (the same outputs for every request of this server)
  Server(Engine* eng, int slots, int64_t max_atoms, int64_t max_systems, int64_t edge_budget,
         bool full = false, bool dist = false, bool disp = false);
  ~Server();
  // returns the request id (> 0) or -1 with the thread's error message set
  // borrow: read the caller's positions until the request completes instead
  // of copying them at submit
  int64_t submit(const double* pos, int64_t n, const double* h, const uint8_t* pbc,
                 double cutoff, bool borrow = false);
  int wait(int64_t id, double timeout_s);  // 0 done, 1 timed out, -1 failed
  int info(int64_t id, int64_t* n_edges, double* latency_s, int* batch_size);
  // device addresses of a completed request's lists (valid until release/take)
  int device_view(int64_t id, const int64_t** ei, const int64_t** ej, const int32_t** sh,
                  const double** dist = nullptr, const double** disp = nullptr);
  int release(int64_t id);
  // copies out (any pointer may be null), then releases
  int take(int64_t id, int64_t* ei, int64_t* ej, int32_t* sh, double* dist = nullptr,
           double* disp = nullptr);
  bool has_dist() const { return dist_; }
  bool has_disp() const { return disp_; }
  void stats(int64_t* completed, int64_t* batches, int64_t* max_inflight, int64_t* max_batch,
             int64_t* refills);
  // device and page-locked allocations made after construction (output
  // blocks, slot buffers that had to grow); a freed buffer synchronises the
  // device, so a warmed-up server makes none
  int64_t allocations() const { return allocs_.load(); }

 private:
  struct Slot;
  void slot_loop(Slot& s);
  bool admissible_locked() const;
  bool launch(Slot& s, std::vector<SrvReq*>& batch);
  int on_counted(Slot& s);
  void deliver(Slot& s, int64_t total);
  void fail_slot(Slot& s, const std::string& msg);
  OutBlock* get_block(int64_t cap);
  SrvReq* find_done(int64_t id);  // mu_ held
  void drop_locked(SrvReq* r);     // mu_ held

  Engine* eng_;
  int64_t max_atoms_, max_systems_, budget_;
  bool full_, dist_, disp_;
  std::vector<Slot*> slots_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<SrvReq*> queue_;
  std::unordered_map<int64_t, SrvReq*> reqs_;
  std::vector<OutBlock*> pool_;  // free output blocks
  bool stop_ = false;
  std::atomic<int64_t> allocs_{0};
  int64_t next_id_ = 1, completed_ = 0, batches_ = 0, max_inflight_ = 0, max_batch_ = 0,
          inflight_est_ = 0, refills_ = 0, busy_ = 0;
};

}  // namespace nl
