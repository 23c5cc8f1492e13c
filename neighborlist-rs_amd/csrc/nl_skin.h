// nl_skin.h -- neighbour lists with a skin for molecular dynamics.
//
// The lists are built at r_list = rc + skin (per list) from the positions of
// the last build (the reference positions) and reused while no atom has moved
// more than half the smallest skin since then: a pair within rc now was
// within rc + skin at the last build.  Each update measures the largest
// displacement on the device (one reduction over the atoms) and replays the
// graph-captured build only when it exceeds that bound, or when the cell
// changes.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "nl_engine.h"

namespace nl {

class Skin {
 public:
  ~Skin();
  // req: device positions, systems, cells and the cutoffs rc_q; skins: one
  // per list (>= 0).  Builds once at rc_q + skin_q.
  int create(Engine* eng, const Request& req, const std::vector<double>& skins,
             hipStream_t ustream);
  // New positions (device, same layout); cells: row-major H per system in
  // the request's order, or null when unchanged.  *rebuilt: whether the
  // lists were rebuilt.
  int update(const double* d_pos, const double* cells, hipStream_t ustream, bool* rebuilt);
  // Largest displacement since the last build, and whether update() would
  // rebuild for these positions; changes nothing.
  int check(const double* d_pos, hipStream_t ustream, double* max_disp, bool* needs);

  int lists() const { return (int)rc_.size(); }
  int64_t len(int list) const { return plan_->len(list); }
  const int64_t* edges(int list) const {
    return (const int64_t*)plan_->edges() + 2 * plan_->base(list);
  }
  const int32_t* shifts(int list) const {
    return (const int32_t*)plan_->shifts() + 3 * plan_->base(list);
  }
  double radius(int list) const { return rc_[list] + skin_[list]; }
  int64_t updates = 0, rebuilds = 0;
  double last_disp = 0.0;

 private:
  int build(const double* d_pos);  // copy positions in, replay the plan
  int max_disp(const double* d_pos, hipStream_t ustream, double* out);
  Engine* eng_ = nullptr;
  Request req_;  // positions point at pos_ (the reference positions)
  std::vector<double> rc_, skin_;
  double trigger_ = 0.0;  // half the smallest skin
  std::unique_ptr<Plan> plan_;
  double* pos_ = nullptr;
  unsigned long long* dmax_ = nullptr;
  unsigned long long* hmax_ = nullptr;  // pinned
  hipStream_t st_ = nullptr;
  hipEvent_t ev_ = nullptr;
};

}  // namespace nl
