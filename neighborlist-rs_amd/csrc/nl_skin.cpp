// nl_skin.cpp -- neighbour lists with a skin for molecular dynamics (the
// dynamic manager of spec.md:41-48, 209-222: skin buffer, reference
// positions, rebuild when the largest displacement exceeds skin / 2).
#include "nl_skin.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "nl_kernels.h"

namespace nl {

Skin::~Skin() {
  if (eng_) (void)hipSetDevice(eng_->device());
  if (st_) (void)hipStreamSynchronize(st_);
  plan_.reset();
  if (pos_) (void)hipFree(pos_);
  if (dmax_) (void)hipFree(dmax_);
  if (hmax_) (void)hipHostFree(hmax_);
  if (ev_) (void)hipEventDestroy(ev_);
  if (st_) (void)hipStreamDestroy(st_);
}

int Skin::create(Engine* eng, const Request& req, const std::vector<double>& skins,
                 hipStream_t ustream) {
  eng_ = eng;
  (void)hipSetDevice(eng->device());
  if (!req.pos_device || req.n <= 0) {
    set_error("the skin manager needs device-resident positions of at least one atom");
    return 1;
  }
  if (skins.size() != req.cutoffs.size()) {
    set_error("one skin per cutoff is needed");
    return 1;
  }
  for (double x : skins)
    if (!(std::isfinite(x) && x >= 0.0)) {
      set_error("Skins must be non-negative and finite");
      return 1;
    }
  rc_ = req.cutoffs;
  skin_ = skins;
  trigger_ = 0.5 * *std::min_element(skins.begin(), skins.end());
  req_ = req;
  for (size_t q = 0; q < rc_.size(); ++q) req_.cutoffs[q] = rc_[q] + skin_[q];
  if (hipStreamCreateWithFlags(&st_, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ev_, hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void**)&pos_, 24 * (size_t)req.n) != hipSuccess ||
      hipMalloc((void**)&dmax_, 16) != hipSuccess ||
      hipHostMalloc((void**)&hmax_, 16, hipHostMallocDefault) != hipSuccess) {
    set_error("out of memory for the skin manager");
    return 1;
  }
  req_.pos = pos_;
  // the caller's positions are read after its queued work
  hipError_t e = hipEventRecord(ev_, ustream);
  if (e == hipSuccess) e = hipStreamWaitEvent(st_, ev_, 0);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pos_, req.pos, 24 * (size_t)req.n, hipMemcpyDeviceToDevice, st_);
  if (e == hipSuccess) e = hipStreamSynchronize(st_);
  if (e != hipSuccess) return eng_->fail(e, "skin manager set-up");
  plan_.reset(new Plan());
  if (plan_->create(eng_, req_, 0)) return 1;
  if (plan_->run()) return 1;
  ++rebuilds;
  return 0;
}

int Skin::max_disp(const double* d_pos, hipStream_t ustream, double* out) {
  hipError_t e = hipEventRecord(ev_, ustream);
  if (e == hipSuccess) e = hipStreamWaitEvent(st_, ev_, 0);
  if (e == hipSuccess) e = launch_max_disp2(d_pos, pos_, req_.n, dmax_, st_);
  if (e == hipSuccess) e = hipMemcpyAsync(hmax_, dmax_, 8, hipMemcpyDeviceToHost, st_);
  if (e == hipSuccess) e = hipStreamSynchronize(st_);
  if (e != hipSuccess) return eng_->fail(e, "displacement check");
  double d2;
  std::memcpy(&d2, hmax_, 8);
  *out = std::sqrt(d2);
  return 0;
}

int Skin::build(const double* d_pos) {
  hipError_t e = hipMemcpyAsync(pos_, d_pos, 24 * (size_t)req_.n, hipMemcpyDeviceToDevice, st_);
  if (e == hipSuccess) e = hipStreamSynchronize(st_);
  if (e != hipSuccess) return eng_->fail(e, "skin manager position copy");
  if (plan_->run()) return 1;
  ++rebuilds;
  return 0;
}

int Skin::check(const double* d_pos, hipStream_t ustream, double* disp, bool* needs) {
  (void)hipSetDevice(eng_->device());
  double m = 0.0;
  if (max_disp(d_pos, ustream, &m)) return 1;
  if (disp) *disp = m;
  if (needs) *needs = !(m <= trigger_);  // NaN rebuilds too
  return 0;
}

int Skin::update(const double* d_pos, const double* cells, hipStream_t ustream, bool* rebuilt) {
  (void)hipSetDevice(eng_->device());
  bool cell_changed = false;
  if (cells) {
    for (size_t s = 0; s < req_.sys.size(); ++s) {
      if (!req_.sys[s].has_cell) continue;
      if (std::memcmp(req_.sys[s].h, cells + 9 * s, sizeof req_.sys[s].h) != 0) {
        double inv[9];
        if (!inverse3(cells + 9 * s, inv)) {
          set_error("Cell matrix is not invertible");  // cell.rs:6
          return 1;
        }
        cell_changed = true;
      }
    }
  }
  double m = 0.0;
  if (max_disp(d_pos, ustream, &m)) return 1;
  last_disp = m;
  ++updates;
  const bool moved = !(m <= trigger_);
  if (cell_changed) {
    // a new cell: a new plan (the cells are part of the captured launch)
    for (size_t s = 0; s < req_.sys.size(); ++s)
      if (req_.sys[s].has_cell) std::memcpy(req_.sys[s].h, cells + 9 * s, sizeof req_.sys[s].h);
    hipError_t e =
        hipMemcpyAsync(pos_, d_pos, 24 * (size_t)req_.n, hipMemcpyDeviceToDevice, st_);
    if (e == hipSuccess) e = hipStreamSynchronize(st_);
    if (e != hipSuccess) return eng_->fail(e, "skin manager position copy");
    std::unique_ptr<Plan> p(new Plan());
    if (p->create(eng_, req_, plan_->capacity())) return 1;
    if (p->run()) return 1;
    plan_ = std::move(p);
    ++rebuilds;
  } else if (moved) {
    if (build(d_pos)) return 1;
  }
  if (rebuilt) *rebuilt = cell_changed || moved;
  return 0;
}

}  // namespace nl
