// nl_kernels.hip -- gfx950 kernels of the neighbour-list engine.
//
// Pipeline per call (one launch per stage for the whole ragged batch):
//   plan      one block per system: auto-box span, H^-1, widths, dispatch
//             (single.rs:19-89), GPU bin grid, FP32 band thresholds
//   bin       one thread per atom: wrapped fractional coordinate -> bin key
//   radix     stable LSD radix sort of (key, atom) pairs, 8-bit digits; per
//             digit: tile counts, one scan over (digit, tile), scatter
//   finalize  one thread per sorted slot: cell offsets from key boundaries,
//             sorted wrapped positions (FP64 + FP32) and wrap shifts
//   pairs     one wavefront per unit (a brute-force row or a sorted atom):
//             stencil segments staged in LDS, 64 candidates per step,
//             ballot-compacted output; count pass then fill pass
//   scan      exclusive scan of per-unit counts (decoupled look-back)
//
// Compiled with -ffp-contract=off: every FP64 expression that feeds a
// decision is the reference's, term for term.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "nl_kernels.h"

namespace nl {

namespace {

constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPre = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1;
constexpr uint64_t kFlagAgg64 = 1ull << 62;
constexpr uint64_t kFlagPre64 = 2ull << 62;
constexpr uint64_t kValMask64 = (1ull << 62) - 1;

__device__ inline int lane_id() { return __lane_id(); }

__device__ inline uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ inline int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline int64_t uniform_i64(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ inline int find_sys(const int64_t* sys_start, int n_sys, int64_t a) {
  int lo = 0, hi = n_sys - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (sys_start[mid] <= a)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Inclusive scans over the wave by DPP lane moves (row shifts 1, 2, 4, 8
// inside each row of 16 lanes, then the row broadcasts of lanes 15 and 31):
// no shuffle addresses in registers.
//
// Running maximum of unsigned values with 0 as the empty value: every DPP
// move fills missing sources with 0 (bound_ctrl, or the row masks of the
// broadcasts keep the old 0), the identity of the unsigned maximum, so each
// move can fold into its v_max_u32.
__device__ __forceinline__ uint32_t wave_incl_umax_dpp(uint32_t v) {
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true));  // row_shr:1
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true));  // row_shr:2
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true));  // row_shr:4
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true));  // row_shr:8
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return v;
}

__device__ __forceinline__ int wave_incl_scan_dpp(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

template <typename T>
__device__ inline T wave_incl_scan(T v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

template <typename T>
__device__ inline T atomic_load_agent(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline void atomic_store_agent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan over one value per thread, any block of up to 16 waves;
// s_tmp holds 16 words.  *total (if given) gets the block's sum.
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int nw = (int)(blockDim.x >> 6);
  const uint32_t incl = (uint32_t)wave_incl_scan_dpp((int)v);
  if (lane == 63) s_tmp[w] = incl;
  __syncthreads();
  uint32_t pre = 0, all = 0;
  for (int i = 0; i < nw; ++i) {
    const uint32_t t = s_tmp[i];
    pre += i < w ? t : 0u;
    all += t;
  }
  if (total) *total = all;
  __syncthreads();
  return pre + incl - v;
}

// Exclusive scan over 256 values held one per thread (block of 256).
__device__ inline uint32_t block_excl_scan_256(uint32_t v, uint32_t* s_tmp /*4*/,
                                               uint32_t* total) {
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  uint32_t incl = wave_incl_scan(v);
  if (lane == 63) s_tmp[w] = incl;
  __syncthreads();
  uint32_t pre = 0;
  for (int i = 0; i < w; ++i) pre += s_tmp[i];
  if (total) *total = s_tmp[0] + s_tmp[1] + s_tmp[2] + s_tmp[3];
  __syncthreads();
  return pre + incl - v;
}

}  // namespace

// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Distances: the fill pass stores d^2 (FP64, the reference expression);
// here |r_ij| = sqrt(d^2), for the edges actually written (none when the
// lists did not fit the output buffers).
__global__ void k_sqrt_dist(double* __restrict__ dist, const int64_t* __restrict__ total,
                            int64_t capacity) {
  const int64_t m = *total;
  if (m > capacity) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride)
    dist[i] = sqrt(dist[i]);
}

// Clears two word ranges in one launch (a kernel, not a memset node, so the
// launch sequence stays safe to capture in a graph), and copies nc words
// from src to dst (a range apart from both) in the same launch.
__global__ void k_zero2(uint32_t* __restrict__ a, int64_t na, uint32_t* __restrict__ b,
                        int64_t nb, uint32_t* __restrict__ dst,
                        const uint32_t* __restrict__ src, int64_t nc) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na + nb + nc;
       i += stride) {
    if (i < na)
      a[i] = 0u;
    else if (i < na + nb)
      b[i - na] = 0u;
    else
      dst[i - na - nb] = src[i - na - nb];
  }
}

// ---------------------------------------------------------------------------
// Planning: one block per system.  Restates single.rs:19-64 (auto-box,
// Cell::new, perpendicular widths, mic_safe, brute-force dispatch) and picks
// the GPU bin grid.  The bin grid never changes which pairs are found: the
// cell-list distance uses wrapped positions and the image shift only
// (search.rs:540-549, 695-697), so any grid with a covering stencil gives
// the reference's edge set.
//
// plan_system: the plan of system `d`, called by every thread of a block of
// up to 1024 (k_plan and the small-system kernels); `band` and `rc2s` are
// this system's [k] thresholds and squared cutoffs.  Thread 0 writes the
// descriptor; the caller synchronises the block before reading it.
__device__ __forceinline__ void plan_system(const double* __restrict__ pos, SysDesc* d,
                                            float2* __restrict__ band,
                                            const double* __restrict__ rc2s, const Params& p) {
  const int64_t a0 = d->atom_start;
  const int n = d->n_atoms;
  const double rc_max = d->rc_max;
  // bounding box of an isolated system: 16 waves, four atoms per thread in
  // flight (a system of 50,000 atoms is 13 rounds of loads, not 200); the
  // fold skips NaN and does not depend on the order
  constexpr int kWavesPlan = 16;
  __shared__ double s_mn[3][kWavesPlan], s_mx[3][kWavesPlan];
  if (!d->has_cell) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i0 = threadIdx.x; i0 < n; i0 += 4 * blockDim.x) {
      double v[4][3];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = i0 + u * (int)blockDim.x;
        const double* q = pos + 3 * (a0 + (i < n ? i : 0));
#pragma unroll
        for (int k = 0; k < 3; ++k) v[u][k] = i < n ? q[k] : NAN;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          mn[k] = v[u][k] < mn[k] ? v[u][k] : mn[k];
          mx[k] = v[u][k] > mx[k] ? v[u][k] : mx[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      for (int off = 32; off > 0; off >>= 1) {
        double a = __shfl_xor(mn[k], off, 64), b = __shfl_xor(mx[k], off, 64);
        mn[k] = a < mn[k] ? a : mn[k];
        mx[k] = b > mx[k] ? b : mx[k];
      }
    }
    if ((threadIdx.x & 63) == 0) {
      for (int k = 0; k < 3; ++k) {
        s_mn[k][threadIdx.x >> 6] = mn[k];
        s_mx[k][threadIdx.x >> 6] = mx[k];
      }
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double h[9];
  double occupied = 1.0;  // volume the atoms occupy (density for the bin size)
  if (!d->has_cell) {
    const double margin = rc_max + 1.0;  // AUTO_BOX_MARGIN (single.rs:6)
    const int waves = (int)(blockDim.x >> 6);
    for (int k = 0; k < 9; ++k) h[k] = 0.0;
    for (int k = 0; k < 3; ++k) {
      double mn = s_mn[k][0], mx = s_mx[k][0];
      for (int w = 1; w < waves; ++w) {
        mn = s_mn[k][w] < mn ? s_mn[k][w] : mn;
        mx = s_mx[k][w] > mx ? s_mx[k][w] : mx;
      }
      h[4 * k] = (mx - mn) + 2.0 * margin;
      // the bounding box, each side at least half a cutoff (flat clusters)
      occupied *= fmax(mx - mn, 0.5 * rc_max);
    }
    for (int k = 0; k < 3; ++k) d->pbc[k] = 0;
  } else {
    for (int k = 0; k < 9; ++k) h[k] = d->h[k];
    occupied = fabs(det3(h));
  }
  double hinv[9];
  if (!inverse3(h, hinv)) {
    d->status = 1;
    d->mode = kBruteIsolated;
    for (int k = 0; k < 3; ++k) {
      d->nb[k] = 1;
      d->ns[k] = 0;
    }
    return;
  }
  d->status = 0;
  double w[3];
  widths3(hinv, w);
  for (int k = 0; k < 9; ++k) {
    d->h[k] = h[k];
    d->hinv[k] = hinv[k];
    d->h32[k] = (float)h[k];
  }
  for (int k = 0; k < 3; ++k) d->w[k] = w[k];
  const double minw = fmin(fmin(w[0], w[1]), w[2]);
  const bool mic_safe = rc_max * 2.0 < minw;
  const uint64_t thr = ((uint64_t)(uint32_t)p.bf_threshold_hi << 32) | (uint32_t)p.bf_threshold_lo;
  const bool any_pbc = d->pbc[0] || d->pbc[1] || d->pbc[2];
  int mode;
  if ((uint64_t)n < thr && mic_safe)
    mode = (p.multi || any_pbc) ? kBruteMic : kBruteIsolated;
  else
    mode = kCellList;
  d->mode = mode;
  if (mode != kCellList) {
    for (int k = 0; k < 3; ++k) {
      d->nb[k] = 1;
      d->ns[k] = 0;
    }
    d->delta = 0.0;
    return;
  }
  // GPU bin grid: edge ~ rc_max / cell_div, at most bin_cap bins.
  const int cell_div =
      p.cell_div > 0 ? p.cell_div : bins_per_cutoff(rc_max, (double)n, occupied);
  const double edge = rc_max / (double)cell_div;
  int nb[3];
  for (int k = 0; k < 3; ++k) {
    double v = floor(w[k] / edge);
    nb[k] = v < 1.0 ? 1 : (v > 4096.0 ? 4096 : (int)v);
  }
  const int64_t cap = d->bin_cap;
  int64_t prod = (int64_t)nb[0] * nb[1] * nb[2];
  if (prod > cap) {
    double f = cbrt((double)cap / (double)prod);
    for (int k = 0; k < 3; ++k) {
      int v = (int)(nb[k] * f);
      nb[k] = v < 1 ? 1 : v;
    }
    prod = (int64_t)nb[0] * nb[1] * nb[2];
    while (prod > cap) {
      int kk = 0;
      for (int k = 1; k < 3; ++k)
        if (nb[k] > nb[kk]) kk = k;
      nb[kk] -= 1;
      prod = (int64_t)nb[0] * nb[1] * nb[2];
    }
  }
  int ns[3];
  for (int k = 0; k < 3; ++k) {
    // covering reach: ceil(rc n / w), widened by one bin when the ratio sits
    // within 1e-9 of an integer (floating-point slack at the boundary)
    double r = rc_max * (double)nb[k] / w[k];
    ns[k] = (int)ceil(r + 1e-9);
    d->nb[k] = nb[k];
    d->ns[k] = ns[k];
  }
  // FP32 filter band: bound on |d32 - d64| for any candidate (see the
  // pair kernel).  P bounds |wrapped coordinate|, O bounds |image offset|.
  double P = 0.0, O = 0.0;
  for (int r = 0; r < 3; ++r) {
    double pr = 0.0, orr = 0.0;
    for (int c = 0; c < 3; ++c) {
      double a = fabs(h[3 * r + c]);
      pr += a;
      orr += a * ((double)ns[c] / (double)nb[c] + 2.0);
    }
    P = fmax(P, pr);
    O = fmax(O, orr);
  }
  const double delta = 64.0 * 5.9604644775390625e-08 * (P + O + rc_max);
  d->delta = delta;
  for (int q = 0; q < p.k; ++q) {
    double rc = sqrt(rc2s[q]);
    double lo = rc - delta;
    lo = lo > 0.0 ? lo * lo : 0.0;
    double hi = (rc + delta) * (rc + delta);
    // round the thresholds outward so the certain-accept / certain-reject
    // decisions stay conservative
    band[q] = make_float2(__double2float_rd(lo), __double2float_ru(hi));
  }
}

__global__ __launch_bounds__(1024) void k_plan(const double* __restrict__ pos,
                                              SysDesc* __restrict__ sys,
                                              float2* __restrict__ band,
                                              const double* __restrict__ rc2, Params p) {
  const int s = blockIdx.x;
  plan_system(pos, sys + s, band + (int64_t)s * p.k, p.per_sys_rc ? rc2 + (int64_t)s * p.k : rc2,
              p);
}

// ---------------------------------------------------------------------------
// Morton order of the bins (search.rs:126-146): one thread per reserved bin;
// the bins of the grid k_plan chose get their rank and the inverse entry.
__global__ __launch_bounds__(256) void k_morton(const SysDesc* __restrict__ sys, int n_sys,
                                                int64_t total_bins, int32_t* __restrict__ bin_rank,
                                                int32_t* __restrict__ rank_bin) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total_bins) return;
  int lo = 0, hi = n_sys - 1;  // the system whose reservation holds bin t
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sys[mid].bin_base <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  const SysDesc* d = sys + lo;
  if (d->mode != kCellList) return;
  const int nx = d->nb[0], ny = d->nb[1], nz = d->nb[2];
  const int64_t l = t - d->bin_base;
  if (l >= (int64_t)nx * ny * nz) return;
  const int x = (int)(l % nx), y = (int)((l / nx) % ny), z = (int)(l / ((int64_t)nx * ny));
  const int32_t r = morton_rank(x, y, z, nx, ny, nz);
  bin_rank[t] = r;
  rank_bin[d->bin_base + r] = (int32_t)l;
}

// Binning: one thread per atom.  Key = bin_base + bin (cell-list systems;
// the linear bin, or its Morton rank when bin_rank is given) or bin_base
// (brute-force systems keep their input order).
__global__ __launch_bounds__(256) void k_bin(const double* __restrict__ pos,
                                             const SysDesc* __restrict__ sys,
                                             const int64_t* __restrict__ sys_start, int n_sys,
                                             int64_t n, const int32_t* __restrict__ bin_rank,
                                             uint32_t* __restrict__ keys,
                                             uint32_t* __restrict__ vals) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int s = find_sys(sys_start, n_sys, a);
  const SysDesc* d = sys + s;
  uint32_t key = (uint32_t)d->bin_base;
  if (d->mode == kCellList) {
    const double* q = pos + 3 * a;
    double f[3];
    matvec3(d->hinv, q[0], q[1], q[2], f);
    int b[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double u = f[k] - floor(f[k]);
      int v = (int)(u * (double)d->nb[k]);  // u >= 0
      b[k] = v < d->nb[k] - 1 ? v : d->nb[k] - 1;
    }
    const int lin = b[0] + d->nb[0] * (b[1] + d->nb[1] * b[2]);
    key += (uint32_t)(bin_rank ? bin_rank[d->bin_base + lin] : lin);
  }
  keys[a] = key;
  vals[a] = (uint32_t)a;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, reduce-then-scan, digits of `bits` <= 12 bits:
// per tile (2048 keys) digit counts, one exclusive scan over (digit, tile)
// in digit-major order (the scan kernel below), then a stable scatter.  No
// block waits on another, so a pass costs three short launches instead of a
// chain of look-backs across tiles.
//
// Digit counts of one tile: cnt[digit * tiles + tile].
__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* __restrict__ kin, int64_t n,
                                                    int shift, int bits, int64_t tiles,
                                                    uint32_t* __restrict__ cnt) {
  __shared__ uint32_t h[kRadixMaxDigits];
  const int tid = threadIdx.x;
  const int nd = 1 << bits;
  const uint32_t mask = (uint32_t)nd - 1u;
  for (int d = tid; d < nd; d += 256) h[d] = 0u;
  __syncthreads();
  const int64_t t0 = (int64_t)blockIdx.x * kRadixTile;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int64_t e = t0 + r * 256 + tid;
    if (e < n) atomicAdd(&h[(kin[e] >> shift) & mask], 1u);
  }
  __syncthreads();
  for (int d = tid; d < nd; d += 256) cnt[(int64_t)d * tiles + blockIdx.x] = h[d];
}

// Ranks the keys of one tile (2048 keys, 256 threads) by digit, stably.
// Wave w owns the tile's keys [512 w, 512 w + 512) in 8 rounds of 64; it
// ranks them with ballots and wave-private digit counters in LDS (the lanes
// of one wave read a counter before its group leader advances it, in
// instruction order, so no barrier is needed).  Then one pass over the
// four waves' counters turns wave ranks into tile ranks.  Element order is
// wave-major, rounds in order, lanes in order, so equal digits keep their
// input order.  On return lp[r] = s_run[digit] on entry + rank within the
// tile, and s_run[d] has advanced by the tile's count of digit d.
// Element of thread (w, lane) in round r: t0 + 512 w + 64 r + lane.
__device__ __forceinline__ void radix_rank_tile(const uint32_t* __restrict__ kin,
                                                const uint32_t* __restrict__ vin, int64_t n,
                                                int shift, int bits, int64_t t0, uint32_t* s_run,
                                                uint32_t (*s_cnt)[kRadixMaxDigits], uint32_t* kk,
                                                uint32_t* vv, uint32_t* lp) {
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int nd = 1 << bits;
  const uint32_t mask = (uint32_t)nd - 1u;
  uint32_t* wc = s_cnt[w];  // this wave's counters, zero on entry
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int64_t e = t0 + 512 * w + 64 * r + lane;
    const bool valid = e < n;
    const uint32_t key = valid ? kin[e] : 0u;
    const uint32_t dig = (key >> shift) & mask;
    kk[r] = key;
    vv[r] = valid ? vin[e] : 0u;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
      const bool bit = (dig >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = lanes_below(peers);
    const uint32_t before = valid ? wc[dig] : 0u;
    __builtin_amdgcn_wave_barrier();
    if (valid && rank == 0) wc[dig] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    lp[r] = before + rank;
  }
  __syncthreads();
  // wave prefix per digit
  for (int d = tid; d < nd; d += 256) {
    const uint32_t c0 = s_cnt[0][d], c1 = s_cnt[1][d], c2 = s_cnt[2][d], c3 = s_cnt[3][d];
    const uint32_t run = s_run[d];
    s_cnt[0][d] = run;
    s_cnt[1][d] = run + c0;
    s_cnt[2][d] = run + c0 + c1;
    s_cnt[3][d] = run + c0 + c1 + c2;
    s_run[d] = run + c0 + c1 + c2 + c3;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) lp[r] += wc[(kk[r] >> shift) & mask];
  __syncthreads();
  for (int d = tid; d < nd; d += 256)
    s_cnt[0][d] = s_cnt[1][d] = s_cnt[2][d] = s_cnt[3][d] = 0u;
  __syncthreads();
}

// Scatter of one tile: off[digit * tiles + tile] is where the tile's keys
// of that digit start.
__global__ __launch_bounds__(256) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                       const uint32_t* __restrict__ vin,
                                                       uint32_t* __restrict__ kout,
                                                       uint32_t* __restrict__ vout, int64_t n,
                                                       int shift, int bits, int64_t tiles,
                                                       const int64_t* __restrict__ off) {
  __shared__ uint32_t s_run[kRadixMaxDigits];
  __shared__ uint32_t s_cnt[4][kRadixMaxDigits];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int nd = 1 << bits;
  const uint32_t mask = (uint32_t)nd - 1u;
  for (int d = tid; d < nd; d += 256) s_run[d] = s_cnt[0][d] = s_cnt[1][d] = s_cnt[2][d] =
      s_cnt[3][d] = 0u;
  __syncthreads();
  const int64_t tile = blockIdx.x;
  const int64_t t0 = tile * kRadixTile;
  uint32_t kk[8], vv[8], lp[8];
  radix_rank_tile(kin, vin, n, shift, bits, t0, s_run, s_cnt, kk, vv, lp);
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int64_t e = t0 + 512 * w + 64 * r + lane;
    if (e < n) {
      const uint32_t dig = (kk[r] >> shift) & mask;
      const int64_t dst = off[(int64_t)dig * tiles + tile] + lp[r];
      kout[dst] = kk[r];
      vout[dst] = vv[r];
    }
  }
}

// A whole pass in one block for small inputs (one or two tiles): digit
// totals, their exclusive scan, then the tiles ranked in order with running
// digit positions.  One launch instead of three.
__global__ __launch_bounds__(256) void k_radix_small(const uint32_t* __restrict__ kin,
                                                     const uint32_t* __restrict__ vin,
                                                     uint32_t* __restrict__ kout,
                                                     uint32_t* __restrict__ vout, int64_t n,
                                                     int shift, int bits) {
  __shared__ uint32_t s_run[kRadixMaxDigits];
  __shared__ uint32_t s_cnt[4][kRadixMaxDigits];
  __shared__ uint32_t s_tmp[4];
  const int tid = threadIdx.x;
  const int nd = 1 << bits;
  const uint32_t mask = (uint32_t)nd - 1u;
  for (int d = tid; d < nd; d += 256) s_run[d] = s_cnt[0][d] = s_cnt[1][d] = s_cnt[2][d] =
      s_cnt[3][d] = 0u;
  __syncthreads();
  for (int64_t e = tid; e < n; e += 256) atomicAdd(&s_run[(kin[e] >> shift) & mask], 1u);
  __syncthreads();
  // exclusive scan of the digit totals: each thread owns a run of digits
  const int per = nd / 256 > 0 ? nd / 256 : 1;
  const int d0 = tid * per;
  uint32_t local = 0;
  for (int i = 0; i < per; ++i)
    if (d0 + i < nd) local += s_run[d0 + i];
  uint32_t acc = block_excl_scan_256(local, s_tmp, nullptr);
  for (int i = 0; i < per; ++i) {
    if (d0 + i < nd) {
      const uint32_t c = s_run[d0 + i];
      s_run[d0 + i] = acc;
      acc += c;
    }
  }
  __syncthreads();
  const int w = tid >> 6, lane = tid & 63;
  for (int64_t t0 = 0; t0 < n; t0 += kRadixTile) {
    uint32_t kk[8], vv[8], lp[8];
    radix_rank_tile(kin, vin, n, shift, bits, t0, s_run, s_cnt, kk, vv, lp);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int64_t e = t0 + 512 * w + 64 * r + lane;
      if (e < n) {
        kout[lp[r]] = kk[r];
        vout[lp[r]] = vv[r];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Segmented sort for batches of small systems: one block per system (at most
// kSegMax atoms), local bin keys (key - bin_base) sorted in LDS by stable LSD
// passes of kSegBits-bit digits, the result written back with the global
// keys.  Systems are contiguous and their bin ranges ascend, so the output is
// the globally sorted order the global radix sort would give, in one launch
// instead of three per pass.  The block also writes the bin offsets of its
// system (cell_start[bin_base + b] for b <= bin_cap) from the sorted keys in
// LDS, which k_finalize then skips.
__global__ __launch_bounds__(256) void k_sort_systems(const uint32_t* __restrict__ kin,
                                                      const uint32_t* __restrict__ vin,
                                                      uint32_t* __restrict__ kout,
                                                      uint32_t* __restrict__ vout,
                                                      const SysDesc* __restrict__ sys,
                                                      const int64_t* __restrict__ sys_start,
                                                      int passes,
                                                      int32_t* __restrict__ cell_start) {
  constexpr int ND = 1 << kSegBits;
  __shared__ uint32_t s_k[2][kSegMax];
  __shared__ uint32_t s_v[2][kSegMax];
  __shared__ uint32_t s_cnt[4][ND];  // per-wave digit counters of the current tile
  __shared__ uint32_t s_run[ND];     // digit offsets so far
  __shared__ uint32_t s_tmp[4];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int s = blockIdx.x;
  const int64_t a0 = sys_start[s];
  const int n = (int)(sys_start[s + 1] - a0);
  const uint32_t base = (uint32_t)sys[s].bin_base;
  for (int i = tid; i < n; i += 256) {
    s_k[0][i] = kin[a0 + i] - base;
    s_v[0][i] = vin[a0 + i];
  }
  int cur = 0;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * kSegBits;
    for (int d = tid; d < ND; d += 256) {
      s_run[d] = 0u;
      s_cnt[0][d] = s_cnt[1][d] = s_cnt[2][d] = s_cnt[3][d] = 0u;
    }
    __syncthreads();
    for (int i = tid; i < n; i += 256) atomicAdd(&s_run[(s_k[cur][i] >> shift) & (ND - 1)], 1u);
    __syncthreads();
    {
      const uint32_t c = tid < ND ? s_run[tid] : 0u;
      const uint32_t x = block_excl_scan_256(c, s_tmp, nullptr);
      __syncthreads();
      if (tid < ND) s_run[tid] = x;
    }
    __syncthreads();
    for (int t0 = 0; t0 < n; t0 += 2048) {
      uint32_t kk[8], vv[8], lp[8];
      // stable ranks: wave w takes elements t0 + 512 w + 64 r + lane in order
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int e = t0 + 512 * w + 64 * r + lane;
        const bool valid = e < n;
        const uint32_t key = valid ? s_k[cur][e] : 0u;
        const uint32_t dig = (key >> shift) & (ND - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kSegBits; ++b) {
          const bool bit = (dig >> b) & 1u;
          const uint64_t bb = __ballot(bit);
          peers &= bit ? bb : ~bb;
        }
        const uint32_t rank = lanes_below(peers);
        const uint32_t before = valid ? s_cnt[w][dig] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) s_cnt[w][dig] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        lp[r] = before + rank;
        kk[r] = key;
        vv[r] = valid ? s_v[cur][e] : 0u;
      }
      __syncthreads();
      for (int d = tid; d < ND; d += 256) {
        const uint32_t c0 = s_cnt[0][d], c1 = s_cnt[1][d], c2 = s_cnt[2][d], c3 = s_cnt[3][d];
        const uint32_t run = s_run[d];
        s_cnt[0][d] = run;
        s_cnt[1][d] = run + c0;
        s_cnt[2][d] = run + c0 + c1;
        s_cnt[3][d] = run + c0 + c1 + c2;
        s_run[d] = run + c0 + c1 + c2 + c3;
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int e = t0 + 512 * w + 64 * r + lane;
        if (e < n) {
          const uint32_t dst = s_cnt[w][(kk[r] >> shift) & (ND - 1)] + lp[r];
          s_k[cur ^ 1][dst] = kk[r];
          s_v[cur ^ 1][dst] = vv[r];
        }
      }
      __syncthreads();
      for (int d = tid; d < ND; d += 256) s_cnt[0][d] = s_cnt[1][d] = s_cnt[2][d] = s_cnt[3][d] = 0u;
      __syncthreads();
    }
    cur ^= 1;
  }
  for (int i = tid; i < n; i += 256) {
    kout[a0 + i] = s_k[cur][i] + base;
    vout[a0 + i] = s_v[cur][i];
  }
  // bin offsets, cell_start[base + b] = first slot whose local key is >= b:
  // slot i starts the bins (key[i-1], key[i]]; the bins up to the first key
  // start at slot 0 and those past the last key at n
  // (an empty system: every bin starts at slot 0 = n)
  __syncthreads();  // with no radix pass the loaded keys are still per-thread
  const int cap = sys[s].bin_cap;
  const int kfirst = n > 0 ? (int)s_k[cur][0] : cap;
  const int klast = n > 0 ? (int)s_k[cur][n - 1] : cap;
  for (int b = tid; b <= cap; b += 256) {
    if (b <= kfirst)
      cell_start[(int64_t)base + b] = (int32_t)a0;
    else if (b > klast)
      cell_start[(int64_t)base + b] = (int32_t)(a0 + n);
  }
  for (int i = tid + 1; i < n; i += 256) {
    const int k = (int)s_k[cur][i], prev = (int)s_k[cur][i - 1];
    for (int b = prev + 1; b <= k; ++b) cell_start[(int64_t)base + b] = (int32_t)(a0 + i);
  }
}

// ---------------------------------------------------------------------------
// The whole build of a batch whose systems all have at most kFusedMax atoms,
// one block per system, in one launch: the plan (plan_system), the bin of
// every atom (as k_bin), the sort by bin, the bin offsets, the sorted wrapped
// positions and wrap shifts (as k_finalize), the system's work items (as
// k_items) and zero counts for its item slots without work.  It replaces
// k_plan, k_bin, the sort (k_sort_systems or the radix passes), k_finalize
// and k_items, and the clearing of the counts, with the same outputs: the
// sort is a counting sort in LDS (bin histogram, exclusive scan, scatter)
// whose bins are then put back in atom order, which is the order the stable
// sort gives.  Bins come from the grid the plan chose (at most bin_cap <=
// 2 kFusedMax + 64 of them).  Local atom indices and bins fit 16 bits, so the
// per-atom arrays are u16 and the bin counts are u16 halves of u32 words
// (a count never exceeds the system's atoms, so an increment never carries
// into the other half): 123 KB of LDS at 12288 atoms.  On the 256-system
// ragged batch (rc 14, bins of rc/3) it takes 28 us where the five launches
// took 54 us.
constexpr int kFusedBins = 2 * kFusedMax + 64;

constexpr int kBuildThreads = 1024;

// bin b's count / offset in the packed halves
__device__ __forceinline__ uint32_t bin_get(const uint32_t* w, int b) {
  return (w[b >> 1] >> ((b & 1) << 4)) & 0xFFFFu;
}
// increments bin b, returns its value before
__device__ __forceinline__ uint32_t bin_inc(uint32_t* w, int b) {
  const int sh = (b & 1) << 4;
  return (atomicAdd(&w[b >> 1], 1u << sh) >> sh) & 0xFFFFu;
}

__global__ __launch_bounds__(kBuildThreads) void k_build_systems(BuildArgs a) {
  __shared__ SysDesc s_d;
  __shared__ uint16_t s_key[kFusedMax];        // bin of local atom i
  __shared__ uint16_t s_any[kFusedMax];        // slots of a bin in scatter order
  __shared__ uint16_t s_val[kFusedMax];        // local atom of sorted slot p
  __shared__ uint32_t s_bin[kFusedBins / 2];   // counts, then starts, then ends (u16 pairs)
  __shared__ uint32_t s_tmp[16];
  __shared__ uint32_t s_base[2];
  constexpr int T = kBuildThreads;
  const int tid = threadIdx.x;
  const int s = blockIdx.x;
  const int k = a.prm.k;
  const double* rc2s = a.prm.per_sys_rc ? a.rc2 + (int64_t)s * k : a.rc2;
  // words the later launches need cleared (when no clearing launch ran)
  for (int64_t i = (int64_t)s * T + tid; i < a.n_zero; i += (int64_t)gridDim.x * T) a.zero[i] = 0u;
  if (tid == 0) s_d = a.sys_in ? a.sys_in[s] : a.sys[s];
  __syncthreads();
  plan_system(a.pos, &s_d, a.band + (int64_t)s * k, rc2s, a.prm);
  __syncthreads();
  if (tid == 0) a.sys[s] = s_d;  // the plan, for the pair kernels
  const int64_t a0 = s_d.atom_start;
  const int n = s_d.n_atoms;
  const uint32_t bb = (uint32_t)s_d.bin_base;
  const int cap = s_d.bin_cap;
  const bool cells = s_d.mode == kCellList;
  if (!cells) {
    // brute force: every atom in bin 0 in input order (as the stable sort
    // of equal keys leaves it)
    for (int i = tid; i < n; i += T) {
      a.keys[a0 + i] = bb;
      a.vals[a0 + i] = (uint32_t)(a0 + i);
    }
    for (int b = tid; b <= cap; b += T) a.cell_start[(int64_t)bb + b] = (int32_t)(b == 0 ? a0 : a0 + n);
  } else {
    const int nb0 = s_d.nb[0], nb1 = s_d.nb[1], nb2 = s_d.nb[2];
    const int nbins = nb0 * nb1 * nb2;  // <= bin_cap
    for (int w = tid; w < (nbins + 1) / 2; w += T) s_bin[w] = 0u;
    __syncthreads();
    for (int i = tid; i < n; i += T) {
      const double* q = a.pos + 3 * (a0 + i);
      double f[3];
      matvec3(s_d.hinv, q[0], q[1], q[2], f);
      int c[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double u = f[d] - floor(f[d]);
        const int v = (int)(u * (double)s_d.nb[d]);  // u >= 0
        c[d] = v < s_d.nb[d] - 1 ? v : s_d.nb[d] - 1;
      }
      const int lin = c[0] + nb0 * (c[1] + nb1 * c[2]);
      s_key[i] = (uint16_t)lin;
      (void)bin_inc(s_bin, lin);
    }
    __syncthreads();
    // exclusive scan of the counts: thread t takes a contiguous run of bins
    // that starts at an even bin (whole words: no two threads share one)
    const int per = (((nbins + T - 1) / T) + 1) & ~1;
    const int b0 = min(nbins, tid * per), b1 = min(nbins, b0 + per);
    uint32_t sum = 0;
    for (int b = b0; b < b1; ++b) sum += bin_get(s_bin, b);
    uint32_t run = block_excl_scan(sum, s_tmp, nullptr);
    for (int b = b0; b < b1; b += 2) {
      const uint32_t wd = s_bin[b >> 1];
      const uint32_t c0 = wd & 0xFFFFu, c1 = wd >> 16;
      const uint32_t st0 = run, st1 = run + c0;
      a.cell_start[(int64_t)bb + b] = (int32_t)(a0 + st0);
      if (b + 1 < b1) a.cell_start[(int64_t)bb + b + 1] = (int32_t)(a0 + st1);
      s_bin[b >> 1] = st0 | (st1 << 16);
      run = st1 + (b + 1 < b1 ? c1 : 0u);
    }
    for (int b = nbins + tid; b <= cap; b += T) a.cell_start[(int64_t)bb + b] = (int32_t)(a0 + n);
    __syncthreads();
    // scatter (any order within a bin), then each atom to its place in atom
    // order within its bin: its rank among the bin's atoms, one thread per
    // slot (an insertion sort in one thread per bin cost 95 us per launch on
    // the ragged batch, whose small periodic systems keep all their atoms in
    // one or a few bins); afterwards bin b holds its end
    for (int i = tid; i < n; i += T) s_any[bin_inc(s_bin, s_key[i])] = (uint16_t)i;
    __syncthreads();
    for (int p = tid; p < n; p += T) {
      const uint32_t v = s_any[p], key = s_key[v];
      const int lo = key > 0 ? (int)bin_get(s_bin, (int)key - 1) : 0;
      const int hi = (int)bin_get(s_bin, (int)key);
      int r = 0;
      for (int x = lo; x < hi; ++x) r += s_any[x] < v ? 1 : 0;
      s_val[lo + r] = (uint16_t)v;
    }
    __syncthreads();
    // sorted keys and atoms; wrapped positions and wrap shifts of each
    // slot, as CellList::build computes them (search.rs:189-204)
    for (int p = tid; p < n; p += T) {
      const int i = (int)s_val[p];
      const int64_t at = a0 + i, slot = a0 + p;
      a.keys[slot] = bb + (uint32_t)s_key[i];
      a.vals[slot] = (uint32_t)at;
      const double* q = a.pos + 3 * at;
      double f[3];
      matvec3(s_d.hinv, q[0], q[1], q[2], f);
      int32_t si[3];
      for (int d = 0; d < 3; ++d) si[d] = as_i32(-floor(f[d]));
      double sv[3];
      matvec3(s_d.h, (double)(-si[0]), (double)(-si[1]), (double)(-si[2]), sv);
      const double wx = q[0] - sv[0], wy = q[1] - sv[1], wz = q[2] - sv[2];
      a.spos64[slot] = Pos64{wx, wy, wz, __hiloint2double(0, i)};  // w: index bits
      a.spos32[slot] = make_float4((float)wx, (float)wy, (float)wz, __int_as_float(i));
      a.sidx[slot] = make_int4(si[0], si[1], si[2], i);
    }
  }
  // work items of the system (items [item0, item0 + n (1 + parts))): its
  // rows on the brute-force path; on the cell-list path the stencil slices
  // of every slot that starts a home group (up to `group` consecutive slots
  // of one bin, aligned to the bin start).  Group items go to the front of
  // the list, rows to the back; the other item slots get zero counts.
  const int parts = a.parts, group = a.group;
  const int64_t item0 = a.item_base[s];
  auto starts_group = [&](int p) {
    const int key = (int)s_key[s_val[p]];
    const int lo = key > 0 ? (int)bin_get(s_bin, key - 1) : 0;
    return (p - lo) % group == 0;
  };
  // thread t takes a contiguous run of slots, so that the list holds the
  // system's items in slot order (neighbouring home groups read the same
  // stencil bins and run close together); systems follow each other in the
  // order their blocks finish
  __syncthreads();
  const int run_len = (n + T - 1) / T;
  const int p_lo = min(n, tid * run_len), p_hi = min(n, p_lo + run_len);
  uint32_t my_g = 0, my_r = 0;
  if (cells) {
    for (int p = p_lo; p < p_hi; ++p) my_g += starts_group(p) ? (uint32_t)parts : 0u;
  } else {
    my_r = (uint32_t)(p_hi - p_lo);
  }
  uint32_t tot_g = 0, tot_r = 0;
  const uint32_t off_g = block_excl_scan(my_g, s_tmp, &tot_g);
  const uint32_t off_r = block_excl_scan(my_r, s_tmp, &tot_r);
  if (tid == 0) {
    s_base[0] = tot_g ? atomicAdd(a.n_real, tot_g) : 0u;
    s_base[1] = tot_r ? atomicAdd(a.n_real + 1, tot_r) : 0u;
  }
  __syncthreads();
  uint32_t wg = s_base[0] + off_g, wr = s_base[1] + off_r;
  const int64_t n_items = a.n_items;
  for (int p = p_lo; p < p_hi; ++p) {
    const int64_t row = item0 + p;
    const int64_t cell0 = item0 + n + (int64_t)p * parts;
    bool grp = false;
    if (cells) {
      grp = starts_group(p);
    } else {
      a.list[n_items - 1 - wr++] = row;
    }
    if (grp) {
      for (int t = 0; t < parts; ++t) a.list[wg++] = cell0 + t;
    }
    // zero counts for the slots without work
    for (int q = 0; q < k; ++q) {
      uint32_t* cq = a.counts + (int64_t)q * n_items;
      if (cells) cq[row] = 0u;
      if (!grp)
        for (int t = 0; t < parts; ++t) cq[cell0 + t] = 0u;
    }
  }
}

// ---------------------------------------------------------------------------
// Finalize: thread t writes the offset of bin t (lower bound of t in the
// sorted keys, for t <= total_bins) and, for t < n, the sorted slot t: the
// wrapped position and wrap shift of its atom, as CellList::build computes
// them (search.rs:189-204): s = -floor(H^-1 r), wrapped = r - H (-s).
__global__ __launch_bounds__(256) void k_finalize(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
    const double* __restrict__ pos, const SysDesc* __restrict__ sys,
    const int64_t* __restrict__ sys_start, int n_sys, int64_t n, int64_t total_bins,
    int32_t* __restrict__ cell_start, Pos64* __restrict__ spos64, float4* __restrict__ spos32,
    int4* __restrict__ sidx, int with_cells) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (with_cells && t <= total_bins) {
    int64_t lo = 0, hi = n;  // first slot with key >= t
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)keys[mid] < t)
        lo = mid + 1;
      else
        hi = mid;
    }
    cell_start[t] = (int32_t)lo;
  }
  if (t >= n) return;
  const int64_t p = t;
  const int64_t a = vals[p];
  const int s = find_sys(sys_start, n_sys, a);
  const SysDesc* d = sys + s;
  if (d->mode != kCellList) return;
  const double* q = pos + 3 * a;
  double f[3];
  matvec3(d->hinv, q[0], q[1], q[2], f);
  int32_t si[3];
  for (int k = 0; k < 3; ++k) si[k] = as_i32(-floor(f[k]));
  double sv[3];
  matvec3(d->h, (double)(-si[0]), (double)(-si[1]), (double)(-si[2]), sv);
  const double wx = q[0] - sv[0], wy = q[1] - sv[1], wz = q[2] - sv[2];
  const int32_t loc = (int32_t)(a - d->atom_start);
  spos64[p] = Pos64{wx, wy, wz, __hiloint2double(0, loc)};  // w: index bits
  spos32[p] = make_float4((float)wx, (float)wy, (float)wz, __int_as_float(loc));
  sidx[p] = make_int4(si[0], si[1], si[2], loc);
}

// ---------------------------------------------------------------------------
// Pair search.  One wavefront per unit; a block holds kWaves units.
namespace {

// Home atoms staged per round of the home-atom loop.
constexpr int kHomeAtoms = 32;

// One wave's LDS: 4.1 KB in the FP64 fill pass, 3.6 KB in the FP64 count
// pass (so registers, not LDS, bound the waves per CU).
template <bool F32, bool FILL>
struct WaveLds {
  int32_t bs[3][64];                  // bin-image shift of the segment
  double off[3][64];                  // H bs in FP64, reference order
  float off32[F32 ? 3 : 1][F32 ? 64 : 1];  // H bs in FP32 (filter)
  // home atoms of the item, read at one uniform address per step (a
  // broadcast LDS read instead of a readlane per coordinate)
  union {
    Pos64 home64[kHomeAtoms];   // FP64 wrapped position, w = local index (FP64 kernels)
    float4 home32[kHomeAtoms];  // FP32 wrapped position, w = local index bits (FP32 kernels)
  };
  int4 homesi[FILL ? kHomeAtoms : 1];      // wrap shifts, w = local index (fill pass)
  int32_t mark[64];                   // candidate lane -> segment starting there (-1: none)
};

// q = v / m, r = v % m for 0 <= v < 2^22, 1 <= m < 2^22: a float estimate
// of the quotient (off by at most one there) and one exact correction step
// (integer division takes several times the instructions).
__device__ __forceinline__ void divmod_small(int v, int m, float inv_m, int* q, int* r) {
  int qq = (int)((float)v * inv_m);
  int rr = v - (int)__umul24((unsigned)qq, (unsigned)m);
  if (rr < 0) {
    qq -= 1;
    rr += m;
  } else if (rr >= m) {
    qq += 1;
    rr -= m;
  }
  *q = qq;
  *r = rr;
}

// The lists handled by one launch: q_lo .. q_lo+cnt-1 in ascending-cutoff
// order.  Held in registers for the whole wave.  For cumulative lists a
// pair belongs to list q iff d^2 < rc_q^2; with `disjoint` only to the
// first list it reaches (search.rs:837-849), i.e. also not to q-1.
// NQ: lists compiled into the kernel (1, 2, 3, 4 exactly; 8 = up to 8);
// lists past cnt have -inf thresholds and never take a pair.
template <int NQ>
struct Cuts {
  double r[NQ];         // rc^2 of the group's lists
  double prev;          // rc^2 of list q_lo-1 (-inf when q_lo == 0)
  float lo[NQ];         // FP32 certain-accept thresholds (per system)
  float hi[NQ];         // FP32 certain-reject thresholds
  float prev_lo, prev_hi;
  int64_t base[NQ];     // first edge of each list (fill pass)
  int64_t len[NQ];      // edges of each list
  int cnt;
  bool disjoint;
};

// The lists of one step as lane masks (bit l: the pair of lane l).  lt[g]:
// the lanes with d^2 < rc_g^2, nested since the cutoffs ascend.  A
// cumulative list takes lt[g]; a disjoint shell takes lt[g] without the
// lanes of the list before it (`below`: those of list q_lo - 1 for g = 0).
// cm: the lanes whose pair is a candidate at all.  Masks live in scalar
// registers: the per-list work of a step is scalar, not per lane.
template <bool DISJ, int NQ>
__device__ __forceinline__ void shells(const uint64_t* lt, uint64_t below, uint64_t cm,
                                       uint64_t* m) {
#pragma unroll
  for (int g = 0; g < NQ; ++g) {
    uint64_t x = lt[g] & cm;
    if (DISJ) x &= ~(g > 0 ? lt[g - 1] : below);
    m[g] = x;
  }
}

template <bool DISJ, int NQ>
__device__ __forceinline__ void lists64(double d2, const Cuts<NQ>& C, uint64_t cm, uint64_t* m) {
  uint64_t lt[NQ];
#pragma unroll
  for (int g = 0; g < NQ; ++g) lt[g] = __ballot(d2 < C.r[g]);
  const uint64_t below = DISJ ? __ballot(d2 < C.prev) : 0ull;
  shells<DISJ, NQ>(lt, below, cm, m);
}

// FP32 filter: per list the lanes certainly inside (e2 < lo); returns the
// lanes where some threshold lies within its band, whose decisions the
// FP64 expression takes (recheck64).
template <bool DISJ, int NQ>
__device__ __forceinline__ uint64_t filter32(float e2, const Cuts<NQ>& C, uint64_t* lt,
                                             uint64_t* below) {
  bool u = false;
#pragma unroll
  for (int g = 0; g < NQ; ++g) {
    const bool a = e2 < C.lo[g];
    lt[g] = __ballot(a);
    u |= !a && e2 <= C.hi[g];
  }
  if (DISJ) {
    const bool a = e2 < C.prev_lo;
    *below = __ballot(a);
    u |= !a && e2 <= C.prev_hi;
  } else {
    *below = 0ull;
  }
  return __ballot(u);
}

// The lanes of `need` decided again from their FP64 d^2 (+inf elsewhere).
template <bool DISJ, int NQ>
__device__ __forceinline__ void recheck64(double d2, const Cuts<NQ>& C, uint64_t need, uint64_t* lt,
                                          uint64_t* below) {
#pragma unroll
  for (int g = 0; g < NQ; ++g) lt[g] = (lt[g] & ~need) | __ballot(d2 < C.r[g]);
  if (DISJ) *below = (*below & ~need) | __ballot(d2 < C.prev);
}

// Count pass: edges per list so far, wave-uniform (scalar registers).
template <int NQ>
struct Counts {
  uint32_t c[NQ];
};

// Fill pass: per list the item's output, as 64-bit bases at the item's
// first edge in each array and a 32-bit count of the edges written since.
// A store takes the base (scalar) plus a 32-bit lane offset (slot x element
// size): no 64-bit address arithmetic per step.  The bases advance by the
// count once it passes kRebase (checked once per candidate chunk, which adds
// at most 64 x 64 x 2 x 2 edges per list), so a slot stays below 2^24 and
// its byte offsets are 24-bit multiplies (full rate; a 32-bit multiply runs
// at a quarter of it).
constexpr uint32_t kRebase = 1u << 23;

template <int NQ, bool DIST, bool DISP>
struct Out {
  int64_t* src[NQ];  // sources; targets at src + len of the list
  int64_t* dst[NQ];
  int32_t* sh[NQ];   // 3 per edge
  double* dist[NQ];  // d^2 per edge (DIST)
  double* disp[NQ];  // r_ij per edge, 3 doubles (DISP)
  uint32_t c[NQ];
  __device__ __forceinline__ void init(const PairArgs& A, const Cuts<NQ>& C, const int64_t* e0) {
#pragma unroll
    for (int g = 0; g < NQ; ++g) {
      src[g] = A.edges + C.base[g] + e0[g];
      dst[g] = src[g] + C.len[g];
      sh[g] = A.shifts + 3 * e0[g];
      dist[g] = DIST ? A.dist + e0[g] : nullptr;
      disp[g] = DISP ? A.disp + 3 * e0[g] : nullptr;
      c[g] = 0;
    }
  }
  __device__ __forceinline__ void rebase() {
#pragma unroll
    for (int g = 0; g < NQ; ++g) {
      if (c[g] >= kRebase) {
        src[g] += c[g];
        dst[g] += c[g];
        sh[g] += 3 * (int64_t)c[g];
        if (DIST) dist[g] += c[g];
        if (DISP) disp[g] += 3 * (int64_t)c[g];
        c[g] = 0;
      }
    }
  }
};

template <bool FILL, int NQ, bool DIST, bool DISP>
using Sink = typename std::conditional<FILL, Out<NQ, DIST, DISP>, Counts<NQ>>::type;

// Writes the pairs of the lanes in m, in lane order, at list g's cursor:
// (gi, gj, S) in a half list; (gi, gj, S) and (gj, gi, -S) at two
// consecutive slots in a full list (so every count is even and a pair goes
// out as 16-byte index stores and three 8-byte shift stores).  Indices are
// below 2^29 (at most 2^28 atoms per launch), so an index is its 32-bit
// value with a zero high word.  Plain stores: L2 merges the partial lines of
// successive steps (non-temporal stores measured 10% slower here).
template <int NQ, bool FULL, bool DIST, bool DISP>
__device__ __forceinline__ void put(Out<NQ, DIST, DISP>& O, int g, uint64_t m, uint32_t gi,
                                    uint32_t gj, int sx, int sy, int sz, double d2, double ex,
                                    double ey, double ez) {
  if (m == 0) return;
  if (__builtin_amdgcn_inverse_ballot_w64(m)) {
    const uint32_t e = O.c[g] + (FULL ? 2u : 1u) * lanes_below(m);  // this lane's slot
    const uint32_t e8 = e << 3;
    char* const s8 = reinterpret_cast<char*>(O.src[g]);
    char* const d8 = reinterpret_cast<char*>(O.dst[g]);
    char* const q12 = reinterpret_cast<char*>(O.sh[g]) + __umul24(e, 12u);
    if (!FULL) {
      *reinterpret_cast<int64_t*>(s8 + e8) = (int64_t)gi;
      *reinterpret_cast<int64_t*>(d8 + e8) = (int64_t)gj;
      int32_t* q = reinterpret_cast<int32_t*>(q12);  // one address, three offsets
      q[0] = sx;
      q[1] = sy;
      q[2] = sz;
      // squared here; square roots in a pass of their own
      if (DIST) *reinterpret_cast<double*>(reinterpret_cast<char*>(O.dist[g]) + e8) = d2;
      if (DISP) {
        double* r = reinterpret_cast<double*>(reinterpret_cast<char*>(O.disp[g]) + __umul24(e, 24u));
        r[0] = ex;
        r[1] = ey;
        r[2] = ez;
      }
    } else {
      *reinterpret_cast<longlong2*>(s8 + e8) = make_longlong2((int64_t)gi, (int64_t)gj);
      *reinterpret_cast<longlong2*>(d8 + e8) = make_longlong2((int64_t)gj, (int64_t)gi);
      int2* q = reinterpret_cast<int2*>(q12);
      q[0] = make_int2(sx, sy);
      q[1] = make_int2(sz, -sx);
      q[2] = make_int2(-sy, -sz);
      if (DIST)
        *reinterpret_cast<double2*>(reinterpret_cast<char*>(O.dist[g]) + e8) = make_double2(d2, d2);
      if (DISP) {
        double2* r = reinterpret_cast<double2*>(reinterpret_cast<char*>(O.disp[g]) + __umul24(e, 24u));
        r[0] = make_double2(ex, ey);
        r[1] = make_double2(ez, -ex);
        r[2] = make_double2(-ey, -ez);
      }
    }
  }
  O.c[g] += (FULL ? 2u : 1u) * (uint32_t)__popcll(m);
}


// Brute-force row i of a system (one image per pair, i < j in row-major
// order, search.rs:1187-1405).
template <bool FILL, int NQ, bool FULL, bool DIST, bool DISP, class SinkT>
__device__ __forceinline__ void unit_bruteforce(const PairArgs& A, const Cuts<NQ>& C,
                                                const SysDesc* d, int64_t u, int mode, SinkT& S) {
  const int lane = lane_id();
  const int64_t a0 = d->atom_start;
  const int n = d->n_atoms;
  const int i = (int)(u - a0);
  const double* pi = A.pos + 3 * u;
  const double xi = pi[0], yi = pi[1], zi = pi[2];
  const uint32_t gi = (uint32_t)(A.global_index ? u : (int64_t)i);
  const bool pb0 = d->pbc[0], pb1 = d->pbc[1], pb2 = d->pbc[2];
  double hv[9], hiv[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    hv[k] = d->h[k];
    hiv[k] = d->hinv[k];
  }
  for (int j0 = i + 1; j0 < n; j0 += 64) {
    const int j = j0 + lane;
    const bool valid = j < n;
    double d2 = INFINITY, dx = 0, dy = 0, dz = 0;
    int sx = 0, sy = 0, sz = 0;
    if (valid) {
      const double* pj = A.pos + 3 * (a0 + j);
      const double xj = pj[0], yj = pj[1], zj = pj[2];
      if (mode == kBruteIsolated) {
        // search.rs:1254-1258, 1278-1281
        dx = xj - xi;
        dy = yj - yi;
        dz = zj - zi;
      } else {
        // Cell::get_shift_and_displacement (cell.rs:74-94)
        double df[3];
        matvec3(hiv, xj - xi, yj - yi, zj - zi, df);
        const double f0 = pb0 ? -round(df[0]) : 0.0;
        const double f1 = pb1 ? -round(df[1]) : 0.0;
        const double f2 = pb2 ? -round(df[2]) : 0.0;
        sx = as_i32(f0);
        sy = as_i32(f1);
        sz = as_i32(f2);
        double v[3];
        matvec3(hv, f0, f1, f2, v);
        dx = (xj + v[0]) - xi;
        dy = (yj + v[1]) - yi;
        dz = (zj + v[2]) - zi;
      }
      d2 = norm2(dx, dy, dz);
    }
    uint64_t m[NQ];
    if (C.disjoint)
      lists64<true, NQ>(d2, C, ~0ull, m);
    else
      lists64<false, NQ>(d2, C, ~0ull, m);
    if constexpr (FILL) {
      uint64_t any = 0;
#pragma unroll
      for (int g = 0; g < NQ; ++g) any |= m[g];
      if (any == 0) continue;
      S.rebase();
      const uint32_t gj = (uint32_t)(A.global_index ? a0 + j : (int64_t)j);
#pragma unroll
      for (int g = 0; g < NQ; ++g)
        put<NQ, FULL>(S, g, m[g], gi, gj, sx, sy, sz, d2, dx, dy, dz);
    } else {
#pragma unroll
      for (int g = 0; g < NQ; ++g) S.c[g] += (uint32_t)__popcll(m[g]);
    }
  }
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Cell-list work item: up to G home atoms of one bin (a "home group", one per
// lane) and one slice of the half stencil.  Candidate atoms stream through
// the lanes 64 at a time, each loaded once per item and tested against every
// home atom in turn.
//
// Half stencil: every unordered (atom, atom, image) triple is examined once,
// from the bin whose stencil offset to the other is lexicographically
// positive (inside the home bin with no image: lower slot first).  The
// reference examines a pair from the lower original index with
// d = (w_j - w_i) + H bs (search.rs:540-543); from the other end the same
// pair gives (w_i - w_j) + H (-bs), the exact negation of every rounded term
// (IEEE subtraction and products are sign-symmetric), so d^2 and every
// decision are bit-identical.  Edges are written in the reference's
// orientation (lower original index first, S = s_j - s_i + bs from it,
// search.rs:695-697); a self image (same atom, bs != 0) is written for both
// S and -S, as the reference emits both (search.rs:549).
// Kinds of candidate chunk (see the home-atom loop of unit_cell).
constexpr int kPlain = 0, kSelf = 1, kHome = 2;

template <bool FILL, int NQ, bool F32, bool FULL, bool DIST, bool DISP, class Lds>
__device__ __forceinline__ void unit_cell(const PairArgs& A, const Cuts<NQ>& C, const SysDesc* d,
                                          int64_t p, int part, Lds& L,
                                          Sink<FILL, NQ, DIST, DISP>& S) {
  const int lane = lane_id();
  const int64_t a0 = d->atom_start;
  const int nbx = d->nb[0], nby = d->nb[1], nbz = d->nb[2];
  const int nsx = d->ns[0], nsy = d->ns[1], nsz = d->ns[2];
  const int64_t bb = d->bin_base;
  // the home bin's key (its Morton rank or linear index) and linear index
  const int kb = (int)((int64_t)A.keys[p] - bb);
  const int lb = A.rank_bin ? A.rank_bin[bb + kb] : kb;
  const int bin_lo = A.cell_start[bb + kb];
  // a group is up to G consecutive slots of one bin, aligned to the bin start
  if (((int)p - bin_lo) % A.group != 0) return;
  const int hs = (int)p, he = min(A.cell_start[bb + kb + 1], (int)p + A.group);
  const int bx = lb % nbx, by = (lb / nbx) % nby, bz = lb / (nbx * nby);
  const double h0 = d->h[0], h1 = d->h[1], h2 = d->h[2], h3 = d->h[3], h4 = d->h[4],
               h5 = d->h[5], h6 = d->h[6], h7 = d->h[7], h8 = d->h[8];
  const int wx = 2 * nsx + 1, wy = 2 * nsy + 1, wz = 2 * nsz + 1;
  const int T = wx * wy * wz;
  const int tc = (nsx * wy + nsy) * wz + nsz;  // the (0,0,0) offset
  // this item's slice of the half stencil [tc, T)
  const int slice = (T - tc + A.parts - 1) / A.parts;
  const int t_lo = tc + part * slice;
  const int t_hi = min(T, t_lo + slice);
  // stencil decode by float reciprocals when every operand is small
  const bool small = T < (1 << 21) && nbx + 2 * nsx < (1 << 20) && nby + 2 * nsy < (1 << 20) &&
                     nbz + 2 * nsz < (1 << 20);
  const float iwz = 1.0f / (float)wz, iwy = 1.0f / (float)wy;
  const float inx = 1.0f / (float)nbx, iny = 1.0f / (float)nby, inz = 1.0f / (float)nbz;
  // shift by whole periods so that the bin index is non-negative
  // an atom can meet its own image only when the stencil spans a period
  const bool self_possible = nsx >= nbx || nsy >= nby || nsz >= nbz;
  const int kx = (nsx + nbx - 1) / nbx + 1, ky = (nsy + nby - 1) / nby + 1,
            kz = (nsz + nbz - 1) / nbz + 1;
  // stencil pruning: perpendicular bin widths, cell shape, bound on d^2
  const float bwx = (float)(d->w[0] / (double)nbx), bwy = (float)(d->w[1] / (double)nby),
              bwz = (float)(d->w[2] / (double)nbz);
  const bool diag = d->h[1] == 0.0 && d->h[2] == 0.0 && d->h[3] == 0.0 && d->h[5] == 0.0 &&
                    d->h[6] == 0.0 && d->h[7] == 0.0;
  const float prune2 = (float)(d->rc_max * d->rc_max * (1.0 + 1e-4));
  const int64_t base_idx = A.global_index ? a0 : 0;
  for (int hb = hs; hb < he; hb += kHomeAtoms) {
    const int hn = uniform_i32(min(kHomeAtoms, he - hb));
    if (lane < hn) {
      if (F32)
        L.home32[lane] = A.spos32[hb + lane];
      else
        L.home64[lane] = A.spos64[hb + lane];
      if constexpr (FILL) L.homesi[lane] = A.sidx[hb + lane];
    }
    for (int t0 = t_lo; t0 < t_hi; t0 += 64) {
      const int t = t0 + lane;
      int len = 0, st = 0, bsx = 0, bsy = 0, bsz = 0;
      double ox = 0, oy = 0, oz = 0;
      if (t < t_hi) {
        int cx, cy, cz;
        int ox_, oy_, oz_;  // stencil offset in bins
        if (small) {
          int rest, dz, dy, dx;
          divmod_small(t, wz, iwz, &rest, &dz);
          divmod_small(rest, wy, iwy, &dx, &dy);
          ox_ = dx - nsx;
          oy_ = dy - nsy;
          oz_ = dz - nsz;
          divmod_small(bx + ox_ + kx * nbx, nbx, inx, &bsx, &cx);
          divmod_small(by + oy_ + ky * nby, nby, iny, &bsy, &cy);
          divmod_small(bz + oz_ + kz * nbz, nbz, inz, &bsz, &cz);
          bsx -= kx;
          bsy -= ky;
          bsz -= kz;
        } else {
          oz_ = t % wz - nsz;
          const int rest = t / wz;
          oy_ = rest % wy - nsy;
          ox_ = rest / wy - nsx;
          div_mod(bx + ox_, nbx, &cx, &bsx);
          div_mod(by + oy_, nby, &cy, &bsy);
          div_mod(bz + oz_, nbz, &cz, &bsz);
        }
        // Bins that cannot hold a partner: a home atom and a candidate
        // |o_k| bins apart along lattice axis k differ by at least
        // (|o_k| - 1) / nb_k in fractional coordinate k, hence by at least
        // (|o_k| - 1) w_k / nb_k in distance (w_k = 1/|row_k(H^-1)|).  The
        // per-axis bounds add in quadrature for a diagonal cell; otherwise
        // the largest one is the bound.  Evaluated in FP32 against a bound
        // widened by 1e-4 (relative), far beyond the FP32 rounding here and
        // that of the wrapped positions.
        const float gx = (float)max(0, abs(ox_) - 1) * bwx;
        const float gy = (float)max(0, abs(oy_) - 1) * bwy;
        const float gz = (float)max(0, abs(oz_) - 1) * bwz;
        const float lb2 = diag ? (gx * gx + gy * gy) + gz * gz
                               : fmaxf(fmaxf(gx * gx, gy * gy), gz * gz);
        if (lb2 <= prune2) {
          const int lin = cx + nbx * (cy + nby * cz);
          const int64_t bin = bb + (A.bin_rank ? A.bin_rank[bb + lin] : lin);
          st = A.cell_start[bin];
          const int en = A.cell_start[bin + 1];
          if (t == tc) st = hb;  // home bin, no image: slots from this group on
          len = en - st;
        }
        // (ox + oy) + oz of search.rs:498 = H (bsx, bsy, bsz) in gemv order
        ox = (h0 * (double)bsx + h1 * (double)bsy) + h2 * (double)bsz;
        oy = (h3 * (double)bsx + h4 * (double)bsy) + h5 * (double)bsz;
        oz = (h6 * (double)bsx + h7 * (double)bsy) + h8 * (double)bsz;
      }
      const int incl = wave_incl_scan_dpp(len);
      // candidate c of this segment list sits at sorted slot base[seg] + c
      const int base = st - (incl - len);
      L.bs[0][lane] = bsx;
      L.bs[1][lane] = bsy;
      L.bs[2][lane] = bsz;
      L.off[0][lane] = ox;
      L.off[1][lane] = oy;
      L.off[2][lane] = oz;
      if constexpr (F32) {
        const float g0 = d->h32[0], g1 = d->h32[1], g2 = d->h32[2], g3 = d->h32[3],
                    g4 = d->h32[4], g5 = d->h32[5], g6 = d->h32[6], g7 = d->h32[7],
                    g8 = d->h32[8];
        L.off32[0][lane] = (g0 * (float)bsx + g1 * (float)bsy) + g2 * (float)bsz;
        L.off32[1][lane] = (g3 * (float)bsx + g4 * (float)bsy) + g5 * (float)bsz;
        L.off32[2][lane] = (g6 * (float)bsx + g7 * (float)bsy) + g8 * (float)bsz;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // readlane, not a shuffle: a scalar, so the chunk loop below is uniform
      // (the fill cursors it carries stay scalar)
      const int total = __builtin_amdgcn_readlane(incl, 63);
      if (total > 0) {
        // Candidates stream through the lanes 64 at a time.  The loads of
        // the next chunk are issued before the current one is tested, so
        // their latency overlaps the home-atom loop.
        int sg = 0, j = 0;
        bool valid = false;
        float4 n32 = make_float4(0.f, 0.f, 0.f, 0.f);
        // the candidate's coordinates and local index (7 of the 8 words of
        // its Pos64) and wrap shift (3 of 4): fewer registers held across
        // the chunk for the next one
        double3 n64 = make_double3(0.0, 0.0, 0.0);
        int n64i = 0;
        int3 nij = make_int3(0, 0, 0);
        // This lane's candidate of the chunk at c0 lies in the last
        // non-empty segment starting at or before it.  Lane s holds segment
        // s.  The segment holding candidate c0 is one ballot; each segment
        // that starts inside the chunk marks its first candidate lane in LDS,
        // and a running maximum over the lanes (DPP) carries every mark
        // forward: a fixed cost per chunk, however many segments end in it.
        const int excl = incl - len;
        auto locate = [&](int c0) {
          const int idx = c0 + lane;
          valid = idx < total;
          const uint64_t before = __ballot(len > 0 && excl <= c0);
          const int s0 = before ? 63 - (int)__builtin_clzll(before) : 0;
          // marks hold segment + 1 (0: no segment starts at that lane)
          L.mark[lane] = 0;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          if (len > 0 && excl > c0 && excl <= c0 + 63) L.mark[excl - c0] = lane + 1;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          sg = max(s0, (int)wave_incl_umax_dpp((uint32_t)L.mark[lane]) - 1);
          j = __shfl(base, sg, 64) + idx;
        };
        auto fetch = [&]() {
          if (valid) {
            // j >= 0: an unsigned index needs no sign extension.  Pointer
            // arithmetic, not an integer address: the loads stay global
            // loads (a flat load also counts against the LDS wait of every
            // step of the home-atom loop, which would expose its latency)
            const uint64_t uj = (uint32_t)j;
            if (F32) {
              n32 = A.spos32[uj];
            } else {
              const Pos64* q = A.spos64 + uj;
              n64 = make_double3(q->x, q->y, q->z);
              n64i = reinterpret_cast<const int*>(q)[6];  // low word of w
            }
            if (FILL) {
              const int* t = reinterpret_cast<const int*>(A.sidx + uj);
              nij = make_int3(t[0], t[1], t[2]);
            }
          }
        };
        locate(0);
        fetch();
        for (int c0 = 0; c0 < total; c0 += 64) {
          const int csg = sg, cjs = j;
          const bool cvalid = valid;
          const float4 c32 = n32;
          double3 c = n64;
          const int ci = n64i;
          const int3 cj = nij;
          if (c0 + 64 < total) {
            locate(c0 + 64);
            fetch();
          }
          if constexpr (FILL) S.rebase();
          const bool home_seg = (t0 == tc) && (csg == 0);
          const int sbx = L.bs[0][csg], sby = L.bs[1][csg], sbz = L.bs[2][csg];
          const double cox = L.off[0][csg], coy = L.off[1][csg], coz = L.off[2][csg];
          // FP32 filter coordinates: candidate plus image offset, once per
          // candidate (the band covers this order of rounding as well)
          float px = 0.f, py = 0.f, pz = 0.f;
          if constexpr (F32) {
            px = c32.x + L.off32[0][csg];
            py = c32.y + L.off32[1][csg];
            pz = c32.z + L.off32[2][csg];
          }
          // lanes past the last candidate: a NaN coordinate fails every test
          if (!cvalid) {
            c.x = NAN;
            px = NAN;
          }
          const int oj = F32 ? __float_as_int(c32.w) : ci;
          // fill pass, per chunk: the candidate's part of S (s_j + bs) and its
          // emitted index
          const int tjx = cj.x + sbx, tjy = cj.y + sby, tjz = cj.z + sbz;
          const uint32_t gj_c = (uint32_t)(base_idx + oj);
          // home atom i is tested against this lane's candidate iff i < lim
          // (in the home bin without image only the later slots)
          const int lim = cvalid ? (home_seg ? cjs - hb : 64) : 0;
          // The home-atom loop, compiled per disjoint mode and per chunk kind:
          //   kPlain  every lane tests every home atom (no home-bin lanes,
          //           no own images possible)
          //   kSelf   own images possible: a step where a lane holds an
          //           image of the home atom counts / writes its -S copy
          //   kHome   home-bin lanes too: lanes at or before the home atom's
          //           slot get d^2 = +inf (no list takes them)
          auto home_loop = [&](auto disj_tag, auto kind_tag) {
            constexpr bool DISJ = decltype(disj_tag)::value;
            constexpr int KIND = decltype(kind_tag)::value;
#pragma unroll 1
            for (int i = 0; i < hn; ++i) {
              // broadcast LDS reads of the home atom (uniform address)
              float4 h32 = make_float4(0.f, 0.f, 0.f, 0.f);
              Pos64 h64 = {0.0, 0.0, 0.0, 0.0};
              if (F32)
                h32 = L.home32[i];
              else
                h64 = L.home64[i];
              const int oi = F32 ? __float_as_int(h32.w) : __double2loint(h64.w);
              const bool cand = KIND != kHome || i < lim;
              uint64_t m[NQ];
              double d2 = INFINITY, ex = 0.0, ey = 0.0, ez = 0.0;
              if (F32) {
                const float dx = px - h32.x, dy = py - h32.y, dz = pz - h32.z;
                float e2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                if (KIND == kHome && !cand) e2 = INFINITY;
                uint64_t lt[NQ], below;
                uint64_t need = filter32<DISJ, NQ>(e2, C, lt, &below);
                if (FILL && (DIST || DISP)) {
                  // exact distances of every pair taken
#pragma unroll
                  for (int g = 0; g < NQ; ++g) need |= lt[g];
                }
                if (need) {
                  const Pos64 hh = A.spos64[hb + i];  // uniform address
                  if (__builtin_amdgcn_inverse_ballot_w64(need)) {
                    // FP64 recheck: the reference expression, exactly
                    const Pos64 cc = A.spos64[cjs];
                    ex = (cc.x - hh.x) + cox;
                    ey = (cc.y - hh.y) + coy;
                    ez = (cc.z - hh.z) + coz;
                    d2 = norm2(ex, ey, ez);
                  }
                  recheck64<DISJ, NQ>(d2, C, need, lt, &below);
                }
                shells<DISJ, NQ>(lt, below, ~0ull, m);
              } else {
                // search.rs:540-543: d = (p_j - p_i) + offset, d^2 = (dx^2 + dy^2) + dz^2
                ex = (c.x - h64.x) + cox;
                ey = (c.y - h64.y) + coy;
                ez = (c.z - h64.z) + coz;
                d2 = norm2(ex, ey, ez);
                if (KIND == kHome && !cand) d2 = INFINITY;
                lists64<DISJ, NQ>(d2, C, ~0ull, m);
              }
              if constexpr (!FILL) {
                // A self image (same atom, other image) counts twice in a
                // half list (both S and -S are emitted); full lists count every
                // hit once and double the item's count at the end.
#pragma unroll
                for (int g = 0; g < NQ; ++g) S.c[g] += (uint32_t)__popcll(m[g]);
                if (KIND != kPlain && !FULL) {
                  const uint64_t self = __ballot(oj == oi);
                  if (self) {
#pragma unroll
                    for (int g = 0; g < NQ; ++g) S.c[g] += (uint32_t)__popcll(m[g] & self);
                  }
                }
              } else {
                uint64_t any = 0;
#pragma unroll
                for (int g = 0; g < NQ; ++g) any |= m[g];
                if (any == 0) continue;
                // S = s_j - s_i + bs in the reference orientation (lower
                // original index first): negated when the candidate is lower
                const int4 hsi = L.homesi[i];
                const bool flip = oj < oi;
                const int sgn = flip ? -1 : 0;  // x -> (x ^ sgn) - sgn negates when flipped
                const int sx = ((tjx - hsi.x) ^ sgn) - sgn;
                const int sy = ((tjy - hsi.y) ^ sgn) - sgn;
                const int sz = ((tjz - hsi.z) ^ sgn) - sgn;
                const uint32_t gi_h = (uint32_t)(base_idx + oi);
                const uint32_t gi = flip ? gj_c : gi_h;
                const uint32_t gj = flip ? gi_h : gj_c;
                // displacement in the edge's orientation
                if (DISP && flip) {
                  ex = -ex;
                  ey = -ey;
                  ez = -ez;
                }
#pragma unroll
                for (int g = 0; g < NQ; ++g)
                  put<NQ, FULL>(S, g, m[g], gi, gj, sx, sy, sz, d2, ex, ey, ez);
                // the self image's -S copy (full lists get it from the mirror)
                if (KIND != kPlain && !FULL) {
                  const uint64_t self = __ballot(oj == oi) & any;
                  if (self) {
#pragma unroll
                    for (int g = 0; g < NQ; ++g)
                      put<NQ, false>(S, g, m[g] & self, gi, gj, -sx, -sy, -sz, d2, -ex, -ey,
                                     -ez);
                  }
                }
              }
            }
          };
          using KPlain = std::integral_constant<int, kPlain>;
          using KSelf = std::integral_constant<int, kSelf>;
          using KHome = std::integral_constant<int, kHome>;
          const bool homes = __ballot(home_seg) != 0;
          if (C.disjoint) {
            if (homes)
              home_loop(std::true_type{}, KHome{});
            else if (self_possible)
              home_loop(std::true_type{}, KSelf{});
            else
              home_loop(std::true_type{}, KPlain{});
          } else {
            if (homes)
              home_loop(std::false_type{}, KHome{});
            else if (self_possible)
              home_loop(std::false_type{}, KSelf{});
            else
              home_loop(std::false_type{}, KPlain{});
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
}

}  // namespace

// Work items: system s owns items [item_base[s], item_base[s+1]): first its
// n_s atoms (brute-force rows when the system takes the brute-force path),
// then n_s * parts (sorted slot, stencil slice) items (home groups when it
// takes the cell-list path; slots that do not start a group are empty).
// Edges of one system are thus contiguous and systems come out in input
// order, as batch.rs:135-144 concatenates them.
template <bool FILL, int NQ, bool F32, bool FULL, bool DIST, bool DISP, class Lds>
__device__ __forceinline__ void pair_item(const PairArgs& A, int64_t u, Lds& L) {
  const int s = uniform_i32(find_sys(A.item_base, A.n_sys, u));
  const SysDesc* d = A.sys + s;
  const int mode = uniform_i32(d->mode);
  const int64_t local = u - A.item_base[s];
  // this launch's lists, in registers
  Cuts<NQ> C;
  C.cnt = A.q_cnt;
  C.disjoint = A.disjoint != 0;
  const double* rc2s = A.per_sys_rc ? A.rc2 + (int64_t)s * A.k : A.rc2;
  C.prev = A.q_lo > 0 ? rc2s[A.q_lo - 1] : -INFINITY;
  const float2* band = A.band + (int64_t)s * A.k;
  const float2 bp = A.q_lo > 0 ? band[A.q_lo - 1] : make_float2(-INFINITY, -INFINITY);
  C.prev_lo = bp.x;
  C.prev_hi = bp.y;
  int64_t e0[NQ];
#pragma unroll
  for (int g = 0; g < NQ; ++g) {
    const bool on = g < A.q_cnt;
    const int q = A.q_lo + (on ? g : 0);
    C.r[g] = on ? rc2s[q] : -INFINITY;
    const float2 b = band[q];
    C.lo[g] = on ? b.x : -INFINITY;
    C.hi[g] = on ? b.y : -INFINITY;
    C.base[g] = FILL && on ? A.bounds[q] : 0;
    C.len[g] = FILL && on ? A.bounds[q + 1] - A.bounds[q] : 0;
    e0[g] = (FILL && on) ? A.offsets[(int64_t)q * A.n + u] : 0;
  }
  Sink<FILL, NQ, DIST, DISP> S;
  if constexpr (FILL) {
    S.init(A, C, e0);
  } else {
#pragma unroll
    for (int g = 0; g < NQ; ++g) S.c[g] = 0;
  }
  const int n_atoms = d->n_atoms;
  if (local < n_atoms) {
    if (mode != kCellList)
      unit_bruteforce<FILL, NQ, FULL, DIST, DISP>(A, C, d, d->atom_start + local, mode, S);
  } else if (mode == kCellList) {
    const int64_t r = local - n_atoms;
    const int64_t p = d->atom_start + r / A.parts;
    unit_cell<FILL, NQ, F32, FULL, DIST, DISP>(A, C, d, p, (int)(r % A.parts), L, S);
  }
  if constexpr (!FILL) {
    // the counts are wave-uniform: lane 0 writes them
#pragma unroll
    for (int g = 0; g < NQ; ++g) {
      const uint32_t t = FULL ? 2u * S.c[g] : S.c[g];  // both directions of every hit
      if (g < A.q_cnt && lane_id() == 0) A.counts[(int64_t)(A.q_lo + g) * A.n + u] = t;
    }
  }
}
// Work list: the items that have work, appended with one atomic per wave
// and kind.  Home groups of cell-list systems (long: a whole half stencil)
// fill the list from the front, brute-force rows (short) from the back, so
// the short ones run last and fill the tail of the launch.  The pair
// kernels map wave w to list[w]: blocks are full of working waves (most
// item slots of a cell-list system start no group).  Which wave takes an
// item never changes its output: offsets are per item.
// count[0] = front length, count[1] = back length.  One atomic per block
// and kind (a single counter taking one atomic per wave serialises).
__global__ __launch_bounds__(1024) void k_items(const uint32_t* __restrict__ keys,
                                               const int32_t* __restrict__ cell_start,
                                               const SysDesc* __restrict__ sys,
                                               const int64_t* __restrict__ item_base, int n_sys,
                                               int64_t n_items, int group, int parts,
                                               int64_t* __restrict__ list,
                                               uint32_t* __restrict__ count) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool row = false, grp = false;
  if (u < n_items) {
    const int s = find_sys(item_base, n_sys, u);
    const SysDesc* d = sys + s;
    const int64_t local = u - item_base[s];
    const int n = d->n_atoms;
    if (local < n) {
      row = d->mode != kCellList;
    } else if (d->mode == kCellList) {
      const int64_t p = d->atom_start + (local - n) / parts;
      const int bin_lo = cell_start[keys[p]];
      grp = ((int)p - bin_lo) % group == 0;
    }
  }
  __shared__ uint32_t s_n[2][16], s_base[2];
  const uint64_t mg = __ballot(grp), mr = __ballot(row);
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) {
    s_n[0][w] = (uint32_t)__popcll(mg);
    s_n[1][w] = (uint32_t)__popcll(mr);
  }
  __syncthreads();
  const int nw = blockDim.x >> 6;
  if (threadIdx.x < 2) {
    uint32_t t = 0;
    for (int i = 0; i < nw; ++i) t += s_n[threadIdx.x][i];
    s_base[threadIdx.x] = t ? atomicAdd(count + threadIdx.x, t) : 0u;
  }
  __syncthreads();
  uint32_t bg = s_base[0], br = s_base[1];
  for (int i = 0; i < w; ++i) {
    bg += s_n[0][i];
    br += s_n[1][i];
  }
  if (grp) list[bg + lanes_below(mg)] = u;
  if (row) list[n_items - 1 - (br + lanes_below(mr))] = u;
}

// One wave per listed work item (a loop over several items per wave keeps
// the per-item state live across iterations and costs about 40 VGPRs).
// F32: FP32 filter with FP64 recheck in the band (cell-list items).
// FULL: full lists (count and fill passes).  DIST / DISP: the fill pass
// also stores squared distances / displacements.
//
// Blocks are dispatched round-robin over the 8 XCDs (block b to XCD b % 8),
// each with its own L2.  With xcd_chunk = c > 0 the work list is dealt to
// the XCDs in runs of 4c items in the fill pass and c items in the count
// pass (XCD x takes runs x, x + 8, ...), so neighbouring home groups, which
// read the same stencil bins, run under one L2 while every XCD still sees
// the whole mix of long and short items; c = 0 keeps the dispatch order.
// On the ragged batch (c = 16) the count ran 5 us faster with runs of 16
// items than of 64, the fill 8 us slower.  Only the blocks that have an
// item take part; the others leave at once.
//
// Fill blocks of several lists hold eight waves, other blocks four: the item
// slots past the work list take a block each way, and dispatching the
// ragged batch's 85,000 four-wave blocks costs ~19 us per launch.  On the
// ragged batch (three lists) eight-wave fill blocks measured 10 us faster
// per fill and eight-wave count blocks 14 us slower (a block keeps its
// slots until its slowest wave ends); on the serving trace (one list per
// system) eight-wave fill blocks cost 5 % of the capacity.
constexpr int kWavesFill = 8;
template <bool FILL, int NQ>
constexpr int waves_per_block() {
  return FILL && NQ > 1 ? kWavesFill : kWaves;
}

template <bool FILL, int NQ, bool F32, bool FULL, bool DIST, bool DISP>
__device__ __forceinline__ void pairs_body(const PairArgs& A) {
  using Lds = WaveLds<F32, FILL>;
  constexpr int WPB = waves_per_block<FILL, NQ>();
  __shared__ Lds s_lds[WPB];
  const int w = uniform_i32(threadIdx.x >> 6);
  const int64_t front = A.n_real[0];
  const int64_t total = front + (int64_t)A.n_real[1];
  const int64_t nblk = (total + WPB - 1) / WPB;
  const int64_t b = blockIdx.x;
  if (b >= nblk) return;
  constexpr int64_t kXcd = 8;
  int64_t bb = b;
  const int64_t run_items = (FILL ? 4 : 1) * (int64_t)A.xcd_chunk;
  const int64_t ch0 = run_items / WPB;  // blocks per run
  const int64_t ch = A.xcd_chunk > 0 ? (ch0 > 0 ? ch0 : 1) : 0;
  if (ch > 0) {
    const int64_t span = kXcd * ch;
    if (b < nblk / span * span) {  // the tail past whole spans keeps its order
      const int64_t q = b / kXcd;
      bb = (q / ch) * span + (b % kXcd) * ch + q % ch;
    }
  }
  const int64_t k = bb * WPB + w;
  if (k >= total) return;
  // wave-uniform item: its loads are scalar and the fill cursors stay scalar
  const int64_t u = uniform_i64(k < front ? A.list[k] : A.list[A.n - 1 - (k - front)]);
  // fill pass launched before the host saw the counts: write nothing when
  // the lists do not fit; the host re-runs the fill with exact buffers
  if (FILL && A.bounds[A.k] > A.capacity) return;
  pair_item<FILL, NQ, F32, FULL, DIST, DISP>(A, u, s_lds[w]);
}

template <bool FILL, int NQ, bool F32, bool FULL, bool DIST = false, bool DISP = false>
__global__ __launch_bounds__((64 * waves_per_block<FILL, NQ>())) void k_pairs(PairArgs A) {
  pairs_body<FILL, NQ, F32, FULL, DIST, DISP>(A);
}

// The same kernel held to 80 VGPRs (6 waves per SIMD; the count pass to 72,
// 7 waves per SIMD: 5 us less per count on the ragged batch, while 8 waves
// spill).  Used for the FP64 path without the full-list fill and without
// distances or displacements, where that fits without scratch and measured a
// few per cent faster.
template <bool FILL, int NQ, bool F32, bool FULL>
__global__ __launch_bounds__((64 * waves_per_block<FILL, NQ>())) __attribute__((amdgpu_waves_per_eu(FILL ? 6 : 7)))
void k_pairs_w6(
    PairArgs A) {
  pairs_body<FILL, NQ, F32, FULL, false, false>(A);
}

// ---------------------------------------------------------------------------
// Single small systems on the brute-force path (single.rs:60-89), without
// the pipeline's nine dependent launches, which cost such a system ~80 us of
// launch and dependency latency around a few microseconds of work:
//   k_small                 one block of 16 waves does everything (best up to
//                           ~130 atoms)
//   k_brute_count + _fill   one row per wave over ceil(n / 16) blocks; the
//                           last count block to finish scans (larger systems)
// All of them plan with plan_system and walk rows with unit_bruteforce in
// the pipeline's row order at the pipeline's offsets, so the lists are
// byte-identical to the pipeline's.  When the plan takes the cell-list path,
// bounds[k] = -1 and nothing else is written.

// In-place exclusive scan of len <= kGroup * kSmallRows u32 entries in LDS
// ([list][row]: offsets over all lists, as the pipeline's scan of
// [list][item] counts, where items off the path count zero) by a block of
// 1024 threads; returns the total.  Ends with the block synchronised.
__device__ __forceinline__ uint32_t small_scan(uint32_t* s_off, int len, uint32_t* s_wsum) {
  constexpr int kPer = kGroup * kSmallRows / 1024;  // entries per thread
  const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
  const int t0 = tid * kPer;
  uint32_t v[kPer], sum = 0;
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    v[e] = t0 + e < len ? s_off[t0 + e] : 0u;
    sum += v[e];
  }
  uint32_t incl = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) s_wsum[w] = incl;
  __syncthreads();
  uint32_t run = 0, total = 0;
  for (int x = 0; x < 16; ++x) {
    const uint32_t y = s_wsum[x];
    run += x < w ? y : 0u;
    total += y;
  }
  run += incl - sum;
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    if (t0 + e < len) s_off[t0 + e] = run;
    run += v[e];
  }
  __syncthreads();
  return total;
}

// The lists of k_small, k_brute_count and k_brute_fill in registers (all
// of them in one launch: k <= kGroup).
template <int NQ>
__device__ __forceinline__ Cuts<NQ> small_cuts(const double* rc2, int k, int disjoint) {
  Cuts<NQ> C;
  C.cnt = k;
  C.disjoint = disjoint != 0;
  C.prev = -INFINITY;
  C.prev_lo = C.prev_hi = -INFINITY;
#pragma unroll
  for (int g = 0; g < NQ; ++g) {
    C.r[g] = g < k ? rc2[g] : -INFINITY;
    C.lo[g] = C.hi[g] = -INFINITY;
    C.base[g] = C.len[g] = 0;
  }
  return C;
}

__device__ __forceinline__ PairArgs small_pair_args(const SmallArgs& a, const double* pos) {
  PairArgs A{};
  A.pos = pos;
  A.global_index = a.prm.global_index;
  A.edges = a.edges;
  A.shifts = a.shifts;
  A.dist = a.dist;
  A.disp = a.disp;
  return A;
}

// Count pass of row i by one wave: per-list counts, written by lane 0.
template <int NQ, bool FULL>
__device__ __forceinline__ void small_count_row(const PairArgs& A, const Cuts<NQ>& C,
                                                const SysDesc* d, int i, int mode, int k,
                                                int n, uint32_t* out) {
  Counts<NQ> cnt;
#pragma unroll
  for (int g = 0; g < NQ; ++g) cnt.c[g] = 0;
  unit_bruteforce<false, NQ, FULL, false, false>(A, C, d, (int64_t)i, mode, cnt);
#pragma unroll
  for (int g = 0; g < NQ; ++g) {
    const uint32_t t = FULL ? 2u * cnt.c[g] : cnt.c[g];  // wave-uniform
    if (g < k && lane_id() == 0) out[g * n + i] = t;
  }
}

// One block: plan, per-row counts, scan, fill.
template <int NQ, bool FULL, bool DIST, bool DISP>
__global__ __launch_bounds__(1024) void k_small(SmallArgs a) {
  __shared__ SysDesc s_d;
  __shared__ double s_rc2[kGroup];
  __shared__ double s_pos[3 * kSmallRows];         // the system's positions
  __shared__ uint32_t s_off[kGroup * kSmallRows];  // [list][row]: count, then offset
  __shared__ uint32_t s_wsum[16];
  __shared__ int64_t s_bounds[kGroup + 1];
  const int tid = threadIdx.x;
  if (tid == 0) {
    s_d = a.desc;
#pragma unroll
    for (int g = 0; g < kGroup; ++g) s_rc2[g] = a.rc2[g];
  }
  __syncthreads();
  plan_system(a.pos, &s_d, a.band, s_rc2, a.prm);
  __syncthreads();
  const int mode = s_d.mode;
  if (mode == kCellList) {
    if (tid == 0) a.bounds[a.k] = -1;
    return;
  }
  const int n = s_d.n_atoms;
  const int k = a.k;
  // the rows read every position many times: from LDS, not through the caches
  for (int e = tid; e < 3 * n; e += 1024) s_pos[e] = a.pos[e];
  __syncthreads();
  const PairArgs A = small_pair_args(a, s_pos);
  Cuts<NQ> C = small_cuts<NQ>(s_rc2, k, a.disjoint);
  const int w = uniform_i32(tid >> 6);
  // count pass: row i by wave i % 16
  for (int i = w; i < n; i += 16) small_count_row<NQ, FULL>(A, C, &s_d, i, mode, k, n, s_off);
  __syncthreads();
  const uint32_t total = small_scan(s_off, k * n, s_wsum);
  if (tid <= k) {
    const int64_t b = tid < k ? (int64_t)s_off[tid * n] : (int64_t)total;
    s_bounds[tid] = b;
    a.bounds[tid] = b;
  }
  __syncthreads();
  if ((int64_t)total > a.capacity) return;  // the caller re-runs with room
  // fill pass
#pragma unroll
  for (int g = 0; g < NQ; ++g) {
    C.base[g] = g < k ? s_bounds[g] : 0;
    C.len[g] = g < k ? s_bounds[g + 1] - s_bounds[g] : 0;
  }
  for (int i = w; i < n; i += 16) {
    int64_t e0[NQ];
#pragma unroll
    for (int g = 0; g < NQ; ++g) e0[g] = g < k ? (int64_t)s_off[g * n + i] : 0;
    Out<NQ, DIST, DISP> O;
    O.init(A, C, e0);
    unit_bruteforce<true, NQ, FULL, DIST, DISP>(A, C, &s_d, (int64_t)i, mode, O);
  }
  if (DIST) {
    // the fill stored squared distances (as the pipeline's fill pass); the
    // block's stores are visible to it after the barrier
    __syncthreads();
    for (int64_t e = tid; e < (int64_t)total; e += 1024) a.dist[e] = sqrt(a.dist[e]);
  }
}

// Count launch: row 16 b + w by wave w of block b.  Every block plans the
// system itself (the same routine on the same inputs: the same plan) and
// stages the positions in LDS; the last block to finish scans all counts,
// publishes offsets, bounds and the planned descriptor, and re-arms the
// ticket for the next call.
template <int NQ, bool FULL>
__global__ __launch_bounds__(1024) void k_brute_count(SmallArgs a) {
  __shared__ SysDesc s_d;
  __shared__ double s_rc2[kGroup];
  __shared__ double s_pos[3 * kSmallRows];
  __shared__ uint32_t s_off[kGroup * kSmallRows];
  __shared__ uint32_t s_wsum[16];
  __shared__ uint32_t s_last;
  const int tid = threadIdx.x;
  if (tid == 0) {
    s_d = a.desc;
#pragma unroll
    for (int g = 0; g < kGroup; ++g) s_rc2[g] = a.rc2[g];
  }
  __syncthreads();
  plan_system(a.pos, &s_d, a.band, s_rc2, a.prm);
  __syncthreads();
  const int mode = s_d.mode;
  if (mode == kCellList) {
    if (blockIdx.x == 0 && tid == 0) a.bounds[a.k] = -1;
    return;
  }
  const int n = s_d.n_atoms;
  const int k = a.k;
  for (int e = tid; e < 3 * n; e += 1024) s_pos[e] = a.pos[e];
  __syncthreads();
  const PairArgs A = small_pair_args(a, s_pos);
  const Cuts<NQ> C = small_cuts<NQ>(s_rc2, k, a.disjoint);
  const int i = (int)blockIdx.x * 16 + uniform_i32(tid >> 6);
  if (i < n) small_count_row<NQ, FULL>(A, C, &s_d, i, mode, k, n, a.counts);
  // publish the counts, then take a ticket
  __threadfence();
  __syncthreads();
  if (tid == 0) s_last = atomicAdd(a.ticket, 1u) == gridDim.x - 1 ? 1u : 0u;
  __syncthreads();
  if (!s_last) return;
  __threadfence();  // every other block's counts are visible past its ticket
  const int len = k * n;
  for (int e = tid; e < len; e += 1024) s_off[e] = __atomic_load_n(a.counts + e, __ATOMIC_RELAXED);
  __syncthreads();
  const uint32_t total = small_scan(s_off, len, s_wsum);
  for (int e = tid; e < len; e += 1024) a.offs[e] = s_off[e];
  if (tid <= k) a.bounds[tid] = tid < k ? (int64_t)s_off[tid * n] : (int64_t)total;
  if (tid == 0) {
    *a.planned = s_d;
    *a.ticket = 0u;
  }
}

// Fill launch: the rows of the count launch at the scanned offsets; nothing
// when the plan declined or the lists exceed the outputs.
template <int NQ, bool FULL, bool DIST, bool DISP>
__global__ __launch_bounds__(1024) void k_brute_fill(SmallArgs a) {
  __shared__ SysDesc s_d;
  __shared__ double s_pos[3 * kSmallRows];
  const int64_t total = a.bounds[a.k];
  if (total < 0 || total > a.capacity) return;
  const int tid = threadIdx.x;
  if (tid == 0) s_d = *a.planned;
  __syncthreads();
  const int n = s_d.n_atoms;
  const int k = a.k;
  for (int e = tid; e < 3 * n; e += 1024) s_pos[e] = a.pos[e];
  __syncthreads();
  const int i = (int)blockIdx.x * 16 + uniform_i32(tid >> 6);
  if (i >= n) return;
  const PairArgs A = small_pair_args(a, s_pos);
  Cuts<NQ> C = small_cuts<NQ>(a.rc2, k, a.disjoint);
  int64_t e0[NQ];
#pragma unroll
  for (int g = 0; g < NQ; ++g) {
    C.base[g] = g < k ? a.bounds[g] : 0;
    C.len[g] = g < k ? a.bounds[g + 1] - a.bounds[g] : 0;
    e0[g] = g < k ? (int64_t)a.offs[g * n + i] : 0;
  }
  Out<NQ, DIST, DISP> O;
  O.init(A, C, e0);
  unit_bruteforce<true, NQ, FULL, DIST, DISP>(A, C, &s_d, (int64_t)i, s_d.mode, O);
}

// ---------------------------------------------------------------------------
// Batches of small systems that all take a brute-force path, without the
// pipeline: the count launch gives each system ceil(n_s / kBatchRows) blocks
// of kBatchRows rows (one row per wave); every block plans its own system
// (plan_system on the same inputs: the same plan) and counts its rows into
// [list][atom].  One block then scans all counts, and the fill launch writes
// the rows at the scanned offsets.  Rows, their order and their offsets are
// the pipeline's (per list: systems in order, rows in order; cell-list items
// of brute-force systems count zero there), so the lists are byte-identical
// to it.  A system whose plan takes the cell-list path sets a flag; the scan
// then writes bounds[k] = -1 and the fill writes nothing.  No block waits for
// another: with a last-block ticket instead of the scan launch (each block
// fencing its counts out of its XCD's L2 before taking the ticket) the count
// launch of 384 blocks took 118 us, without it 21 us.  Blocks are small
// because each pays the plan's latency once: 16-wave blocks fit one per CU
// at these registers, so 384 of them ran in two rounds.

template <int NQ, bool FULL>
__global__ __launch_bounds__(64 * kBatchRows) void k_batch_brute_count(BatchSmallArgs a) {
  __shared__ SysDesc s_d;
  __shared__ double s_rc2[kGroup];
  const int tid = threadIdx.x;
  const int b = (int)blockIdx.x;
  const int s = uniform_i32(a.blk_sys[b]);
  if (tid == 0) {
    s_d = a.sys[s];
#pragma unroll
    for (int g = 0; g < kGroup; ++g) s_rc2[g] = a.rc2[g];
  }
  __syncthreads();
  plan_system(a.pos, &s_d, a.band + (int64_t)s * a.k, s_rc2, a.prm);
  __syncthreads();
  const int mode = s_d.mode;
  if (mode == kCellList) {
    if (tid == 0) atomicOr(a.flag, 1u);
    return;
  }
  const int64_t a0 = s_d.atom_start;
  const int i = (b - a.blk_start[s]) * kBatchRows + uniform_i32(tid >> 6);
  if (b == a.blk_start[s] && tid == 0) a.planned[s] = s_d;
  if (i >= s_d.n_atoms) return;
  PairArgs A{};
  A.pos = a.pos;
  A.global_index = a.prm.global_index;
  const Cuts<NQ> C = small_cuts<NQ>(s_rc2, a.k, a.disjoint);
  Counts<NQ> cnt;
#pragma unroll
  for (int g = 0; g < NQ; ++g) cnt.c[g] = 0;
  unit_bruteforce<false, NQ, FULL, false, false>(A, C, &s_d, a0 + i, mode, cnt);
#pragma unroll
  for (int g = 0; g < NQ; ++g) {
    const uint32_t t = FULL ? 2u * cnt.c[g] : cnt.c[g];  // wave-uniform
    if (g < a.k && lane_id() == 0) a.counts[(int64_t)g * a.n + a0 + i] = t;
  }
}

// One block: exclusive scan of the [list][atom] counts, kChunk entries at a
// time, into offsets and list bounds (or bounds[k] = -1 when a system took
// the cell-list path; the flag is re-armed for the next call).
__global__ __launch_bounds__(1024) void k_batch_scan(BatchSmallArgs a) {
  constexpr int kChunk = kGroup * kSmallRows;  // entries small_scan takes at once
  __shared__ uint32_t s_off[kChunk];
  __shared__ uint32_t s_wsum[16];
  __shared__ int64_t s_b[kGroup];
  __shared__ uint32_t s_flag;
  const int tid = threadIdx.x;
  if (tid == 0) s_flag = *a.flag;
  __syncthreads();
  if (s_flag != 0u) {
    if (tid == 0) {
      a.bounds[a.k] = -1;
      *a.flag = 0u;
    }
    return;
  }
  const int64_t len = (int64_t)a.k * a.n;
  uint64_t carry = 0;
  for (int64_t c0 = 0; c0 < len; c0 += kChunk) {
    const int m = (int)(len - c0 < kChunk ? len - c0 : kChunk);
    for (int e = tid; e < m; e += 1024) s_off[e] = a.counts[c0 + e];
    __syncthreads();
    const uint32_t tot = small_scan(s_off, m, s_wsum);
    for (int e = tid; e < m; e += 1024) a.offs[c0 + e] = s_off[e] + (uint32_t)carry;
    if (tid < a.k) {
      const int64_t p = (int64_t)tid * a.n;  // first entry of list tid
      if (p >= c0 && p < c0 + m) s_b[tid] = (int64_t)carry + s_off[p - c0];
    }
    carry += tot;
    __syncthreads();
  }
  if (tid <= a.k) a.bounds[tid] = tid < a.k ? s_b[tid] : (int64_t)carry;
}

template <int NQ, bool FULL, bool DIST, bool DISP>
__global__ __launch_bounds__(64 * kBatchRows) void k_batch_brute_fill(BatchSmallArgs a) {
  __shared__ SysDesc s_d;
  const int64_t total = a.bounds[a.k];
  if (total < 0 || total > a.capacity) return;
  const int tid = threadIdx.x;
  const int b = (int)blockIdx.x;
  const int s = uniform_i32(a.blk_sys[b]);
  if (tid == 0) s_d = a.planned[s];
  __syncthreads();
  const int i = (b - a.blk_start[s]) * kBatchRows + uniform_i32(tid >> 6);
  if (i >= s_d.n_atoms) return;
  const int64_t u = s_d.atom_start + i;
  PairArgs A{};
  A.pos = a.pos;
  A.global_index = a.prm.global_index;
  A.edges = a.edges;
  A.shifts = a.shifts;
  A.dist = a.dist;
  A.disp = a.disp;
  Cuts<NQ> C = small_cuts<NQ>(a.rc2, a.k, a.disjoint);
  int64_t e0[NQ];
#pragma unroll
  for (int g = 0; g < NQ; ++g) {
    C.base[g] = g < a.k ? a.bounds[g] : 0;
    C.len[g] = g < a.k ? a.bounds[g + 1] - a.bounds[g] : 0;
    e0[g] = g < a.k ? (int64_t)a.offs[(int64_t)g * a.n + u] : 0;
  }
  Out<NQ, DIST, DISP> O;
  O.init(A, C, e0);
  unit_bruteforce<true, NQ, FULL, DIST, DISP>(A, C, &s_d, u, s_d.mode, O);
}

// ---------------------------------------------------------------------------
// The count scan runs after the count pass, the last reader of the work
// list's lengths n_real = rearm[0..1]: block 0 hands them on to the fill
// pass in rearm[2..3] and clears them for the next build (a graph replay
// then needs no clearing launch before its build).
__device__ __forceinline__ void rearm_list(uint32_t* rearm) {
  if (rearm && blockIdx.x == 0 && threadIdx.x == 0) {
    rearm[2] = rearm[0];
    rearm[3] = rearm[1];
    rearm[0] = 0u;
    rearm[1] = 0u;
  }
}
__global__ void k_rearm(uint32_t* rearm) { rearm_list(rearm); }

// Exclusive scan of u32 counts into int64 offsets, decoupled look-back.
// bounds[q] = offset of element q*seg (q < nseg), bounds[nseg] = total.
__global__ __launch_bounds__(256) void k_scan(const uint32_t* __restrict__ in,
                                              int64_t* __restrict__ out, int64_t len,
                                              int64_t seg, int nseg, int64_t* __restrict__ bounds,
                                              uint64_t* __restrict__ status,
                                              uint32_t* __restrict__ ticket,
                                              uint32_t* __restrict__ rearm) {
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_w[4];
  __shared__ uint64_t s_excl;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  rearm_list(rearm);
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const int64_t e0 = (int64_t)tile * kRadixTile + 8 * tid;
  uint64_t v[8];
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    v[i] = (e0 + i < len) ? (uint64_t)in[e0 + i] : 0ull;
    sum += v[i];
  }
  const uint64_t winc = wave_incl_scan(sum);
  if (lane == 63) s_w[w] = winc;
  __syncthreads();
  uint64_t pre = winc - sum;
  for (int i = 0; i < w; ++i) pre += s_w[i];
  const uint64_t tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
  // decoupled look-back by the first wavefront: 64 predecessors per probe
  if (w == 0) {
    uint64_t excl = 0;
    if (tile == 0) {
      if (lane == 0) atomic_store_agent(status, kFlagPre64 | tot);
    } else {
      if (lane == 0) atomic_store_agent(status + tile, kFlagAgg64 | tot);
      int64_t t_hi = (int64_t)tile - 1;  // newest predecessor of this window
      while (true) {
        const int64_t t = t_hi - lane;
        uint64_t x = t >= 0 ? atomic_load_agent(status + t) : kFlagPre64;  // before 0: prefix 0
        const uint64_t flag = x & ~kValMask64;
        const uint64_t pre_mask = __ballot(flag == kFlagPre64);
        const uint64_t zero_mask = __ballot(flag == 0);
        // lanes up to and including the nearest inclusive prefix must be ready
        const int first_pre = pre_mask ? __builtin_ctzll(pre_mask) : 64;
        const uint64_t need = first_pre >= 63 ? ~0ull : ((1ull << (first_pre + 1)) - 1);
        if (zero_mask & need) continue;  // a predecessor is still ranking: probe again
        uint64_t v = (lane <= first_pre) ? (x & kValMask64) : 0ull;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        excl += v;
        if (first_pre < 64) break;
        t_hi -= 64;
      }
      if (lane == 0) atomic_store_agent(status + tile, kFlagPre64 | (excl + tot));
    }
    if (lane == 0) s_excl = excl;
  }
  __syncthreads();
  uint64_t run = s_excl + pre;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int64_t e = e0 + i;
    if (e < len) {
      out[e] = (int64_t)run;
      if (e % seg == 0) bounds[e / seg] = (int64_t)run;
      run += v[i];
      if (e == len - 1) {
        out[len] = (int64_t)run;
        bounds[nseg] = (int64_t)run;
      }
    }
  }
}

// The same scan for long inputs, reduce-then-scan: tile sums, one
// exclusive scan over them, then each tile's own scan from its offset.  No
// block waits on another (a look-back chain across hundreds of tiles costs
// more than the two extra launches).
__device__ inline uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_scan_sums(const uint32_t* __restrict__ in, int64_t len,
                                                   uint64_t* __restrict__ sums) {
  __shared__ uint64_t s_w[4];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int64_t e0 = (int64_t)blockIdx.x * kRadixTile + 8 * tid;
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) sum += (e0 + i < len) ? (uint64_t)in[e0 + i] : 0ull;
  sum = wave_sum_u64(sum);
  if (lane == 0) s_w[w] = sum;
  __syncthreads();
  if (tid == 0) sums[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// In-place exclusive scan of the n tile sums, one block, 4096 per round.
__global__ __launch_bounds__(1024) void k_scan_sums_excl(uint64_t* __restrict__ sums, int64_t n) {
  __shared__ uint64_t s_w[16];
  __shared__ uint64_t s_carry;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  if (tid == 0) s_carry = 0;
  __syncthreads();
  for (int64_t b = 0; b < n; b += 4096) {
    const int64_t e0 = b + 4 * tid;
    uint64_t v[4];
    uint64_t t = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = (e0 + i < n) ? sums[e0 + i] : 0ull;
      t += v[i];
    }
    const uint64_t winc = wave_incl_scan(t);
    if (lane == 63) s_w[w] = winc;
    __syncthreads();
    uint64_t run = s_carry + winc - t;
    uint64_t all = 0;
    for (int i = 0; i < 16; ++i) {
      if (i < w) run += s_w[i];
      all += s_w[i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (e0 + i < n) sums[e0 + i] = run;
      run += v[i];
    }
    __syncthreads();
    if (tid == 0) s_carry += all;
    __syncthreads();
  }
}

// Tile `blockIdx.x` of the scan, starting from its offset pre[tile].
// bounds[q] = out[q * seg] for q < nseg; bounds[nseg] = out[len] = total.
// RAW: pre holds the tile sums themselves, and the block adds up those
// before its tile (up to kScanRawTiles tiles: one launch fewer than the
// separate scan of the sums).
constexpr int64_t kScanRawTiles = 2048;
template <bool RAW>
__global__ __launch_bounds__(256) void k_scan_apply(const uint32_t* __restrict__ in,
                                                    int64_t* __restrict__ out, int64_t len,
                                                    int64_t seg, int nseg,
                                                    int64_t* __restrict__ bounds,
                                                    const uint64_t* __restrict__ pre,
                                                    uint32_t* __restrict__ rearm) {
  __shared__ uint64_t s_w[4];
  __shared__ uint64_t s_pre;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  rearm_list(rearm);
  const int64_t t0 = (int64_t)blockIdx.x * kRadixTile;
  const int64_t e0 = t0 + 8 * tid;
  if constexpr (RAW) {
    uint64_t part = 0;
    for (int64_t t = tid; t < (int64_t)blockIdx.x; t += 256) part += pre[t];
    part = wave_sum_u64(part);
    if (lane == 0) s_w[w] = part;
    __syncthreads();
    if (tid == 0) s_pre = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    __syncthreads();
  }
  uint64_t v[8];
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    v[i] = (e0 + i < len) ? (uint64_t)in[e0 + i] : 0ull;
    sum += v[i];
  }
  const uint64_t winc = wave_incl_scan(sum);
  if (lane == 63) s_w[w] = winc;
  __syncthreads();
  uint64_t run = (RAW ? s_pre : pre[blockIdx.x]) + winc - sum;
  for (int i = 0; i < w; ++i) run += s_w[i];
  // list bounds that fall in this tile (block-uniform range of q)
  const int64_t t1 = min(t0 + (int64_t)kRadixTile, len);
  const int64_t q_lo = (t0 + seg - 1) / seg;
  const int64_t q_hi = min((int64_t)nseg, (t1 + seg - 1) / seg);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int64_t e = e0 + i;
    if (e < len) {
      out[e] = (int64_t)run;
      for (int64_t q = q_lo; q < q_hi; ++q)
        if (q * seg == e) bounds[q] = (int64_t)run;
      run += v[i];
      if (e == len - 1) {
        out[len] = (int64_t)run;
        bounds[nseg] = (int64_t)run;
      }
    }
  }
}

// Largest squared displacement |pos - ref|^2 over n atoms, into *out as the
// bit pattern of a non-negative double (ordered like the value, so an
// unsigned atomic max finds it).  *out must be zero before the launch.
__global__ __launch_bounds__(256) void k_max_disp2(const double* __restrict__ pos,
                                                   const double* __restrict__ ref, int64_t n,
                                                   unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s_w[4];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  unsigned long long best = 0ull;
  for (int64_t a = (int64_t)blockIdx.x * blockDim.x + tid; a < n;
       a += (int64_t)gridDim.x * blockDim.x) {
    const double dx = pos[3 * a] - ref[3 * a], dy = pos[3 * a + 1] - ref[3 * a + 1],
                 dz = pos[3 * a + 2] - ref[3 * a + 2];
    const double d2 = (dx * dx + dy * dy) + dz * dz;
    // NaN positions count as an unbounded displacement
    const unsigned long long b = d2 == d2 ? (unsigned long long)__double_as_longlong(d2)
                                          : 0x7ff0000000000000ull;
    best = b > best ? b : best;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long t = __shfl_xor(best, o, 64);
    best = t > best ? t : best;
  }
  if (lane == 0) s_w[w] = best;
  __syncthreads();
  if (tid == 0) {
    unsigned long long m = s_w[0];
    for (int i = 1; i < 4; ++i) m = s_w[i] > m ? s_w[i] : m;
    if (m) atomicMax(out, m);
  }
}

// Per-system edge ranges: ranges[(s*k + q)*2 + {0,1}].
__global__ void k_sys_ranges(const int64_t* __restrict__ off,
                             const int64_t* __restrict__ item_base, int n_sys, int k, int64_t n,
                             int64_t* __restrict__ ranges) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_sys * k) return;
  const int s = (int)(t / k), q = (int)(t % k);
  ranges[2 * t] = off[(int64_t)q * n + item_base[s]];
  ranges[2 * t + 1] = off[(int64_t)q * n + item_base[s + 1]];
}

// ---------------------------------------------------------------------------
// launch wrappers
static inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)g), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_sqrt_dist(double* dist, const int64_t* total, int64_t capacity,
                            hipStream_t st) {
  if (capacity <= 0) return hipSuccess;
  int64_t g = (capacity + 255) / 256;
  if (g > 2048) g = 2048;
  hipLaunchKernelGGL(k_sqrt_dist, dim3((unsigned)g), dim3(256), 0, st, dist, total, capacity);
  return hipGetLastError();
}

hipError_t launch_zero2(uint32_t* a, int64_t na, uint32_t* b, int64_t nb, hipStream_t st,
                        uint32_t* dst, const uint32_t* src, int64_t nc) {
  if (!dst || !src) nc = 0;
  if (na + nb + nc <= 0) return hipSuccess;
  int64_t g = (na + nb + nc + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_zero2, dim3((unsigned)g), dim3(256), 0, st, a, na, b, nb, dst, src, nc);
  return hipGetLastError();
}

hipError_t launch_plan(const double* pos, SysDesc* sys, float2* band, const double* rc2,
                       const Params& p, hipStream_t st) {
  if (p.n_sys <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_plan, dim3(p.n_sys), dim3(1024), 0, st, pos, sys, band, rc2, p);
  return hipGetLastError();
}

hipError_t launch_bin(const double* pos, const SysDesc* sys, const int64_t* sys_start, int n_sys,
                      int64_t n, const int32_t* bin_rank, uint32_t* keys, uint32_t* vals,
                      hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bin, dim3(grid_for(n, 256)), dim3(256), 0, st, pos, sys, sys_start, n_sys,
                     n, bin_rank, keys, vals);
  return hipGetLastError();
}

hipError_t launch_morton(const SysDesc* sys, int n_sys, int64_t total_bins, int32_t* bin_rank,
                         int32_t* rank_bin, hipStream_t st) {
  if (total_bins <= 0 || n_sys <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton, dim3(grid_for(total_bins, 256)), dim3(256), 0, st, sys, n_sys,
                     total_bins, bin_rank, rank_bin);
  return hipGetLastError();
}

hipError_t launch_radix_pass(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                             uint32_t* vout, int64_t n, int shift, int bits, uint32_t* cnt,
                             int64_t* off, int64_t* bounds, uint64_t* status, uint32_t* ticket,
                             hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (bits < 1 || bits > kRadixMaxBits) return hipErrorInvalidValue;
  const int64_t tiles = radix_tiles(n);
  if (tiles <= 2) {  // one block ranks the tiles in order: one launch
    // (one block is slower than three launches from about 3 tiles on)
    hipLaunchKernelGGL(k_radix_small, dim3(1), dim3(256), 0, st, kin, vin, kout, vout, n, shift,
                       bits);
    return hipGetLastError();
  }
  const int64_t len = ((int64_t)1 << bits) * tiles;
  hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)tiles), dim3(256), 0, st, kin, n, shift, bits,
                     tiles, cnt);
  const hipError_t e = launch_scan(cnt, off, len, len, 1, bounds, status, ticket, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)tiles), dim3(256), 0, st, kin, vin, kout,
                     vout, n, shift, bits, tiles, (const int64_t*)off);
  return hipGetLastError();
}

hipError_t launch_finalize(const uint32_t* keys, const uint32_t* vals, const double* pos,
                           const SysDesc* sys, const int64_t* sys_start, int n_sys, int64_t n,
                           int64_t total_bins, int32_t* cell_start, Pos64* spos64,
                           float4* spos32, int4* sidx, bool with_cells, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int64_t threads = (with_cells && total_bins + 1 > n) ? total_bins + 1 : n;
  hipLaunchKernelGGL(k_finalize, dim3(grid_for(threads, 256)), dim3(256), 0, st, keys, vals, pos, sys,
                     sys_start, n_sys, n, total_bins, cell_start, spos64, spos32, sidx,
                     with_cells ? 1 : 0);
  return hipGetLastError();
}

// Fill kernels by (FP32 filter, full lists, distances, displacements); the
// plain FP64 half-list fill is the 6-waves-per-SIMD build.
template <int NQ, bool F32, bool FULL, bool DIST, bool DISP>
static void fill_launch(const dim3& g, const dim3& b, const PairArgs& a, hipStream_t st) {
  if constexpr (!F32 && !FULL && !DIST && !DISP)
    hipLaunchKernelGGL((k_pairs_w6<true, NQ, false, false>), g, b, 0, st, a);
  else
    hipLaunchKernelGGL((k_pairs<true, NQ, F32, FULL, DIST, DISP>), g, b, 0, st, a);
}
template <int NQ, bool F32, bool FULL, bool DIST>
static void fill_disp(const dim3& g, const dim3& b, const PairArgs& a, hipStream_t st) {
  if (a.disp)
    fill_launch<NQ, F32, FULL, DIST, true>(g, b, a, st);
  else
    fill_launch<NQ, F32, FULL, DIST, false>(g, b, a, st);
}
template <int NQ, bool F32, bool FULL>
static void fill_dist(const dim3& g, const dim3& b, const PairArgs& a, hipStream_t st) {
  if (a.dist)
    fill_disp<NQ, F32, FULL, true>(g, b, a, st);
  else
    fill_disp<NQ, F32, FULL, false>(g, b, a, st);
}
template <int NQ, bool F32>
static void fill_full(const dim3& g, const dim3& b, const PairArgs& a, hipStream_t st) {
  if (a.full)
    fill_dist<NQ, F32, true>(g, b, a, st);
  else
    fill_dist<NQ, F32, false>(g, b, a, st);
}

template <int NQ>
static void launch_pairs_nq(bool fill, const PairArgs& a, hipStream_t st) {
  // one wave per item that can have work (the blocks past the work list
  // leave at once, but each costs dispatch time)
  const int64_t items = a.max_items > 0 ? std::min(a.max_items, a.n) : a.n;
  if (fill) {
    constexpr int wf = waves_per_block<true, NQ>();
    const dim3 g(grid_for(items, wf)), b(64 * wf);
    if (a.use_fp32)
      fill_full<NQ, true>(g, b, a, st);
    else
      fill_full<NQ, false>(g, b, a, st);
    return;
  }
  const dim3 g(grid_for(items, kWaves)), b(64 * kWaves);
  if (a.use_fp32) {
    if (a.full)
      hipLaunchKernelGGL((k_pairs<false, NQ, true, true>), g, b, 0, st, a);
    else
      hipLaunchKernelGGL((k_pairs<false, NQ, true, false>), g, b, 0, st, a);
  } else {
    if (a.full)
      hipLaunchKernelGGL((k_pairs_w6<false, NQ, false, true>), g, b, 0, st, a);
    else
      hipLaunchKernelGGL((k_pairs_w6<false, NQ, false, false>), g, b, 0, st, a);
  }
}

hipError_t launch_items(const uint32_t* keys, const int32_t* cell_start, const SysDesc* sys,
                        const int64_t* item_base, int n_sys, int64_t n_items, int group,
                        int parts, int64_t* list, uint32_t* count, hipStream_t st) {
  if (n_items <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_items, dim3(grid_for(n_items, 1024)), dim3(1024), 0, st, keys, cell_start,
                     sys, item_base, n_sys, n_items, group, parts, list, count);
  return hipGetLastError();
}

template <int NQ, bool FULL, bool DIST>
static void small_disp(const SmallArgs& a, hipStream_t st) {
  if (a.disp)
    hipLaunchKernelGGL((k_small<NQ, FULL, DIST, true>), dim3(1), dim3(1024), 0, st, a);
  else
    hipLaunchKernelGGL((k_small<NQ, FULL, DIST, false>), dim3(1), dim3(1024), 0, st, a);
}
template <int NQ>
static void launch_small_nq(const SmallArgs& a, hipStream_t st) {
  if (a.full) {
    if (a.dist)
      small_disp<NQ, true, true>(a, st);
    else
      small_disp<NQ, true, false>(a, st);
  } else {
    if (a.dist)
      small_disp<NQ, false, true>(a, st);
    else
      small_disp<NQ, false, false>(a, st);
  }
}

template <int NQ, bool FULL, bool DIST>
static void brute_fill_disp(const dim3& g, const dim3& b, const SmallArgs& a, hipStream_t st) {
  if (a.disp)
    hipLaunchKernelGGL((k_brute_fill<NQ, FULL, DIST, true>), g, b, 0, st, a);
  else
    hipLaunchKernelGGL((k_brute_fill<NQ, FULL, DIST, false>), g, b, 0, st, a);
}
template <int NQ>
static void launch_brute_nq(bool fill, const SmallArgs& a, hipStream_t st) {
  const dim3 g((unsigned)((a.desc.n_atoms + 15) / 16)), b(1024);
  if (!fill) {
    if (a.full)
      hipLaunchKernelGGL((k_brute_count<NQ, true>), g, b, 0, st, a);
    else
      hipLaunchKernelGGL((k_brute_count<NQ, false>), g, b, 0, st, a);
  } else if (a.full) {
    if (a.dist)
      brute_fill_disp<NQ, true, true>(g, b, a, st);
    else
      brute_fill_disp<NQ, true, false>(g, b, a, st);
  } else {
    if (a.dist)
      brute_fill_disp<NQ, false, true>(g, b, a, st);
    else
      brute_fill_disp<NQ, false, false>(g, b, a, st);
  }
}

template <int NQ, bool FULL, bool DIST>
static void batch_fill_disp(const dim3& g, const dim3& b, const BatchSmallArgs& a,
                            hipStream_t st) {
  if (a.disp)
    hipLaunchKernelGGL((k_batch_brute_fill<NQ, FULL, DIST, true>), g, b, 0, st, a);
  else
    hipLaunchKernelGGL((k_batch_brute_fill<NQ, FULL, DIST, false>), g, b, 0, st, a);
}
template <int NQ>
static void launch_batch_brute_nq(bool fill, const BatchSmallArgs& a, int blocks,
                                  hipStream_t st) {
  const dim3 g((unsigned)blocks), b(64 * kBatchRows);
  if (!fill) {
    if (a.full)
      hipLaunchKernelGGL((k_batch_brute_count<NQ, true>), g, b, 0, st, a);
    else
      hipLaunchKernelGGL((k_batch_brute_count<NQ, false>), g, b, 0, st, a);
    hipLaunchKernelGGL(k_batch_scan, dim3(1), dim3(1024), 0, st, a);
  } else if (a.full) {
    if (a.dist)
      batch_fill_disp<NQ, true, true>(g, b, a, st);
    else
      batch_fill_disp<NQ, true, false>(g, b, a, st);
  } else {
    if (a.dist)
      batch_fill_disp<NQ, false, true>(g, b, a, st);
    else
      batch_fill_disp<NQ, false, false>(g, b, a, st);
  }
}

hipError_t launch_batch_brute(bool fill, const BatchSmallArgs& a, int blocks, hipStream_t st) {
  if (a.k < 1 || a.k > kGroup || a.n_sys < 1 || a.n < 1 || blocks < 1 || !a.sys ||
      !a.blk_start || !a.blk_sys || !a.counts || !a.offs || !a.planned || !a.flag || !a.bounds)
    return hipErrorInvalidValue;
  if (a.k == 1)
    launch_batch_brute_nq<1>(fill, a, blocks, st);
  else if (a.k <= 4)
    launch_batch_brute_nq<4>(fill, a, blocks, st);
  else
    launch_batch_brute_nq<8>(fill, a, blocks, st);
  return hipGetLastError();
}

static bool small_args_ok(const SmallArgs& a) {
  return a.k >= 1 && a.k <= kGroup && a.desc.n_atoms >= 1 && a.desc.n_atoms <= kSmallRows &&
         a.desc.atom_start == 0;
}

hipError_t launch_brute(bool fill, const SmallArgs& a, hipStream_t st) {
  if (!small_args_ok(a) || !a.counts || !a.offs || !a.planned || !a.ticket)
    return hipErrorInvalidValue;
  if (a.k == 1)
    launch_brute_nq<1>(fill, a, st);
  else if (a.k <= 4)
    launch_brute_nq<4>(fill, a, st);
  else
    launch_brute_nq<8>(fill, a, st);
  return hipGetLastError();
}

hipError_t launch_small(const SmallArgs& a, hipStream_t st) {
  // one block of 1024 threads: the plan's reduction and the scan assume it
  if (!small_args_ok(a)) return hipErrorInvalidValue;
  if (a.k == 1)
    launch_small_nq<1>(a, st);
  else if (a.k <= 4)
    launch_small_nq<4>(a, st);
  else
    launch_small_nq<8>(a, st);
  return hipGetLastError();
}

hipError_t launch_pairs(bool fill, const PairArgs& a, hipStream_t st) {
  if (a.n <= 0) return hipSuccess;
  switch (a.q_cnt) {
    case 1: launch_pairs_nq<1>(fill, a, st); break;
    case 2: launch_pairs_nq<2>(fill, a, st); break;
    case 3: launch_pairs_nq<3>(fill, a, st); break;
    case 4: launch_pairs_nq<4>(fill, a, st); break;
    default: launch_pairs_nq<8>(fill, a, st); break;
  }
  return hipGetLastError();
}

hipError_t launch_scan(const uint32_t* in, int64_t* out, int64_t len, int64_t seg, int nseg,
                       int64_t* bounds, uint64_t* status, uint32_t* ticket, hipStream_t st,
                       uint32_t* rearm) {
  if (len <= 0) {
    if (rearm) hipLaunchKernelGGL(k_rearm, dim3(1), dim3(1), 0, st, rearm);
    return hipGetLastError();
  }
  const int64_t tiles = radix_tiles(len);
  if (tiles <= kScanChainTiles) {
    hipLaunchKernelGGL(k_scan, dim3((unsigned)tiles), dim3(256), 0, st, in, out, len, seg, nseg,
                       bounds, status, ticket, rearm);
    return hipGetLastError();
  }
  // status (one u64 per tile) holds the tile sums, then (longer scans)
  // their prefixes
  hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)tiles), dim3(256), 0, st, in, len, status);
  if (tiles <= kScanRawTiles) {
    hipLaunchKernelGGL(k_scan_apply<true>, dim3((unsigned)tiles), dim3(256), 0, st, in, out, len,
                       seg, nseg, bounds, (const uint64_t*)status, rearm);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_scan_sums_excl, dim3(1), dim3(1024), 0, st, status, tiles);
  hipLaunchKernelGGL(k_scan_apply<false>, dim3((unsigned)tiles), dim3(256), 0, st, in, out, len,
                     seg, nseg, bounds, (const uint64_t*)status, rearm);
  return hipGetLastError();
}

hipError_t launch_max_disp2(const double* pos, const double* ref, int64_t n,
                            unsigned long long* out, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, (uint32_t*)out, (int64_t)2, 0u);
  if (n <= 0) return hipGetLastError();
  const int64_t blocks = std::min<int64_t>(grid_for(n, 256), 2048);
  hipLaunchKernelGGL(k_max_disp2, dim3((unsigned)blocks), dim3(256), 0, st, pos, ref, n, out);
  return hipGetLastError();
}

hipError_t launch_sort_systems(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                               uint32_t* vout, const SysDesc* sys, const int64_t* sys_start,
                               int n_sys, int passes, int32_t* cell_start, hipStream_t st) {
  if (n_sys <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_systems, dim3((unsigned)n_sys), dim3(256), 0, st, kin, vin, kout, vout,
                     sys, sys_start, passes, cell_start);
  return hipGetLastError();
}

hipError_t launch_build_systems(const BuildArgs& a, int n_sys, hipStream_t st) {
  if (n_sys <= 0) return hipSuccess;
  if (a.prm.k < 1 || a.group < 1 || a.parts < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_build_systems, dim3((unsigned)n_sys), dim3(kBuildThreads), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_sys_ranges(const int64_t* offsets, const int64_t* item_base, int n_sys, int k,
                             int64_t n, int64_t* ranges, hipStream_t st) {
  const int64_t t = (int64_t)n_sys * k;
  if (t <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sys_ranges, dim3(grid_for(t, 256)), dim3(256), 0, st, offsets, item_base,
                     n_sys, k, n, ranges);
  return hipGetLastError();
}

}  // namespace nl
