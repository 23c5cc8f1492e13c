// nl_kernels.h -- launch wrappers for the gfx950 kernels in nl_kernels.hip.
#pragma once
#include "nl_common.h"

namespace nl {

struct alignas(16) Pos64 {  // sorted wrapped position + local original index
  double x, y, z, w;
};

// Device pointers handed to the pair kernels.
struct PairArgs {
  const double* pos;         // raw positions (N,3), f64
  const SysDesc* sys;
  const int64_t* sys_start;  // n_sys + 1
  const int64_t* item_base;  // n_sys + 1: work items of system s (atoms, then bins)
  const uint32_t* keys;      // sorted bin keys
  const int32_t* bin_rank;   // Morton order: [bin_base + linear bin] -> rank (null: linear)
  const int32_t* rank_bin;   // Morton order: [bin_base + rank] -> linear bin
  const int32_t* cell_start; // total_bins + 1
  const Pos64* spos64;       // sorted wrapped positions (FP64), w = local index
  const float4* spos32;      // sorted wrapped positions (FP32), w = local index bits
  const int4* sidx;          // sorted wrap shifts s_a, w = local index
  const double* rc2;         // k squared cutoffs, ascending ([n_sys][k] if per_sys_rc)
  const float2* band;        // [n_sys][k] FP32 accept/reject thresholds
  uint32_t* counts;          // [k][n] (count pass; zero for items without work)
  const int64_t* list;       // items with work, one per wave (any order)
  const uint32_t* n_real;    // [2] list lengths from the front and from the back (device)
  const int64_t* offsets;    // [k][n] + 1 (fill pass)
  const int64_t* bounds;     // [k + 1] first edge of each list (sorted-cutoff order)
  int64_t* edges;            // list q: sources at base_q + e, targets at base_q + M_q + e
  int32_t* shifts;           // list q: 3 per edge at 3 e
  double* dist;              // optional per edge: d^2 from the fill pass, then |r_ij|
  double* disp;              // optional per edge: r_ij = r_j + H S - r_i, 3 doubles
  int64_t n;                 // work items
  int64_t max_items;         // bound on the items with work (the launch grid)
  int64_t capacity;          // edges the output buffers hold (fill pass)
  int32_t n_sys, k, q_lo, q_cnt, disjoint, use_fp32, global_index;
  int32_t per_sys_rc;        // rc2 is [n_sys][k] (per-system cutoffs)
  int32_t full;              // full lists: every pair in both directions
  int32_t group;             // home atoms per cell-list work item (<= 64)
  int32_t parts;             // slices of the half stencil per home group
  int32_t xcd_chunk;         // XCD run length: count items (fill: 4x), 0: dispatch order
};

// One single system of at most kSmallRows atoms for k_small: its descriptor
// before planning (atom_start 0), ascending squared cutoffs and the outputs.
struct SmallArgs {
  const double* pos;    // raw positions (n, 3), f64
  SysDesc desc;
  double rc2[kGroup];   // k squared cutoffs, ascending
  Params prm;
  float2* band;         // [k], written by the plan (cell-list path only)
  int64_t* bounds;      // [k + 1]: first edge of each list, total; bounds[k] = -1
                        // when the plan takes the cell-list path (nothing written)
  int64_t* edges;       // outputs laid out as for the pair kernels
  int32_t* shifts;
  double* dist;         // optional: |r_ij|
  double* disp;         // optional: r_ij
  int64_t capacity;     // edges the outputs hold; a larger total writes only bounds
  int32_t k, disjoint, full;
  // two-launch form (launch_brute) only
  uint32_t* counts;     // [k][n] per-row counts
  uint32_t* offs;       // [k][n] per-row offsets over all lists
  SysDesc* planned;     // the descriptor after planning (count -> fill)
  uint32_t* ticket;     // zero before the count launch; the last block re-arms it
};

// A batch of small systems that all take a brute-force path (each at most
// kSmallRows atoms): ceil(n_s / kBatchRows) blocks of kBatchRows rows (one
// wave each) per system.
constexpr int kBatchRows = 4;
struct BatchSmallArgs {
  const double* pos;          // raw positions of the batch (n, 3), f64
  const SysDesc* sys;         // [n_sys] descriptors before planning
  const int32_t* blk_start;   // [n_sys + 1] first block of each system
  const int32_t* blk_sys;     // [blocks] system of each block
  int32_t n_sys;
  int64_t n;                  // atoms in the batch
  double rc2[kGroup];         // k squared cutoffs, ascending
  Params prm;
  float2* band;               // [n_sys][k], written by the plan (cell-list path only)
  int64_t* bounds;            // [k + 1]; bounds[k] = -1 when some system takes the cell list
  int64_t* edges;             // outputs laid out as for the pair kernels
  int32_t* shifts;
  double* dist;
  double* disp;
  int64_t capacity;           // edges the outputs hold; a larger total writes only bounds
  int32_t k, disjoint, full;
  uint32_t* counts;           // [k][n] per-row counts
  uint32_t* offs;             // [k][n] per-row offsets over all lists (total < 2^32)
  SysDesc* planned;           // [n_sys] descriptors after planning (count -> fill)
  uint32_t* flag;             // some system took the cell-list path: zero before the
                              // count launch, re-armed by the scan
};
// count and scan launches (fill = false; writes bounds), then the fill
// launch (fill = true)
hipError_t launch_batch_brute(bool fill, const BatchSmallArgs& a, int blocks, hipStream_t st);

// The build of a batch whose systems all have at most kFusedMax atoms in one
// launch (one block per system): plan, bins, sort, bin offsets, sorted
// positions and wrap shifts, work list and zero counts of item slots without
// work.  n_real must be zero before the launch.
struct BuildArgs {
  const double* pos;         // raw positions (n, 3), f64
  SysDesc* sys;              // [n_sys]: descriptors in, plans out
  float2* band;              // [n_sys][k] FP32 filter thresholds (written)
  const double* rc2;         // [k], or [n_sys][k] with per-system cutoffs
  Params prm;
  const int64_t* item_base;  // [n_sys + 1]
  int64_t n_items;
  int32_t group, parts;
  uint32_t* keys;            // sorted bin keys (bin_base + bin)
  uint32_t* vals;            // sorted atoms (global index)
  int32_t* cell_start;       // [total bins + 1]
  Pos64* spos64;
  float4* spos32;
  int4* sidx;
  int64_t* list;             // work list: group items from the front, rows from the back
  uint32_t* n_real;          // [2] lengths of both ends (zero before the launch)
  uint32_t* counts;          // [k][n_items]
  const SysDesc* sys_in;     // descriptors to plan from (null: sys itself)
  uint32_t* zero;            // words cleared by the launch (n_zero of them; not n_real)
  int64_t n_zero;
};
hipError_t launch_build_systems(const BuildArgs& a, int n_sys, hipStream_t st);

// one block (k_small)
hipError_t launch_small(const SmallArgs& a, hipStream_t st);
// count launch (fill = false; writes bounds), then fill launch (fill = true)
hipError_t launch_brute(bool fill, const SmallArgs& a, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t st);
// clears a[0, na) and b[0, nb), and copies src[0, nc) to dst (u32 words)
hipError_t launch_zero2(uint32_t* a, int64_t na, uint32_t* b, int64_t nb, hipStream_t st,
                        uint32_t* dst = nullptr, const uint32_t* src = nullptr, int64_t nc = 0);
hipError_t launch_sqrt_dist(double* dist, const int64_t* total, int64_t capacity,
                            hipStream_t st);
hipError_t launch_plan(const double* pos, SysDesc* sys, float2* band, const double* rc2,
                       const Params& p, hipStream_t st);
hipError_t launch_bin(const double* pos, const SysDesc* sys, const int64_t* sys_start, int n_sys,
                      int64_t n, const int32_t* bin_rank, uint32_t* keys, uint32_t* vals,
                      hipStream_t st);
// Morton ranks of the bins of every cell-list system: bin_rank[bin_base +
// linear] = rank, rank_bin[bin_base + rank] = linear
hipError_t launch_morton(const SysDesc* sys, int n_sys, int64_t total_bins, int32_t* bin_rank,
                         int32_t* rank_bin, hipStream_t st);
// one stable pass over digit bits [shift, shift + bits), bits <= 12:
// cnt [2^bits * tiles] u32, off [2^bits * tiles + 1] i64, bounds [2] i64 and
// a zeroed scan status [radix_tiles(2^bits * tiles)] + ticket
hipError_t launch_radix_pass(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                             uint32_t* vout, int64_t n, int shift, int bits, uint32_t* cnt,
                             int64_t* off, int64_t* bounds, uint64_t* status, uint32_t* ticket,
                             hipStream_t st);
hipError_t launch_finalize(const uint32_t* keys, const uint32_t* vals, const double* pos,
                           const SysDesc* sys, const int64_t* sys_start, int n_sys, int64_t n,
                           int64_t total_bins, int32_t* cell_start, Pos64* spos64,
                           float4* spos32, int4* sidx, bool with_cells, hipStream_t st);
hipError_t launch_items(const uint32_t* keys, const int32_t* cell_start, const SysDesc* sys,
                        const int64_t* item_base, int n_sys, int64_t n_items, int group,
                        int parts, int64_t* list, uint32_t* count, hipStream_t st);
hipError_t launch_pairs(bool fill, const PairArgs& a, hipStream_t st);
// rearm (or null): the work list's lengths [2], handed on to rearm[2..3] for
// the fill pass and cleared for the next build
hipError_t launch_scan(const uint32_t* in, int64_t* out, int64_t len, int64_t seg, int nseg,
                       int64_t* bounds, uint64_t* status, uint32_t* ticket, hipStream_t st,
                       uint32_t* rearm = nullptr);
// *out = bits of max_a |pos_a - ref_a|^2 (cleared first, in the same stream)
hipError_t launch_max_disp2(const double* pos, const double* ref, int64_t n,
                            unsigned long long* out, hipStream_t st);
// stable sort of each system's keys (<= kSegMax atoms each) in one launch
hipError_t launch_sort_systems(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                               uint32_t* vout, const SysDesc* sys, const int64_t* sys_start,
                               int n_sys, int passes, int32_t* cell_start, hipStream_t st);
hipError_t launch_sys_ranges(const int64_t* offsets, const int64_t* item_base, int n_sys, int k,
                             int64_t n, int64_t* ranges, hipStream_t st);

inline int64_t radix_tiles(int64_t n) { return (n + kRadixTile - 1) / kRadixTile; }

}  // namespace nl
