// nl_common.h -- data layout shared by the host engine and the gfx950 kernels.
//
// Every translation unit that includes this header is compiled with
// -ffp-contract=off: the FP64 expressions below are the reference's
// (Rust never fuses a*b+c), and near-cutoff decisions depend on each
// rounding step.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NL_HD __host__ __device__ inline

namespace nl {

constexpr int kGroup = 8;           // cutoffs classified per pair-kernel launch
constexpr int kWaves = 4;           // waves (units) per pair-kernel block
constexpr int kRadixMaxBits = 12;   // LSD radix digit width limit (4096 digits)
constexpr int kRadixMaxDigits = 1 << kRadixMaxBits;
constexpr int kRadixTile = 2048;    // keys per radix / scan tile (256 threads x 8)
// scans of up to this many tiles use one decoupled look-back launch; longer
// ones reduce-then-scan (three launches, no block waits on another)
constexpr int kScanChainTiles = 32;
// segmented sort of small systems: atoms per system, digit width
constexpr int kSegMax = 4096;
constexpr int kSegBits = 7;
// one-launch build (k_build_systems): atoms per system (16-bit local indices)
constexpr int kFusedMax = 12288;
// single systems of at most this many atoms on the brute-force path: one
// block plans, counts, scans and fills them in one launch (k_small)
constexpr int kSmallRows = 1024;

// Search mode chosen per system, exactly as single.rs:60-89 / batch.rs:100-124.
enum Mode : int32_t {
  kBruteIsolated = 0,  // search.rs:1187-1292: raw positions, zero shifts
  kBruteMic = 1,       // search.rs:1319-1405: minimum image per pair
  kCellList = 2,       // search.rs:92-727: all images through wrapped positions
};

// One system of a (possibly ragged) batch, 16-byte aligned for vector loads.
struct alignas(16) SysDesc {
  double h[9];      // row-major H; lattice vectors are the columns
  double hinv[9];   // row-major H^-1 (reference cofactor formula)
  double w[3];      // perpendicular widths 1/|row_k(H^-1)|
  double delta;     // FP32 filter band half-width (A), cell-list path
  double rc_max;    // largest cutoff of this system (all systems share it
                    // unless the request gives per-system cutoffs)
  float h32[9];     // H rounded to FP32 for the filter
  int64_t atom_start;
  int64_t bin_base;  // first global bin id of this system
  int32_t n_atoms;
  int32_t has_cell;  // 0: auto-box from the span of the positions
  int32_t pbc[3];
  int32_t mode;      // Mode
  int32_t nb[3];     // bins per dimension used on the GPU
  int32_t ns[3];     // stencil reach per dimension (bins)
  int32_t bin_cap;   // bins reserved for this system
  int32_t status;    // 0 ok, 1 singular cell
};

// Call-wide parameters of one launch sequence.
struct Params {
  int64_t n_atoms;     // total atoms (= pair units)
  int32_t n_sys;
  int32_t k;           // number of cutoffs
  int32_t disjoint;
  int32_t multi;       // search_*_multi rules
  int32_t bf_threshold_lo;  // brute-force threshold, split into two
  int32_t bf_threshold_hi;  // 32-bit halves (size_t semantics)
  int32_t cell_div;    // GPU bin edge = rc_max / cell_div; 0: bins_per_cutoff per system
  int32_t use_fp32;    // FP32 filter + FP64 band recheck on the cell-list path
  int32_t global_index;// add atom_start to emitted indices (batch API)
  int32_t per_sys_rc;  // rc2 holds [n_sys][k] squared cutoffs instead of [k]
  int32_t bin_order;   // 0: bins in linear order, 1: ranked by Morton code
  double rc_max;       // max cutoff (NaN-ignoring fold, single.rs:106)
};

// ---- FP64 geometry, restated from the reference (no FMA contraction) ----

// y = M x in nalgebra's gemv order: (m0*x0 + m1*x1) + m2*x2 per row.
NL_HD void matvec3(const double* m, double x0, double x1, double x2, double* y) {
  y[0] = (m[0] * x0 + m[1] * x1) + m[2] * x2;
  y[1] = (m[3] * x0 + m[4] * x1) + m[5] * x2;
  y[2] = (m[6] * x0 + m[7] * x1) + m[8] * x2;
}

// nalgebra try_inverse for 3x3 (cofactor formula), used by Cell::new
// (cell.rs:18-21).  Returns false when the determinant is exactly zero.
NL_HD bool inverse3(const double* h, double* out) {
  double m11 = h[0], m12 = h[1], m13 = h[2];
  double m21 = h[3], m22 = h[4], m23 = h[5];
  double m31 = h[6], m32 = h[7], m33 = h[8];
  double c1 = m22 * m33 - m32 * m23;
  double c2 = m21 * m33 - m31 * m23;
  double c3 = m21 * m32 - m31 * m22;
  double det = m11 * c1 - m12 * c2 + m13 * c3;
  if (det == 0.0) return false;
  out[0] = c1 / det;
  out[1] = (m13 * m32 - m33 * m12) / det;
  out[2] = (m12 * m23 - m22 * m13) / det;
  out[3] = -c2 / det;
  out[4] = (m11 * m33 - m31 * m13) / det;
  out[5] = (m13 * m21 - m23 * m11) / det;
  out[6] = c3 / det;
  out[7] = (m12 * m31 - m32 * m11) / det;
  out[8] = (m11 * m22 - m21 * m12) / det;
  return true;
}

// cell.rs:45-50
NL_HD void widths3(const double* hinv, double* w) {
  for (int k = 0; k < 3; ++k) {
    const double* r = hinv + 3 * k;
    double n2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
    w[k] = 1.0 / sqrt(n2);
  }
}

NL_HD double norm2(double x, double y, double z) { return (x * x + y * y) + z * z; }

// Bins per cutoff length chosen per system (cell_div 0): the largest of 1..4
// whose bins still hold kBinAtoms atoms at the system's density n / volume.
// Finer bins test fewer candidates per pair but add segments and shrink the
// home groups; on gfx950 the finer grid pays once its bins hold about six
// atoms (measured on periodic systems of 20,000 atoms at densities 0.03 and
// 0.1 per cubic angstrom, cutoffs 3 to 14 angstrom).  Host and device use
// the same rule; the grid never changes a result.
constexpr double kBinAtoms = 6.0;
NL_HD int bins_per_cutoff(double rc, double n_atoms, double volume) {
  const double q = n_atoms * rc * rc * rc;  // atoms per bin at cd = 1, times volume
  for (int cd = 4; cd > 1; --cd)
    if (q >= kBinAtoms * (double)(cd * cd * cd) * volume) return cd;
  return 1;  // also for a NaN or zero volume
}
// Rank of bin (x, y, z) of an nx x ny x nz grid in the Morton order of its
// coordinates (x in the lowest bit: interleave(x) | interleave(y) << 1 |
// interleave(z) << 2, as the reference ranks its bins, search.rs:126-146 and
// :1177-1185), counted without sorting: going down the code bits, each set
// bit adds the grid cells that share the higher bits and have a 0 there
// (the per-axis counts of the sub-box multiply).
NL_HD int32_t morton_rank(int32_t x, int32_t y, int32_t z, int32_t nx, int32_t ny, int32_t nz) {
  const int32_t v[3] = {x, y, z}, n[3] = {nx, ny, nz};
  int B = 0;
  for (int d = 0; d < 3; ++d) {
    int b = 0;
    while (b < 31 && ((int64_t)1 << b) < (int64_t)n[d]) ++b;  // bits of n - 1
    B = b > B ? b : B;
  }
  int64_t q[3] = {0, 0, 0};  // bits of each coordinate taken so far
  int f[3] = {B, B, B};      // bits of each coordinate still free
  int64_t r = 0;
  for (int l = B - 1; l >= 0; --l) {
    for (int d = 2; d >= 0; --d) {  // code bits 3l+2 (z), 3l+1 (y), 3l (x)
      const int64_t bit = (v[d] >> l) & 1;
      if (bit) {
        int64_t prod = 1;
        for (int e = 0; e < 3; ++e) {
          const int fe = e == d ? f[e] - 1 : f[e];
          const int64_t lo = e == d ? (q[d] << f[d]) : (q[e] << f[e]);
          int64_t c = (int64_t)n[e] - lo;
          c = c < 0 ? 0 : c;
          const int64_t span = (int64_t)1 << fe;
          prod *= c < span ? c : span;
        }
        r += prod;
      }
      q[d] = (q[d] << 1) | bit;
      f[d] -= 1;
    }
  }
  return (int32_t)r;
}

NL_HD double det3(const double* h) {
  return h[0] * (h[4] * h[8] - h[7] * h[5]) - h[1] * (h[3] * h[8] - h[6] * h[5]) +
         h[2] * (h[3] * h[7] - h[6] * h[4]);
}

// Rust `as i32` (saturating, truncating toward zero, NaN -> 0).
NL_HD int32_t as_i32(double x) {
  if (x != x) return 0;
  if (x >= 2147483647.0) return 2147483647;
  if (x <= -2147483648.0) return (-2147483647 - 1);
  return (int32_t)x;
}

// Euclidean division (search.rs:1162-1166).
NL_HD void div_mod(int32_t val, int32_t m, int32_t* rem, int32_t* q) {
  int32_t qq = val / m, r = val % m;
  if (r < 0) {
    r += m;
    qq -= 1;
  }
  *rem = r;
  *q = qq;
}

}  // namespace nl
