/*
 * nl_oracle.c -- CPU oracle for radius-cutoff neighbour-list construction.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * algorithm of neighborlist-rs 0.1.9 (src/cell.rs, src/search.rs,
 * src/single.rs, src/batch.rs, src/config.rs).  It is the checker that the
 * HIP engine is compared against, and (built with OpenMP) the CPU baseline
 * timed next to the GPU.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -shared -fPIC (oracle/build_oracle.py).
 * -ffp-contract=off is required: the reference is Rust, which never contracts
 * a*b+c into an FMA, and every near-cutoff decision depends on that.
 *
 * Arithmetic-order conventions restated here (each pinned as noted in the
 * tests/ suite; see the pin table at the bottom of this header):
 *   - 3x3 matrix . vector: y_r = (m_r0*x0 + m_r1*x1) + m_r2*x2
 *     (nalgebra 0.34.1 gemv column-axpy order; used by cell.rs:23-29, 79, 90
 *     and search.rs:103, 154-160, 194, 203, 1113).
 *   - 3x3 inverse: nalgebra's explicit cofactor formula (try_inverse for
 *     dimension 3), determinant (m11*c1 - m12*c2) + m13*c3, exact-zero test.
 *   - |v|^2 = (x*x + y*y) + z*z (search.rs:543, 1285; nalgebra norm_squared).
 *   - rust f64::round = round half away from zero (cell.rs:81-83).
 *   - `as i32` / `as usize` casts saturate; float->int truncates toward 0.
 *
 * Parity pins (checks that do not merely retype these formulas): the
 * reference's unit-test known answers (cell.rs:102-239, search.rs:1413-1910,
 * single.rs:177-284, batch.rs:297-564, tests/ *.py), the pair counts the
 * reference logged for its own fixture files (benchmarks/benchmark_EPYC7R13.txt
 * :7-98), and brute-force/multi-image invariants on random inputs.  The ulp
 * level of H^-1 (nalgebra formula, absent here) is "parity unpinned": it is
 * only checked through those set-level answers.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NLO_OK 0
#define NLO_ERR 1

/* ---- config.rs:4-39 ----------------------------------------------------- */
#define NLO_MAX_STACK_SIZE 1024
static size_t g_bf_threshold = 1000;   /* config.rs:8  */
static size_t g_par_threshold = 300;   /* config.rs:9  */
static size_t g_stack_threshold = 1000;/* config.rs:10 */

size_t nlo_get_brute_force_threshold(void) { return g_bf_threshold; }
void nlo_set_brute_force_threshold(size_t v) { g_bf_threshold = v; }
size_t nlo_get_parallel_threshold(void) { return g_par_threshold; }
void nlo_set_parallel_threshold(size_t v) { g_par_threshold = v; }
size_t nlo_get_stack_threshold(void) { return g_stack_threshold; }
void nlo_set_stack_threshold(size_t v) {
  g_stack_threshold = v < NLO_MAX_STACK_SIZE ? v : NLO_MAX_STACK_SIZE; /* config.rs:36-38 */
}

static _Thread_local char g_err[256];
const char *nlo_last_error(void) { return g_err; }
static int set_err(const char *msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return NLO_ERR;
}

/* ---- results ------------------------------------------------------------- */
typedef struct {
  int64_t n;
  int64_t *ei;
  int64_t *ej;
  int32_t *s; /* 3 per edge */
} nlo_list;

typedef struct {
  int64_t k;
  nlo_list *lists;
} nlo_result;

void nlo_result_free(nlo_result *r) {
  if (!r || !r->lists) return;
  for (int64_t k = 0; k < r->k; ++k) {
    free(r->lists[k].ei);
    free(r->lists[k].ej);
    free(r->lists[k].s);
  }
  free(r->lists);
  r->lists = NULL;
  r->k = 0;
}

static void list_alloc(nlo_list *l, int64_t n) {
  l->n = n;
  size_t m = (size_t)(n > 0 ? n : 1);
  l->ei = (int64_t *)malloc(m * sizeof(int64_t));
  l->ej = (int64_t *)malloc(m * sizeof(int64_t));
  l->s = (int32_t *)malloc(3 * m * sizeof(int32_t));
}

/* growable list for the serial/brute-force paths */
typedef struct {
  int64_t n, cap;
  int64_t *ei, *ej;
  int32_t *s;
} vlist;

static void vl_init(vlist *v, int64_t cap) {
  v->n = 0;
  v->cap = cap > 16 ? cap : 16;
  v->ei = (int64_t *)malloc(v->cap * sizeof(int64_t));
  v->ej = (int64_t *)malloc(v->cap * sizeof(int64_t));
  v->s = (int32_t *)malloc(3 * v->cap * sizeof(int32_t));
}
static void vl_push(vlist *v, int64_t i, int64_t j, int32_t sx, int32_t sy, int32_t sz) {
  if (v->n == v->cap) {
    v->cap *= 2;
    v->ei = (int64_t *)realloc(v->ei, v->cap * sizeof(int64_t));
    v->ej = (int64_t *)realloc(v->ej, v->cap * sizeof(int64_t));
    v->s = (int32_t *)realloc(v->s, 3 * v->cap * sizeof(int32_t));
  }
  v->ei[v->n] = i;
  v->ej[v->n] = j;
  v->s[3 * v->n] = sx;
  v->s[3 * v->n + 1] = sy;
  v->s[3 * v->n + 2] = sz;
  v->n++;
}
static nlo_list vl_take(vlist *v) {
  nlo_list l = {v->n, v->ei, v->ej, v->s};
  return l;
}

/* ---- saturating casts (Rust `as`) ---------------------------------------- */
static int32_t as_i32(double x) {
  if (x != x) return 0;
  if (x >= 2147483647.0) return INT32_MAX;
  if (x <= -2147483648.0) return INT32_MIN;
  return (int32_t)x; /* truncation toward zero */
}
static uint64_t as_usize(double x) {
  if (x != x || x <= 0.0) return 0;
  if (x >= 18446744073709551615.0) return UINT64_MAX;
  return (uint64_t)x;
}
static uint64_t as_u64(double x) { return as_usize(x); }

/* rust f64::round: half away from zero */
static double rust_round(double x) { return round(x); }

/* ---- cell.rs ------------------------------------------------------------- */
typedef struct {
  double h[9];    /* row-major H[r][c]; lattice vectors are the columns */
  double hinv[9]; /* row-major H^-1 */
  int pbc[3];
} cell_t;

/* y = M x, nalgebra gemv order */
static inline void matvec(const double *m, const double *x, double *y) {
  y[0] = (m[0] * x[0] + m[1] * x[1]) + m[2] * x[2];
  y[1] = (m[3] * x[0] + m[4] * x[1]) + m[5] * x[2];
  y[2] = (m[6] * x[0] + m[7] * x[1]) + m[8] * x[2];
}

/* cell.rs:18-21 -> nalgebra try_inverse, 3x3 cofactor formula */
int nlo_cell_inverse(const double *h, double *out) {
  double m11 = h[0], m12 = h[1], m13 = h[2];
  double m21 = h[3], m22 = h[4], m23 = h[5];
  double m31 = h[6], m32 = h[7], m33 = h[8];
  double minor_m12_m23 = m22 * m33 - m32 * m23;
  double minor_m11_m23 = m21 * m33 - m31 * m23;
  double minor_m11_m22 = m21 * m32 - m31 * m22;
  double det = m11 * minor_m12_m23 - m12 * minor_m11_m23 + m13 * minor_m11_m22;
  if (det == 0.0) return NLO_ERR;
  out[0] = minor_m12_m23 / det;
  out[1] = (m13 * m32 - m33 * m12) / det;
  out[2] = (m12 * m23 - m22 * m13) / det;
  out[3] = -minor_m11_m23 / det;
  out[4] = (m11 * m33 - m31 * m13) / det;
  out[5] = (m13 * m21 - m23 * m11) / det;
  out[6] = minor_m11_m22 / det;
  out[7] = (m12 * m31 - m32 * m11) / det;
  out[8] = (m11 * m22 - m21 * m12) / det;
  return NLO_OK;
}

static int cell_new(const double *h, const int *pbc, cell_t *c) {
  memcpy(c->h, h, 9 * sizeof(double));
  for (int k = 0; k < 3; ++k) c->pbc[k] = pbc ? (pbc[k] != 0) : 1;
  if (nlo_cell_inverse(h, c->hinv) != NLO_OK)
    return set_err("Cell matrix is not invertible"); /* cell.rs:6 */
  return NLO_OK;
}

/* cell.rs:45-50: w_k = 1/|row_k(H^-1)| */
void nlo_perpendicular_widths(const double *hinv, double *w) {
  for (int k = 0; k < 3; ++k) {
    const double *r = hinv + 3 * k;
    double n2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
    w[k] = 1.0 / sqrt(n2);
  }
}

/* cell.rs:52-72 */
int nlo_cell_wrap(const double *h, const int *pbc, const double *pos, double *out) {
  cell_t c;
  if (cell_new(h, pbc, &c) != NLO_OK) return NLO_ERR;
  double f[3];
  matvec(c.hinv, pos, f);
  for (int k = 0; k < 3; ++k)
    if (c.pbc[k]) f[k] = f[k] - floor(f[k]);
  matvec(c.h, f, out);
  return NLO_OK;
}

/* cell.rs:74-94 */
static inline void shift_and_disp(const cell_t *c, const double *ri, const double *rj,
                                  int32_t *shift, double *disp) {
  double d[3] = {rj[0] - ri[0], rj[1] - ri[1], rj[2] - ri[2]};
  double df[3];
  matvec(c->hinv, d, df);
  double sf[3];
  for (int k = 0; k < 3; ++k) sf[k] = c->pbc[k] ? -rust_round(df[k]) : 0.0;
  for (int k = 0; k < 3; ++k) shift[k] = as_i32(sf[k]);
  double v[3];
  matvec(c->h, sf, v);
  double img[3] = {rj[0] + v[0], rj[1] + v[1], rj[2] + v[2]};
  for (int k = 0; k < 3; ++k) disp[k] = img[k] - ri[k];
}

int nlo_cell_mic(const double *h, const int *pbc, const double *ri, const double *rj,
                 int32_t *shift, double *disp) {
  cell_t c;
  if (cell_new(h, pbc, &c) != NLO_OK) return NLO_ERR;
  shift_and_disp(&c, ri, rj, shift, disp);
  return NLO_OK;
}

static inline double norm2(const double *d) { return (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]; }

/* ---- search.rs: Morton ---------------------------------------------------- */
/* search.rs:1177-1185 */
static uint64_t interleave_3(uint64_t x) {
  x &= 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffULL;
  x = (x | x << 16) & 0x1f0000ff0000ffULL;
  x = (x | x << 8) & 0x100f00f00f00f00fULL;
  x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
  x = (x | x << 2) & 0x1249249249249249ULL;
  return x;
}
static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
/* search.rs:1169-1175 */
uint64_t nlo_z_order(const double *f) {
  const double scale = (double)(1ULL << 21);
  uint64_t x = as_u64(clampd(f[0], 0.0, 0.999999) * scale);
  uint64_t y = as_u64(clampd(f[1], 0.0, 0.999999) * scale);
  uint64_t z = as_u64(clampd(f[2], 0.0, 0.999999) * scale);
  return interleave_3(x) | (interleave_3(y) << 1) | (interleave_3(z) << 2);
}

/* search.rs:1162-1166 rem_euclid/div_euclid */
static inline void div_mod(int32_t val, int32_t max, int32_t *rem, int32_t *shift) {
  int32_t q = val / max, r = val % max;
  if (r < 0) {
    r += max;
    q -= 1;
  }
  *rem = r;
  *shift = q;
}

/* ---- CellList (search.rs:68-264) ----------------------------------------- */
typedef struct {
  int64_t n;
  int64_t *particles;   /* sorted -> original */
  int64_t *cell_starts; /* by bin rank, total_bins+1 */
  int64_t *bin_ranks;   /* linear bin -> rank */
  double *px, *py, *pz;
  int32_t *sx, *sy, *sz;
  int64_t nb[3];
  int32_t ns[3];
  double basis[3][3];
  double origin[3];
  double radius;
} celllist_t;

typedef struct {
  uint64_t key;
  int64_t idx;
} kv_t;
static int cmp_kv(const void *a, const void *b) {
  const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

static void cl_free(celllist_t *cl) {
  free(cl->particles);
  free(cl->cell_starts);
  free(cl->bin_ranks);
  free(cl->px);
  free(cl->py);
  free(cl->pz);
  free(cl->sx);
  free(cl->sy);
  free(cl->sz);
}

static void bin_of(const cell_t *c, const int64_t *nb, const double *pos, int64_t *lin) {
  double f[3];
  matvec(c->hinv, pos, f);
  int64_t b[3];
  for (int k = 0; k < 3; ++k) {
    double u = f[k] - floor(f[k]);
    uint64_t bk = as_usize(u * (double)nb[k]);
    b[k] = (int64_t)(bk < (uint64_t)(nb[k] - 1) ? bk : (uint64_t)(nb[k] - 1));
  }
  *lin = b[0] + nb[0] * (b[1] + nb[1] * b[2]);
}

static void cl_build(const cell_t *c, const double *pos, int64_t n, double cutoff, int par,
                     celllist_t *cl) {
  memset(cl, 0, sizeof *cl);
  cl->n = n;
  /* 1. Morton codes (search.rs:96-110) */
  kv_t *ad = (kv_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(kv_t));
#pragma omp parallel for if (par)
  for (int64_t i = 0; i < n; ++i) {
    double f[3];
    matvec(c->hinv, pos + 3 * i, f);
    for (int k = 0; k < 3; ++k) f[k] -= floor(f[k]);
    ad[i].key = nlo_z_order(f);
    ad[i].idx = i;
  }
  /* 2. sort (search.rs:112-116; sort_unstable tie order is irrelevant to the set) */
  qsort(ad, (size_t)n, sizeof(kv_t), cmp_kv);
  /* 3. bins (search.rs:118-124) */
  double w[3];
  nlo_perpendicular_widths(c->hinv, w);
  for (int k = 0; k < 3; ++k) {
    uint64_t nk = as_usize(floor(w[k] / cutoff));
    cl->nb[k] = (int64_t)(nk > 1 ? nk : 1);
  }
  int64_t total = cl->nb[0] * cl->nb[1] * cl->nb[2];
  /* bin Morton ranks (search.rs:126-146) */
  cl->bin_ranks = (int64_t *)malloc((size_t)total * sizeof(int64_t));
  kv_t *bm = (kv_t *)malloc((size_t)total * sizeof(kv_t));
  for (int64_t i = 0; i < total; ++i) {
    uint64_t bx = (uint64_t)(i % cl->nb[0]);
    uint64_t by = (uint64_t)((i / cl->nb[0]) % cl->nb[1]);
    uint64_t bz = (uint64_t)(i / (cl->nb[0] * cl->nb[1]));
    bm[i].key = interleave_3(bx) | (interleave_3(by) << 1) | (interleave_3(bz) << 2);
    bm[i].idx = i;
  }
  qsort(bm, (size_t)total, sizeof(kv_t), cmp_kv);
  for (int64_t r = 0; r < total; ++r) cl->bin_ranks[bm[r].idx] = r;
  free(bm);
  /* n_search (search.rs:148-152) */
  for (int k = 0; k < 3; ++k) cl->ns[k] = as_i32(ceil(cutoff * (double)cl->nb[k] / w[k]));
  /* 4. counts + prefix (search.rs:156-178) */
  int64_t *counts = (int64_t *)calloc((size_t)total, sizeof(int64_t));
  for (int64_t a = 0; a < n; ++a) {
    int64_t lin;
    bin_of(c, cl->nb, pos + 3 * ad[a].idx, &lin);
    counts[cl->bin_ranks[lin]]++;
  }
  cl->cell_starts = (int64_t *)malloc((size_t)(total + 1) * sizeof(int64_t));
  int64_t acc = 0;
  for (int64_t i = 0; i < total; ++i) {
    cl->cell_starts[i] = acc;
    acc += counts[i];
  }
  cl->cell_starts[total] = acc;
  /* 5. fill (search.rs:180-227) */
  size_t m = (size_t)(n > 0 ? n : 1);
  cl->px = (double *)malloc(m * sizeof(double));
  cl->py = (double *)malloc(m * sizeof(double));
  cl->pz = (double *)malloc(m * sizeof(double));
  cl->sx = (int32_t *)malloc(m * sizeof(int32_t));
  cl->sy = (int32_t *)malloc(m * sizeof(int32_t));
  cl->sz = (int32_t *)malloc(m * sizeof(int32_t));
  cl->particles = (int64_t *)malloc(m * sizeof(int64_t));
  int64_t *fillp = counts; /* reuse as cursor */
  for (int64_t i = 0; i < total; ++i) fillp[i] = cl->cell_starts[i];
  for (int64_t a = 0; a < n; ++a) {
    int64_t oi = ad[a].idx;
    const double *p = pos + 3 * oi;
    double f[3];
    matvec(c->hinv, p, f);
    double fl[3] = {floor(f[0]), floor(f[1]), floor(f[2])};
    int32_t s[3] = {as_i32(-fl[0]), as_i32(-fl[1]), as_i32(-fl[2])};
    double ms[3] = {(double)(-s[0]), (double)(-s[1]), (double)(-s[2])};
    double sv[3];
    matvec(c->h, ms, sv);
    double wp[3] = {p[0] - sv[0], p[1] - sv[1], p[2] - sv[2]};
    int64_t b[3];
    for (int k = 0; k < 3; ++k) {
      double u = f[k] - fl[k];
      uint64_t bk = as_usize(u * (double)cl->nb[k]);
      b[k] = (int64_t)(bk < (uint64_t)(cl->nb[k] - 1) ? bk : (uint64_t)(cl->nb[k] - 1));
    }
    int64_t lin = b[0] + cl->nb[0] * (b[1] + cl->nb[1] * b[2]);
    int64_t loc = fillp[cl->bin_ranks[lin]]++;
    cl->px[loc] = wp[0];
    cl->py[loc] = wp[1];
    cl->pz[loc] = wp[2];
    cl->sx[loc] = s[0];
    cl->sy[loc] = s[1];
    cl->sz[loc] = s[2];
    cl->particles[loc] = oi;
  }
  free(counts);
  free(ad);
  /* 6. prune geometry (search.rs:229-246) */
  double e0[3] = {1.0 / (double)cl->nb[0], 0.0, 0.0};
  double e1[3] = {0.0, 1.0 / (double)cl->nb[1], 0.0};
  double e2[3] = {0.0, 0.0, 1.0 / (double)cl->nb[2]};
  double eo[3] = {0.5 / (double)cl->nb[0], 0.5 / (double)cl->nb[1], 0.5 / (double)cl->nb[2]};
  matvec(c->h, e0, cl->basis[0]);
  matvec(c->h, e1, cl->basis[1]);
  matvec(c->h, e2, cl->basis[2]);
  matvec(c->h, eo, cl->origin);
  const double *vx = cl->basis[0], *vy = cl->basis[1], *vz = cl->basis[2];
  double t[3], dd[4];
  for (int q = 0; q < 4; ++q) {
    double sy_ = (q & 2) ? -1.0 : 1.0, sz_ = (q & 1) ? -1.0 : 1.0;
    for (int k = 0; k < 3; ++k) {
      double a = vx[k] + (sy_ > 0 ? vy[k] : -vy[k]);
      t[k] = sz_ > 0 ? a + vz[k] : a - vz[k];
    }
    dd[q] = norm2(t);
  }
  double mx = fmax(fmax(fmax(dd[0], dd[1]), dd[2]), dd[3]);
  cl->radius = 0.5 * sqrt(mx);
}

/* One atom's stencil traversal, shared by count, fill and the serial search
 * (search.rs:448-727, 729-1078, 1080-1159: identical traversal order and
 * identical arithmetic; the pair rule and d^2 expression are at :540-549).
 * sorted_cutoffs_sq ascending; max_sq = last.  Emits via callback-free
 * switches: if `counts` is non-NULL, it counts per list; else writes.      */
typedef struct {
  const cell_t *c;
  const celllist_t *cl;
  const double *rc2; /* sorted ascending */
  int k;
  int disjoint;
  int64_t idx_offset; /* batch offset added to emitted indices */
} trav_t;

static void atom_traverse(const trav_t *t, int64_t i, int64_t *counts, nlo_list *out,
                          const int64_t *offs, vlist *vout) {
  const celllist_t *cl = t->cl;
  const cell_t *c = t->c;
  double pi[3] = {cl->px[i], cl->py[i], cl->pz[i]};
  int32_t si[3] = {cl->sx[i], cl->sy[i], cl->sz[i]};
  double fi[3];
  matvec(c->hinv, pi, fi);
  int32_t b[3];
  for (int k = 0; k < 3; ++k) {
    int32_t v = as_i32(fi[k] * (double)cl->nb[k]);
    int32_t mx = (int32_t)cl->nb[k] - 1;
    b[k] = v < mx ? v : mx;
  }
  int64_t io = cl->particles[i];
  double max_sq = t->rc2[t->k - 1];
  double plim = sqrt(max_sq) + cl->radius;
  double plim2 = plim * plim;
  int64_t local[t->k]; /* one counter per list (any number of cutoffs) */
  for (int q = 0; q < t->k; ++q) local[q] = 0;
  for (int32_t dx = -cl->ns[0]; dx <= cl->ns[0]; ++dx) {
    int32_t nbx, sxb;
    div_mod(b[0] + dx, (int32_t)cl->nb[0], &nbx, &sxb);
    double ox[3] = {c->h[0] * (double)sxb, c->h[3] * (double)sxb, c->h[6] * (double)sxb};
    double cx[3];
    for (int k = 0; k < 3; ++k) cx[k] = cl->origin[k] + cl->basis[0][k] * (double)nbx;
    for (int32_t dy = -cl->ns[1]; dy <= cl->ns[1]; ++dy) {
      int32_t nby, syb;
      div_mod(b[1] + dy, (int32_t)cl->nb[1], &nby, &syb);
      double oy[3] = {c->h[1] * (double)syb, c->h[4] * (double)syb, c->h[7] * (double)syb};
      double cxy[3];
      for (int k = 0; k < 3; ++k) cxy[k] = cx[k] + cl->basis[1][k] * (double)nby;
      for (int32_t dz = -cl->ns[2]; dz <= cl->ns[2]; ++dz) {
        int32_t nbz, szb;
        div_mod(b[2] + dz, (int32_t)cl->nb[2], &nbz, &szb);
        double oz[3] = {c->h[2] * (double)szb, c->h[5] * (double)szb, c->h[8] * (double)szb};
        double tc[3], off[3], diff[3];
        for (int k = 0; k < 3; ++k) {
          tc[k] = cxy[k] + cl->basis[2][k] * (double)nbz;
          off[k] = (ox[k] + oy[k]) + oz[k];
          diff[k] = (tc[k] + off[k]) - pi[k];
        }
        if (norm2(diff) > plim2) continue; /* search.rs:500-505 */
        int64_t lin = nbx + cl->nb[0] * (nby + cl->nb[1] * nbz);
        int64_t rank = cl->bin_ranks[lin];
        int64_t s0 = cl->cell_starts[rank], s1 = cl->cell_starts[rank + 1];
        int shifted = (sxb != 0 || syb != 0 || szb != 0);
        for (int64_t j = s0; j < s1; ++j) {
          int64_t jo = cl->particles[j];
          if (!(io < jo || (io == jo && shifted))) continue;
          double d[3] = {(cl->px[j] - pi[0]) + off[0], (cl->py[j] - pi[1]) + off[1],
                         (cl->pz[j] - pi[2]) + off[2]};
          double d2 = norm2(d);
          if (!(d2 < max_sq)) continue;
          for (int q = 0; q < t->k; ++q) {
            if (d2 < t->rc2[q]) {
              if (counts) {
                counts[q]++;
              } else {
                int32_t S[3] = {cl->sx[j] - si[0] + sxb, cl->sy[j] - si[1] + syb,
                                cl->sz[j] - si[2] + szb};
                if (vout) {
                  vl_push(&vout[q], io + t->idx_offset, jo + t->idx_offset, S[0], S[1], S[2]);
                } else {
                  int64_t w = offs[q] + local[q]++;
                  out[q].ei[w] = io + t->idx_offset;
                  out[q].ej[w] = jo + t->idx_offset;
                  out[q].s[3 * w] = S[0];
                  out[q].s[3 * w + 1] = S[1];
                  out[q].s[3 * w + 2] = S[2];
                }
              }
              if (t->disjoint) break;
            }
          }
        }
      }
    }
  }
}

/* search.rs:291-342 / 344-446 (parallel two-pass) and 275-289 (serial). */
static void cl_search(const cell_t *c, const celllist_t *cl, const double *rc2, int k,
                      int disjoint, int par, int64_t idx_offset, nlo_list *out) {
  trav_t t = {c, cl, rc2, k, disjoint, idx_offset};
  int64_t n = cl->n;
  if (!par) {
    vlist *v = (vlist *)malloc((size_t)k * sizeof(vlist));
    for (int q = 0; q < k; ++q) vl_init(&v[q], n * 8);
    for (int64_t i = 0; i < n; ++i) atom_traverse(&t, i, NULL, NULL, NULL, v);
    for (int q = 0; q < k; ++q) out[q] = vl_take(&v[q]);
    free(v);
    return;
  }
  int64_t *cnt = (int64_t *)calloc((size_t)(n > 0 ? n : 1) * k, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < n; ++i) atom_traverse(&t, i, cnt + i * k, NULL, NULL, NULL);
  /* per-cutoff serial prefix (search.rs:386-396) */
  int64_t *offs = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * k * sizeof(int64_t));
  for (int q = 0; q < k; ++q) {
    int64_t acc = 0;
    for (int64_t i = 0; i < n; ++i) {
      offs[i * k + q] = acc;
      acc += cnt[i * k + q];
    }
    list_alloc(&out[q], acc);
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < n; ++i) atom_traverse(&t, i, NULL, out, offs + i * k, NULL);
  free(cnt);
  free(offs);
}

/* ---- brute force (search.rs:1187-1405) ----------------------------------- */
/* isolated SIMD kernel semantics: raw positions, i<j row-major, zero shifts */
static void bf_isolated(const double *pos, int64_t n, double rc2, int64_t off, nlo_list *out) {
  vlist v;
  vl_init(&v, n * 10);
  for (int64_t i = 0; i < n; ++i) {
    const double *pi = pos + 3 * i;
    for (int64_t j = i + 1; j < n; ++j) {
      const double *pj = pos + 3 * j;
      double d[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
      if (norm2(d) < rc2) vl_push(&v, i + off, j + off, 0, 0, 0);
    }
  }
  *out = vl_take(&v);
}

/* minimum-image double loop for one or more sorted cutoffs (search.rs:1319-1405) */
static void bf_mic(const cell_t *c, const double *pos, int64_t n, const double *rc2, int k,
                   int disjoint, int64_t off, nlo_list *out) {
  vlist *v = (vlist *)malloc((size_t)k * sizeof(vlist));
  for (int q = 0; q < k; ++q) vl_init(&v[q], n * 10);
  double max_sq = rc2[k - 1];
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t j = i + 1; j < n; ++j) {
      int32_t s[3];
      double d[3];
      shift_and_disp(c, pos + 3 * i, pos + 3 * j, s, d);
      double d2 = norm2(d);
      if (!(d2 < max_sq)) continue;
      for (int q = 0; q < k; ++q) {
        if (d2 < rc2[q]) {
          vl_push(&v[q], i + off, j + off, s[0], s[1], s[2]);
          if (disjoint) break;
        }
      }
    }
  }
  for (int q = 0; q < k; ++q) out[q] = vl_take(&v[q]);
  free(v);
}

/* ---- single-system dispatch (single.rs:8-164) ----------------------------- */
static void auto_box(const double *pos, int64_t n, double rc, double *h) {
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      double p = pos[3 * i + k];
      if (p < mn[k]) mn[k] = p;
      if (p > mx[k]) mx[k] = p;
    }
  double margin = rc + 1.0; /* AUTO_BOX_MARGIN, single.rs:6 */
  memset(h, 0, 9 * sizeof(double));
  for (int k = 0; k < 3; ++k) h[4 * k] = (mx[k] - mn[k]) + 2.0 * margin;
}

/* stable ascending order of cutoffs (Rust sort_by is stable) */
static void sort_cutoffs(const double *cut, int k, int *order) {
  for (int q = 0; q < k; ++q) order[q] = q;
  for (int a = 1; a < k; ++a) { /* insertion sort: stable */
    int v = order[a], b = a - 1;
    while (b >= 0 && cut[order[b]] > cut[v]) {
      order[b + 1] = order[b];
      --b;
    }
    order[b + 1] = v;
  }
}

/* One system: `cutoffs` in input order, k >= 1.  multi selects the
 * search_single_multi rules (MIC brute force for every system, grid at
 * max rc, always the parallel cell-list path). */
static int system_search(const double *pos, int64_t n, const double *h, const int *pbc,
                         const double *cutoffs, int k, int disjoint, int parallel, int multi,
                         int64_t idx_offset, nlo_list *out_input_order, int inner_par) {
  double rmax = NAN;
  for (int q = 0; q < k; ++q) rmax = isnan(rmax) ? cutoffs[q] : fmax(rmax, cutoffs[q]);
  cell_t c;
  int fpbc[3] = {0, 0, 0};
  double hb[9];
  if (h) {
    if (cell_new(h, pbc, &c) != NLO_OK) return NLO_ERR;
  } else {
    auto_box(pos, n, rmax, hb);
    if (cell_new(hb, fpbc, &c) != NLO_OK) return NLO_ERR;
  }
  double w[3];
  nlo_perpendicular_widths(c.hinv, w);
  double minw = fmin(fmin(w[0], w[1]), w[2]);
  int mic_safe = rmax * 2.0 < minw;
  int order[k];
  sort_cutoffs(cutoffs, k, order);
  double rc2[k];
  for (int q = 0; q < k; ++q) rc2[q] = cutoffs[order[q]] * cutoffs[order[q]];
  nlo_list tmp[k];
  if ((size_t)n < g_bf_threshold && mic_safe) {
    if (!multi && !c.pbc[0] && !c.pbc[1] && !c.pbc[2]) {
      bf_isolated(pos, n, rc2[0], idx_offset, &tmp[0]);
    } else {
      bf_mic(&c, pos, n, rc2, k, disjoint, idx_offset, tmp);
    }
  } else {
    celllist_t cl;
    int par = multi ? 1 : (parallel && (size_t)n >= g_par_threshold);
    cl_build(&c, pos, n, rmax, inner_par && par, &cl);
    cl_search(&c, &cl, rc2, k, disjoint, inner_par ? par : 0, idx_offset, tmp);
    cl_free(&cl);
  }
  for (int q = 0; q < k; ++q) out_input_order[order[q]] = tmp[q];
  return NLO_OK;
}

static int check_k(int k) {
  /* any number of cutoffs, as search.rs:344-361 (per-list state is sized
   * by k on the stack) */
  if (k > 100000) return set_err("too many cutoffs");
  return NLO_OK;
}

static void empty_lists(nlo_result *out, int k) {
  out->k = k;
  out->lists = (nlo_list *)calloc((size_t)(k > 0 ? k : 1), sizeof(nlo_list));
  for (int q = 0; q < k; ++q) list_alloc(&out->lists[q], 0);
}

/* single.rs:8-90 */
int nlo_search_single(const double *pos, int64_t n, const double *h, const int *pbc,
                      double cutoff, int parallel, nlo_result *out) {
  if (n == 0) {
    empty_lists(out, 1);
    return NLO_OK;
  }
  empty_lists(out, 0);
  out->k = 1;
  out->lists = (nlo_list *)calloc(1, sizeof(nlo_list));
  if (system_search(pos, n, h, pbc, &cutoff, 1, 0, parallel, 0, 0, out->lists, 1) != NLO_OK) {
    free(out->lists);
    out->lists = NULL;
    out->k = 0;
    return NLO_ERR;
  }
  return NLO_OK;
}

/* single.rs:92-164 */
int nlo_search_single_multi(const double *pos, int64_t n, const double *h, const int *pbc,
                            const double *cutoffs, int k, int disjoint, nlo_result *out) {
  if (check_k(k) != NLO_OK) return NLO_ERR;
  if (n == 0 || k == 0) {
    empty_lists(out, n == 0 ? k : 0);
    return NLO_OK;
  }
  out->k = k;
  out->lists = (nlo_list *)calloc((size_t)k, sizeof(nlo_list));
  if (system_search(pos, n, h, pbc, cutoffs, k, disjoint, 1, 1, 0, out->lists, 1) != NLO_OK) {
    free(out->lists);
    out->lists = NULL;
    out->k = 0;
    return NLO_ERR;
  }
  return NLO_OK;
}

/* batch.rs:10-290.  cells: n_cells x 9 row-major H (lattice vectors as
 * columns), has_cell[s] != 0 selects it, pbc: n_cells x 3.  cells == NULL
 * means every system is auto-boxed. */
static int batch_impl(const double *pos, int64_t n, const int32_t *batch, int64_t n_cells,
                      const double *cells, const int *has_cell, const int *pbc,
                      const double *cutoffs, int k, int disjoint, int parallel, int multi,
                      nlo_result *out) {
  if (check_k(k) != NLO_OK) return NLO_ERR;
  if (n == 0 || (multi && k == 0)) {
    empty_lists(out, k);
    return NLO_OK;
  }
  /* segments at value changes (batch.rs:25-37) */
  int64_t ns = 1;
  for (int64_t i = 1; i < n; ++i) ns += batch[i] != batch[i - 1];
  int64_t *st = (int64_t *)malloc((size_t)(ns + 1) * sizeof(int64_t));
  int64_t s = 0;
  st[0] = 0;
  for (int64_t i = 1; i < n; ++i)
    if (batch[i] != batch[i - 1]) st[++s] = i;
  st[ns] = n;
  if (cells && n_cells < ns) {
    char msg[128];
    snprintf(msg, sizeof msg, "Expected at least %lld cells, but got %lld", (long long)ns,
             (long long)n_cells);
    free(st);
    return set_err(msg);
  }
  nlo_list *per = (nlo_list *)calloc((size_t)(ns * k), sizeof(nlo_list));
  int err = 0;
  char errmsg[256] = {0};
#pragma omp parallel for schedule(dynamic, 1) if (parallel)
  for (int64_t q = 0; q < ns; ++q) {
    const double *h = (cells && has_cell[q]) ? cells + 9 * q : NULL;
    const int *pb = (cells && has_cell[q]) ? pbc + 3 * q : NULL;
    int rc = system_search(pos + 3 * st[q], st[q + 1] - st[q], h, pb, cutoffs, k, disjoint,
                           parallel, multi, st[q], per + q * k, 0);
    if (rc != NLO_OK) {
#pragma omp critical
      {
        err = 1;
        snprintf(errmsg, sizeof errmsg, "%s", g_err);
      }
    }
  }
  if (err) {
    for (int64_t q = 0; q < ns * k; ++q) {
      free(per[q].ei);
      free(per[q].ej);
      free(per[q].s);
    }
    free(per);
    free(st);
    return set_err(errmsg);
  }
  out->k = k;
  out->lists = (nlo_list *)calloc((size_t)k, sizeof(nlo_list));
  for (int c = 0; c < k; ++c) {
    int64_t tot = 0;
    for (int64_t q = 0; q < ns; ++q) tot += per[q * k + c].n;
    list_alloc(&out->lists[c], tot);
    int64_t w = 0;
    for (int64_t q = 0; q < ns; ++q) {
      nlo_list *l = &per[q * k + c];
      memcpy(out->lists[c].ei + w, l->ei, (size_t)l->n * sizeof(int64_t));
      memcpy(out->lists[c].ej + w, l->ej, (size_t)l->n * sizeof(int64_t));
      memcpy(out->lists[c].s + 3 * w, l->s, (size_t)l->n * 3 * sizeof(int32_t));
      w += l->n;
      free(l->ei);
      free(l->ej);
      free(l->s);
    }
  }
  free(per);
  free(st);
  return NLO_OK;
}

int nlo_search_batch(const double *pos, int64_t n, const int32_t *batch, int64_t n_cells,
                     const double *cells, const int *has_cell, const int *pbc, double cutoff,
                     int parallel, nlo_result *out) {
  return batch_impl(pos, n, batch, n_cells, cells, has_cell, pbc, &cutoff, 1, 0, parallel, 0,
                    out);
}

int nlo_search_batch_multi(const double *pos, int64_t n, const int32_t *batch, int64_t n_cells,
                           const double *cells, const int *has_cell, const int *pbc,
                           const double *cutoffs, int k, int disjoint, nlo_result *out) {
  return batch_impl(pos, n, batch, n_cells, cells, has_cell, pbc, cutoffs, k, disjoint, 1, 1,
                    out);
}

/* ---- introspection for the reference's cell-list unit tests -------------- */
/* search.rs:248-256 get_atoms_in_bin + num_bins; returns count written. */
int64_t nlo_celllist_bin_atoms(const double *pos, int64_t n, const double *h, const int *pbc,
                               double cutoff, int64_t bx, int64_t by, int64_t bz,
                               int64_t *nbins_out, int64_t *atoms_out, int64_t cap) {
  cell_t c;
  if (cell_new(h, pbc, &c) != NLO_OK) return -1;
  celllist_t cl;
  cl_build(&c, pos, n, cutoff, 0, &cl);
  for (int k = 0; k < 3; ++k) nbins_out[k] = cl.nb[k];
  int64_t cnt = 0;
  if (bx < cl.nb[0] && by < cl.nb[1] && bz < cl.nb[2]) {
    int64_t lin = bx + cl.nb[0] * (by + cl.nb[1] * bz);
    int64_t r = cl.bin_ranks[lin];
    for (int64_t j = cl.cell_starts[r]; j < cl.cell_starts[r + 1]; ++j) {
      if (cnt < cap) atoms_out[cnt] = cl.particles[j];
      ++cnt;
    }
  }
  cl_free(&cl);
  return cnt;
}

int nlo_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void nlo_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
