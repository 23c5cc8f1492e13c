/*
 * nl_engine_ext.h -- engine entry points of libnlmi355 that have no
 * counterpart in neighborlist-rs.  They serve callers that keep data on the
 * GPU (device-resident positions and outputs), the batched and served forms
 * built on the plain path, and measurement.  Same error convention as
 * nl_api.h (0 = success, message in nl_last_error()).
 */
#ifndef NL_ENGINE_EXT_H
#define NL_ENGINE_EXT_H

#include <stddef.h>
#include <stdint.h>

#include "nl_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Ragged batch whose positions are already in device memory.  sys_sizes
 * (host) gives the atom count of each of n_sys consecutive systems.
 * global_index != 0 offsets indices by the system start (batch semantics);
 * 0 keeps each system's local indices (serving semantics).  want_ranges
 * records per-system edge ranges (nl_result_system_ranges).  want_dist
 * also emits the FP64 length |r_ij| of every edge (nl_result_copy_dist). */
int nl_search_batch_device(const double* d_positions, size_t n, const int64_t* sys_sizes,
                           size_t n_sys, const double* cells, const uint8_t* has_cell,
                           const uint8_t* pbc, const double* cutoffs, size_t n_cutoffs,
                           int disjoint, int multi, int global_index, int want_ranges,
                           int want_dist, nl_result** out);

/* Host-input variant of the above (the batch entry point with explicit
 * system sizes instead of batch ids, plus the range / distance options). */
int nl_search_systems(const double* positions, size_t n, const int64_t* sys_sizes,
                      size_t n_sys, const double* cells, const uint8_t* has_cell,
                      const uint8_t* pbc, const double* cutoffs, size_t n_cutoffs, int disjoint,
                      int multi, int global_index, int want_ranges, int want_dist,
                      nl_result** out);

/* The general form of the two above, for callers that keep data on the GPU
 * (a model's step, a serving loop).  options is a bit set of the NL_* flags
 * below.  stream is the caller's HIP stream (hipStream_t; NULL = the null
 * stream): the search reads device-resident positions after the work
 * already queued there, and output blocks released through DLPack views
 * are handed out again only after the work queued there before the release.
 *
 * NL_FULL gives full (directed) lists: every pair of the half list in both
 * directions, (i, j, S) at an even position e and (j, i, -S) at e + 1; a
 * self image (i, i, S) appears once per S (the half list already holds both
 * S and -S).  Per-system ranges, distances and displacements cover the full
 * lists.
 *
 * NL_DISP records each edge's displacement r_ij = r_j + H S - r_i in FP64,
 * the vector whose squared norm the search compared with the cutoff
 * ((w_j - w_i) + H bs on wrapped positions, in the edge's own orientation),
 * so sqrt((x^2 + y^2) + z^2) of it is the NL_DIST distance exactly. */
#define NL_DEVICE_POSITIONS 1u /* positions is a device pointer */
#define NL_LOCAL_INDEX 2u      /* indices local to each system (else offset by its start) */
#define NL_RANGES 4u           /* per-system edge ranges (nl_result_system_ranges) */
#define NL_DIST 8u             /* FP64 |r_ij| per edge (nl_result_copy_dist) */
#define NL_FULL 16u            /* full (directed) lists */
#define NL_DISP 32u            /* FP64 r_ij per edge (nl_result_copy_disp) */
int nl_search_ex(const double* positions, size_t n, const int64_t* sys_sizes, size_t n_sys,
                 const double* cells, const uint8_t* has_cell, const uint8_t* pbc,
                 const double* cutoffs, size_t n_cutoffs, int disjoint, int multi,
                 unsigned options, void* stream, nl_result** out);

/* A zero-copy DLPack view (DLManagedTensor*, ROCm device) of one array of
 * list `list`: what = 0 edge_index (2, M) int64, 1 shift (M, 3) int32,
 * 2 distance (M,) float64, 3 displacement (M, 3) float64.  The view shares ownership of the result: the
 * lists stay valid until nl_result_free and the deleter of every view have
 * run, whichever comes last.  Work on the arrays belongs on the stream given
 * to nl_search_ex; the view's deleter fences that stream. */
int nl_result_dlpack(const nl_result* r, size_t list, int what, void** managed);

/* ranges[(s * n_lists + list) * 2 + {0,1}] = [first, end) edge of system s
 * within list `list` (input cutoff order); mode[s] = 0 brute force on raw
 * positions, 1 brute force minimum image, 2 cell list.  Either may be NULL. */
int nl_result_system_ranges(const nl_result* r, int64_t* ranges, int32_t* modes);
/* FP64 |r_ij| of list `list` into dist[M] (requires want_dist). */
int nl_result_copy_dist(const nl_result* r, size_t list, double* dist);
/* FP64 r_ij of list `list` into disp[M][3] (requires NL_DISP). */
int nl_result_copy_disp(const nl_result* r, size_t list, double* disp);

/* Stage timing with hipEvents on the engine stream (milliseconds). */
void nl_set_timing(int on);
int nl_last_timing(float* build_ms, float* count_ms, float* fill_ms, float* total_ms,
                   int64_t* edges, int64_t* bins);

/* Tunables: GPU bin edge = max cutoff / cell_div (1..4), or 0 (the default):
 * each system takes the most bins per cutoff, up to 4, that still hold six
 * atoms at its density; FP32 filter with FP64 band recheck on the cell-list
 * path (0/1).  Neither changes results. */
void nl_set_cell_div(int d);
int nl_get_cell_div(void);
/* Single systems of at most 1024 atoms below the brute-force threshold run
 * without the nine-launch pipeline: up to 128 atoms (74 with the
 * minimum-image branch) as one kernel launch (plan, counts, scan and fill in
 * one block), above that as two (counts over many blocks with the scan in
 * the last block, then the fill).  The lists
 * are byte-identical either way.  1 (the default): on, for isolated systems
 * and for cells whose widths leave twice the largest cutoff clear of the
 * minimum-image bound; 2: one launch, 3: two launches, for every single
 * system of at most 1024 atoms (the kernels' plan hands cell-list systems
 * back to the pipeline); 0: off.  Batches of two or more systems, each of at
 * most 1024 atoms and below the brute-force threshold (at most 65,536
 * atoms x lists in all, no per-system ranges), take three launches whenever
 * the mode is not 0: counts (one row per wave, four rows per block, every
 * block planning its own system), a one-block scan, the fill; modes 2 and 3
 * skip the host's minimum-image screen there too.  No counterpart in the
 * reference (a GPU launch choice). */
void nl_set_small_path(int mode);
int nl_get_small_path(void);
/* Order of the cell-list bins, and so of the sorted atoms and the work items:
 * 1 ranks each system's bins by the Morton code of their coordinates, as
 * CellList::build ranks its bins (src/search.rs:126-146, interleave_3
 * :1177-1185), so neighbouring bins sit close together in memory and in the
 * order the work items run; 0 (the default, about 1% faster on the ragged
 * batch) keeps the linear bin index.  Never changes an edge set (only the
 * order of edges within a list). */
void nl_set_bin_order(int morton);
int nl_get_bin_order(void);
/* Build stages of a batch whose systems all have at most 12288 atoms (linear
 * bin order): one launch, one block per system (plan, counting sort by bin in
 * LDS, bin offsets, sorted positions, work items), or the separate launches
 * of the general path.  1 (the default) takes the one launch when every
 * system has at most 4096 atoms, or at most 12288 and the batch holds at
 * least four times its largest system's atoms (one block per system: a lone
 * large system builds faster over many blocks); 2 whenever the systems fit;
 * 0 never.  Never changes a result (the lists are byte-identical). */
void nl_set_fused_build(int on);
int nl_get_fused_build(void);
void nl_set_fp32_filter(int on);
int nl_get_fp32_filter(void);
/* Work decomposition of the cell-list pair kernels: `group` home atoms of
 * one bin per wavefront (1..64) and `parts` slices of the half stencil per
 * home group (1..16).  Results never depend on them. */
void nl_set_work_split(int group, int parts);
int nl_get_work_split(int* group, int* parts);

/* Autotuning: on a new workload class (number of lists, multi, log2 of
 * atoms, systems and max cutoff) the engine times a staged search over the
 * four tunables above and keeps the fastest; results never change. */
void nl_set_autotune(int on);
int nl_get_autotune(void);
void nl_clear_tuning(void);
/* one line per tuned class into buf (truncated to len); returns the length */
size_t nl_tune_report(char* buf, size_t len);

/* ---- graph-captured steady state (batching form) -------------------------
 * A plan fixes a batch shape (system sizes, cells, cutoffs) whose positions
 * live at a fixed device address and change between runs (an MD loop, a
 * serving slot).  nl_plan_create sizes the output (capacity <= 0: from one
 * eager count, plus 1/8 headroom) and captures the whole sequence (uploads,
 * plan, bin, radix sort, finalize, count, scan, fill, bounds download) in a
 * hipGraph; nl_plan_run replays it with one launch.  When the lists outgrow
 * the capacity the fill writes nothing, the plan grows its buffers,
 * re-captures and replays once more (reported by nl_plan_stats). */
typedef struct nl_plan nl_plan;
int nl_plan_create(const double* d_positions, size_t n, const int64_t* sys_sizes, size_t n_sys,
                   const double* cells, const uint8_t* has_cell, const uint8_t* pbc,
                   const double* cutoffs, size_t n_cutoffs, int disjoint, int multi,
                   int global_index, int64_t capacity, nl_plan** out);
int nl_plan_run(nl_plan* p);
int64_t nl_plan_num_edges(const nl_plan* p, size_t list);
int nl_plan_device(const nl_plan* p, size_t list, const int64_t** edge_index,
                   const int32_t** shifts);
int nl_plan_copy(const nl_plan* p, size_t list, int64_t* edge_i, int64_t* edge_j,
                 int32_t* shifts);
int nl_plan_stats(const nl_plan* p, int64_t* replays, int64_t* recaptures, int64_t* capacity);
void nl_plan_free(nl_plan* p);
/* The general form of nl_plan_create: options are the NL_* flags of
 * nl_search_ex (NL_LOCAL_INDEX, NL_RANGES, NL_DIST, NL_FULL, NL_DISP;
 * positions are device-resident in any case), captured into the graph with
 * the rest, so a replay also emits full lists, distances or displacements
 * and downloads the per-system ranges. */
int nl_plan_create_ex(const double* d_positions, size_t n, const int64_t* sys_sizes,
                      size_t n_sys, const double* cells, const uint8_t* has_cell,
                      const uint8_t* pbc, const double* cutoffs, size_t n_cutoffs, int disjoint,
                      int multi, unsigned options, int64_t capacity, nl_plan** out);
/* Device address of one array of list `list` after the last run: what = 0
 * edge_index (2, M) int64, 1 shift (M, 3) int32, 2 distance (M,) float64
 * (NL_DIST), 3 displacement (M, 3) float64 (NL_DISP).  Valid until the next
 * nl_plan_run (which moves the arrays when the lists outgrow the capacity)
 * or nl_plan_free. */
int nl_plan_array(const nl_plan* p, size_t list, int what, const void** ptr);
/* Per-system ranges and modes of the last run, as nl_result_system_ranges
 * (NL_RANGES). */
int nl_plan_system_ranges(const nl_plan* p, int64_t* ranges, int32_t* modes);

/* ---- serving: continuous batching -----------------------------------------
 * No counterpart in the reference (its API is synchronous, one call per
 * request); this is the serving form of nl_search_single with one cutoff.
 * Requests are submitted from any thread and return an id at once.  Up to
 * `slots` ragged batches are in flight, each on its own HIP stream with its
 * own buffers and host thread; a free slot takes the queued requests in
 * arrival order, up to max_atoms atoms and max_systems systems per batch,
 * while the estimated edges of all batches on the device stay within
 * edge_budget (a request alone is always admitted).  Each request keeps its
 * own cell, periodicity and cutoff, its lists use its own atom indices, its
 * edges are nl_search_single's for the same input, and its arrays (order
 * included) are the same whatever batch it ran in.  Results stay in device
 * memory until taken or released.  nl_server_wait: 0 done, 1 timed out
 * (timeout_s >= 0), nonzero with nl_last_error on failure.  nl_server_take
 * copies the lists to host memory (any of the outputs may be NULL) and
 * releases the request; nl_server_result_device gives their device
 * addresses instead (edge_i and edge_j: M int64 each, shifts: M x 3 int32),
 * valid until nl_server_release, which the caller issues once its own device
 * work on them has completed.  The edge budget covers the batches on the
 * device, not results held by callers. */
typedef struct nl_server nl_server;
nl_server* nl_server_create(int slots, int64_t max_atoms, int64_t max_systems,
                            int64_t edge_budget);
int64_t nl_server_submit(nl_server* s, const double* positions, size_t n, const double* cell,
                         const uint8_t* pbc, double cutoff);
/* nl_server_submit with flags.  NL_SUBMIT_BORROW: the server does not copy
 * the positions at submit; it reads them from the caller's buffer when the
 * request's batch is staged, so the buffer must stay valid and unchanged
 * until nl_server_wait reports the request done or failed.  Saves a host
 * copy of every request (submission is otherwise bound by that copy: about
 * 10 GB/s on one host thread). */
#define NL_SUBMIT_BORROW 1u
int64_t nl_server_submit_ex(nl_server* s, const double* positions, size_t n, const double* cell,
                            const uint8_t* pbc, double cutoff, unsigned flags);
int nl_server_wait(nl_server* s, int64_t id, double timeout_s);
int nl_server_result_info(nl_server* s, int64_t id, int64_t* n_edges, double* latency_s,
                          int* batch_size);
int nl_server_take(nl_server* s, int64_t id, int64_t* edge_i, int64_t* edge_j,
                   int32_t* shifts);
int nl_server_result_device(nl_server* s, int64_t id, const int64_t** edge_i,
                            const int64_t** edge_j, const int32_t** shifts);
int nl_server_release(nl_server* s, int64_t id);
/* counters since creation; refills = batches whose edge estimate was short
 * (their fill pass ran a second time into larger buffers) */
int nl_server_stats(nl_server* s, int64_t* completed, int64_t* batches, int64_t* max_inflight,
                    int64_t* max_batch, int64_t* refills);
/* device and page-locked allocations the server made after its creation
 * (output blocks, slot buffers that grew); freeing a device buffer
 * synchronises the device, so a warmed-up server should report no new ones */
int nl_server_allocations(nl_server* s, int64_t* n);
/* waits for every submitted request, stops the slot threads, frees the buffers */
void nl_server_free(nl_server* s);
/* A server whose requests all get the outputs in `options` (NL_FULL: full
 * lists; NL_DIST: FP64 |r_ij|; NL_DISP: FP64 r_ij, as nl_search_ex).  The
 * edge estimate of a full-list request is twice that of its half list. */
nl_server* nl_server_create_ex(int slots, int64_t max_atoms, int64_t max_systems,
                               int64_t edge_budget, unsigned options);
/* nl_server_take and nl_server_result_device with the distances (M float64)
 * and displacements (M x 3 float64) too; requesting either from a server
 * created without it is an error, NULL skips it. */
int nl_server_take_ex(nl_server* s, int64_t id, int64_t* edge_i, int64_t* edge_j,
                      int32_t* shifts, double* dist, double* disp);
int nl_server_result_arrays(nl_server* s, int64_t id, const int64_t** edge_i,
                            const int64_t** edge_j, const int32_t** shifts, const double** dist,
                            const double** disp);

/* ---- skin lists for molecular dynamics -------------------------------------
 * The dynamic manager of the reference's specification (spec.md:41-48,
 * 209-222; never implemented there): lists built at r_list = rc_q + skin_q
 * from the positions of the last build (the reference positions) and kept
 * while no atom has moved more than min_q(skin_q) / 2 since then, so every
 * pair now within rc_q is in list q.  The lists are half (or, with NL_FULL,
 * full) lists at r_list; a consumer applies rc itself.
 *
 * nl_skin_create copies the device positions and builds once (a
 * graph-captured plan).  nl_skin_update measures the largest displacement
 * from the reference positions on the device and rebuilds when it exceeds
 * the bound (or is NaN), or when `cells` (row-major H per system, NULL =
 * unchanged) differs from the current cells; *rebuilt says whether it did.
 * nl_skin_check measures without changing anything.  Device list pointers
 * (nl_skin_device, same layout as nl_result_device) stay valid until the
 * next update that rebuilds.  options: NL_LOCAL_INDEX, NL_FULL.  disjoint
 * must be 0: the lists are cumulative (a disjoint shell would lose pairs
 * that cross a shell boundary between rebuilds), and a nonzero value is
 * rejected with an error.  stream:
 * the caller's stream, as for nl_search_ex.  Positions must keep one
 * continuous trajectory (no re-wrapping into the cell between updates, or a
 * wrapped atom counts as a large displacement and forces a rebuild). */
typedef struct nl_skin nl_skin;
int nl_skin_create(const double* d_positions, size_t n, const int64_t* sys_sizes, size_t n_sys,
                   const double* cells, const uint8_t* has_cell, const uint8_t* pbc,
                   const double* cutoffs, const double* skins, size_t n_cutoffs, int disjoint,
                   int multi, unsigned options, void* stream, nl_skin** out);
int nl_skin_update(nl_skin* s, const double* d_positions, const double* cells, void* stream,
                   int* rebuilt);
int nl_skin_check(nl_skin* s, const double* d_positions, void* stream, double* max_disp,
                  int* needs_rebuild);
int64_t nl_skin_num_edges(const nl_skin* s, size_t list);
int nl_skin_device(const nl_skin* s, size_t list, const int64_t** edge_index,
                   const int32_t** shifts);
/* updates since creation, builds (the first included), last measured
 * largest displacement */
int nl_skin_stats(const nl_skin* s, int64_t* updates, int64_t* rebuilds, double* last_disp);
void nl_skin_free(nl_skin* s);

/* ---- allocation-free search --------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * or page-locked memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): nl_search_ex on device-resident
 * positions (options: NL_LOCAL_INDEX only), with its device workspace in a
 * 64 MiB array of the library (one per GPU, held until the process ends),
 * host staging in ordinary memory, and outputs in the caller's buffers.  It
 * calls no allocation routine.  Requests whose workspace exceeds 64 MiB
 * (roughly 200,000 atoms with three lists) are rejected with an error.
 *
 * d_edges holds 2 * capacity int64 and d_shifts 3 * capacity int32 (device
 * memory).  On return list q (input cutoff order) holds list_len[q] edges:
 * its edge_index (2, list_len[q]) at d_edges + 2 * list_base[q] and its
 * shifts (list_len[q], 3) at d_shifts + 3 * list_base[q], the same edges as
 * nl_search_ex.  The call runs on `stream` (hipStream_t, NULL = the null
 * stream) after the work queued there and synchronises it before returning.
 * Returns 0 on success, 2 when the lists exceed the capacity (nothing is
 * written; list_base / list_len give the sizes), 1 on other errors.  Every
 * call on a GPU uses that GPU's one workspace: calls are serialised within
 * the process, and calls that overlap on one GPU from elsewhere would race. */
int nl_search_noalloc(const double* d_positions, size_t n, const int64_t* sys_sizes,
                      size_t n_sys, const double* cells, const uint8_t* has_cell,
                      const uint8_t* pbc, const double* cutoffs, size_t n_cutoffs, int disjoint,
                      int multi, unsigned options, int64_t* d_edges, int32_t* d_shifts,
                      int64_t capacity, int64_t* list_base, int64_t* list_len, void* stream);

/* Host-side cell helpers with the reference's arithmetic (no GPU needed):
 * H^-1 by the cofactor formula of Cell::new (cell.rs:18-21; returns nonzero
 * with "Cell matrix is not invertible" when det == 0) and Cell::wrap
 * (cell.rs:52-72).  h is row-major H, lattice vectors as columns. */
int nl_cell_inverse(const double* h, double* hinv);
int nl_cell_wrap(const double* h, const uint8_t* pbc, const double* pos, double* out);

/* Device memory plumbing for callers without a framework of their own
 * (benchmarks, tests): plain hipMalloc / hipFree / synchronous copies. */
void* nl_device_malloc(size_t bytes);
int nl_device_free(void* p);
int nl_memcpy_h2d(void* dst, const void* src, size_t bytes);
int nl_memcpy_d2h(void* dst, const void* src, size_t bytes);
int nl_memcpy_d2d(void* dst, const void* src, size_t bytes);
int nl_device_synchronize(void);

/* Page-locked host memory for lists copied to host arrays (no counterpart in
 * the reference, whose lists never leave host memory): a device-to-host copy
 * into it runs at the link's DMA rate.  nl_host_alloc returns NULL when the
 * pool's limit (8 GiB handed out) is reached or the runtime refuses; the
 * caller then uses ordinary memory.  Blocks are recycled by nl_host_free. */
void* nl_host_alloc(size_t bytes);
void nl_host_free(void* p);

/* device selection for subsequent calls on this thread */
int nl_set_device(int device);
int nl_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* NL_ENGINE_EXT_H */
