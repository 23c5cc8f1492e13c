/*
 * nl_api.h -- C ABI of libnlmi355, the MI355X (gfx950) neighbour-list engine.
 *
 * Drop-in for the construction path of neighborlist-rs 0.1.9.  The four
 * search entry points replace the reference's dispatch layer (the functions
 * its Rust API and PyO3 module call):
 *
 *   nl_search_single        <- single::search_single        src/single.rs:8-13
 *   nl_search_single_multi  <- single::search_single_multi  src/single.rs:92-97
 *   nl_search_batch         <- batch::search_batch          src/batch.rs:10-16
 *   nl_search_batch_multi   <- batch::search_batch_multi    src/batch.rs:149-155
 *
 * The reference returns EdgeResult = (Vec<i64> edge_i, Vec<i64> edge_j,
 * Vec<i32> shifts) per cutoff (src/search.rs:14).  Here the callee keeps the
 * lists in device memory behind an nl_result handle; the caller sizes its own
 * buffers with nl_result_num_edges and copies with nl_result_copy (or takes
 * the device pointers with nl_result_device).  Semantics match the reference
 * exactly: half list, r_ij = pos_j + H S - pos_i, i < j or (i == j and S != 0),
 * edge sets bit-identical after a canonical sort on (i, j, S).
 *
 * Conventions
 *   positions  n x 3 doubles, row-major (C-contiguous)
 *   cell h     9 doubles, row-major H, lattice vectors are the COLUMNS
 *              (the Matrix3 that src/api.rs:93-98 builds); NULL = auto-box
 *   pbc        3 bytes, nonzero = periodic (ignored when h is NULL)
 *   batch      n int32 ids; systems are maximal runs of equal ids
 *              (src/batch.rs:25-37)
 *   cells      n_cells x 9 doubles; has_cell[s] selects cells[s] (else the
 *              system is auto-boxed, the Rust None); pbc n_cells x 3 bytes.
 *              has_cell == NULL means every system is auto-boxed (the
 *              cells=None of src/api.rs:66-73).
 *
 * Errors: every function returning int returns 0 on success and nonzero on
 * failure; nl_last_error() gives the message of the calling thread's last
 * failure, with the reference's texts where it has one ("Cell matrix is not
 * invertible", "positions and batch must have the same length",
 * "Expected at least {n} cells, but got {m}").
 *
 * Threading: calls are serialised per device inside the library; results
 * are independent objects and may be used from any thread.
 */
#ifndef NL_API_H
#define NL_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nl_result nl_result;

/* ---- search entry points (reference dispatch layer) --------------------- */

/* single::search_single (src/single.rs:8-90).  `parallel` is accepted for
 * signature parity and has no effect on the GPU. */
int nl_search_single(const double* positions, size_t n, const double* h, const uint8_t* pbc,
                     double cutoff, int parallel, nl_result** out);

/* single::search_single_multi (src/single.rs:92-164).  One list per cutoff,
 * returned in input order; disjoint != 0 assigns each pair only to the
 * smallest cutoff it satisfies. */
int nl_search_single_multi(const double* positions, size_t n, const double* h,
                           const uint8_t* pbc, const double* cutoffs, size_t n_cutoffs,
                           int disjoint, nl_result** out);

/* batch::search_batch (src/batch.rs:10-147).  Indices are global (offset by
 * the start of each system), systems concatenated in order. */
int nl_search_batch(const double* positions, size_t n, const int32_t* batch, size_t n_cells,
                    const double* cells, const uint8_t* has_cell, const uint8_t* pbc,
                    double cutoff, int parallel, nl_result** out);

/* batch::search_batch_multi (src/batch.rs:149-290). */
int nl_search_batch_multi(const double* positions, size_t n, const int32_t* batch,
                          size_t n_cells, const double* cells, const uint8_t* has_cell,
                          const uint8_t* pbc, const double* cutoffs, size_t n_cutoffs,
                          int disjoint, nl_result** out);

/* ---- results -------------------------------------------------------------- */

size_t nl_result_num_lists(const nl_result* r);
/* number of edges in list `list` (input cutoff order) */
int64_t nl_result_num_edges(const nl_result* r, size_t list);
/* Copy list `list` into host buffers: edge_i[M], edge_j[M], shifts[3M]
 * (x, y, z per edge).  Any pointer may be NULL to skip that array. */
int nl_result_copy(const nl_result* r, size_t list, int64_t* edge_i, int64_t* edge_j,
                   int32_t* shifts);
/* Device pointers of list `list`: edge_index (2 x M int64, sources then
 * targets, the PyG layout of src/python_api.rs:177-183) and shifts (M x 3
 * int32).  Valid until nl_result_free. */
int nl_result_device(const nl_result* r, size_t list, const int64_t** edge_index,
                     const int32_t** shifts);
void nl_result_free(nl_result* r);

/* ---- configuration (src/config.rs:16-39; src/python_api.rs:548-591) ------ */
size_t nl_get_brute_force_threshold(void);
void nl_set_brute_force_threshold(size_t v);
size_t nl_get_parallel_threshold(void);
void nl_set_parallel_threshold(size_t v);
size_t nl_get_stack_threshold(void);
void nl_set_stack_threshold(size_t v); /* clamped to 1024 as in config.rs:36-38 */

/* thread-local message of the last failure ("" if none) */
const char* nl_last_error(void);
/* library version string */
const char* nl_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NL_API_H */
