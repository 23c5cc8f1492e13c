"""neighborlist_rs -- drop-in Python surface of neighborlist-rs, served by the
MI355X engine libnlmi355 (hand-written HIP kernels for gfx950).

Function names, argument order, defaults, dict layout, dtypes and error texts
follow the reference's PyO3 module (src/python_api.rs:15-626,
neighborlist_rs.pyi:1-56).  Every search runs on the GPU through the C ABI in
include/nl_api.h; there is no CPU fallback.

Outputs: {"edge_index": (2, M) int64, "shift": (M, 3) int32}, a half list with
r_ij = pos_j + S . H - pos_i (src/python_api.rs:177-183).
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import lib

__all__ = [
    "PyCell", "build_neighborlists", "build_from_ase", "build_multi_from_ase",
    "build_neighborlists_multi", "build_neighborlists_batch", "build_neighborlists_batch_multi",
    "get_num_threads", "set_num_threads", "init_logging", "get_brute_force_threshold",
    "set_brute_force_threshold", "get_parallel_threshold", "set_parallel_threshold",
    "get_stack_threshold", "set_stack_threshold",
]

_num_threads = None
_log_level = None


# ---------------------------------------------------------------------------
def _fmt(x: float) -> str:
    """Rust `{}` formatting of an f64 (shortest positional form)."""
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "inf" if x > 0 else "-inf"
    return np.format_float_positional(float(x), trim="-")


class PyCell:
    """Periodic cell (src/python_api.rs:15-99).  `h` is H with the lattice
    vectors as columns; pbc defaults to all periodic."""

    __slots__ = ("_h", "_pbc")

    def __init__(self, h, pbc=None):
        try:
            rows = [list(r) for r in h]
        except TypeError:
            raise TypeError("h must be a 3x3 sequence of floats") from None
        if len(rows) != 3 or any(len(r) != 3 for r in rows):
            raise ValueError("Cell matrix must be 3x3")
        H = np.ascontiguousarray(np.array(rows, dtype=np.float64).reshape(9))
        if pbc is None:
            P = (True, True, True)
        else:
            P = tuple(bool(v) for v in pbc)
            if len(P) != 3:
                raise ValueError("pbc must have 3 entries")
        inv = np.zeros(9)
        if lib.nl_cell_inverse(H.ctypes.data_as(_lib.dp), inv.ctypes.data_as(_lib.dp)) != 0:
            raise ValueError(_lib.last_error())
        self._h = H
        self._pbc = P

    @property
    def h(self):
        return self._h.reshape(3, 3).copy()

    @property
    def pbc(self):
        return list(self._pbc)

    def wrap(self, pos):
        """Cell::wrap (cell.rs:52-72)."""
        p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(3))
        out = np.zeros(3)
        pb = (C.c_uint8 * 3)(*[1 if v else 0 for v in self._pbc])
        rc = lib.nl_cell_wrap(self._h.ctypes.data_as(_lib.dp), pb, p.ctypes.data_as(_lib.dp),
                              out.ctypes.data_as(_lib.dp))
        if rc != 0:
            raise ValueError(_lib.last_error())
        return [float(v) for v in out]

    def __repr__(self):
        h = self._h
        rows = ", ".join("[" + ", ".join(_fmt(h[3 * r + c]) for c in range(3)) + "]"
                         for r in range(3))
        p = ", ".join("true" if v else "false" for v in self._pbc)
        return f"PyCell(h=[{rows}], pbc=[{p}])"

    @staticmethod
    def from_ase(atoms):
        """PyCell from an ASE-like object: ASE rows are transposed into H
        (src/python_api.rs:70-98)."""
        cell_obj = atoms.get_cell()
        arr = cell_obj.__array__() if hasattr(cell_obj, "__array__") else cell_obj
        c = np.asarray(arr, dtype=np.float64)
        if c.shape != (3, 3):
            raise ValueError("Cell must be 3x3")
        pbc = [bool(v) for v in np.asarray(atoms.get_pbc()).reshape(3)]
        return PyCell(c.T.tolist(), pbc)


# ---------------------------------------------------------------------------
def _positions(positions):
    """get_positions (src/python_api.rs:116-137): an (N, 3) float64 array;
    non-contiguous input is copied."""
    a = positions if isinstance(positions, np.ndarray) else np.asarray(positions)
    if a.dtype != np.float64:
        raise TypeError("argument 'positions': array must be float64")
    if a.ndim != 2:
        raise TypeError(f"argument 'positions': expected a 2-d array, got {a.ndim}-d")
    if a.shape[1] != 3:
        raise ValueError("Positions must be (N, 3)")
    return np.ascontiguousarray(a)


def _batch_ids(batch):
    b = batch if isinstance(batch, np.ndarray) else np.asarray(batch)
    if b.dtype != np.int32:
        raise TypeError("argument 'batch': array must be int32")
    if b.ndim != 1:
        raise TypeError("argument 'batch': expected a 1-d array")
    return np.ascontiguousarray(b)


def _check(rc):
    if rc != 0:
        raise ValueError(_lib.last_error())


def _empty():
    return {"edge_index": np.zeros((2, 0), np.int64), "shift": np.zeros((0, 3), np.int32)}


_PINNED_MIN = 1 << 16  # bytes; smaller lists are copied into ordinary numpy memory


class _Pinned:
    """Owner of a page-locked block of the engine's pool: the block goes back
    to the pool when the last array over it is gone."""
    __slots__ = ("p",)

    def __init__(self, p):
        self.p = p

    def __del__(self):
        free = getattr(lib, "nl_host_free", None)  # None at interpreter exit
        if free is not None and self.p:
            free(C.c_void_p(self.p))
            self.p = None


def _host_block(specs):
    """numpy arrays for [(shape, dtype), ...], back to back in one page-locked
    block (a device-to-host copy into it runs at the link's DMA rate), or in
    ordinary memory below _PINNED_MIN bytes or when the pool has no block.
    List 8-byte types first: every array then starts aligned."""
    sizes = [int(np.prod(shape)) * np.dtype(dt).itemsize for shape, dt in specs]
    total = sum(sizes)
    p = lib.nl_host_alloc(total) if total >= _PINNED_MIN else None
    if not p:
        return [np.empty(shape, dt) for shape, dt in specs]
    buf = (C.c_char * total).from_address(p)
    buf._owner = _Pinned(p)
    out, off = [], 0
    for (shape, dt), nb in zip(specs, sizes):
        out.append(np.frombuffer(buf, dt, count=nb // np.dtype(dt).itemsize,
                                 offset=off).reshape(shape))
        off += nb
    return out


class _Result:
    def __init__(self):
        self.ptr = C.c_void_p()

    def __del__(self):
        if self.ptr:
            lib.nl_result_free(self.ptr)
            self.ptr = C.c_void_p()

    def list(self, q):
        m = int(lib.nl_result_num_edges(self.ptr, q))
        ei, sh = _host_block([((2, m), np.int64), ((m, 3), np.int32)])
        if m:
            _check(lib.nl_result_copy(self.ptr, q, ei[0].ctypes.data_as(_lib.i64p),
                                      ei[1].ctypes.data_as(_lib.i64p),
                                      sh.ctypes.data_as(_lib.i32p)))
        return {"edge_index": ei, "shift": sh}

    def nlists(self):
        return int(lib.nl_result_num_lists(self.ptr))


def _cell_args(cell):
    if cell is None:
        return None, None
    if not isinstance(cell, PyCell):
        raise TypeError("argument 'cell': expected PyCell or None")
    pb = (C.c_uint8 * 3)(*[1 if v else 0 for v in cell._pbc])
    return cell._h.ctypes.data_as(_lib.dp), pb


def _cutoff_list(cutoffs):
    return [float(c) for c in cutoffs]


def _keyed(results, labels, n):
    out = {}
    for k in range(n):
        out[labels[k] if labels is not None else k] = results[k]
    return out


# ---------------------------------------------------------------------------
def build_neighborlists(cell, positions, cutoff, parallel=True):
    """src/python_api.rs:139-186."""
    cutoff = float(cutoff)
    if not np.isfinite(cutoff) or cutoff <= 0.0:
        raise ValueError("Cutoff must be positive and finite")
    pos = _positions(positions)
    n = pos.shape[0]
    if n == 0:
        return _empty()
    h, pb = _cell_args(cell)
    r = _Result()
    _check(lib.nl_search_single(pos.ctypes.data_as(_lib.dp), n, h, pb, cutoff, int(parallel),
                                C.byref(r.ptr)))
    _log("build_neighborlists", n)
    return r.list(0)


def build_neighborlists_multi(cell, positions, cutoffs, labels=None, disjoint=False):
    """src/python_api.rs:235-309."""
    pos = _positions(positions)
    n = pos.shape[0]
    cutoffs = _cutoff_list(cutoffs)
    if len(cutoffs) == 0:
        return {}
    if any((not np.isfinite(c)) or c <= 0.0 for c in cutoffs):
        raise ValueError("All cutoffs must be positive and finite")
    if labels is not None:
        labels = [str(l) for l in labels]
        if len(labels) != len(cutoffs):
            raise ValueError("Length of labels must match length of cutoffs")
    if n == 0:
        return _keyed([_empty() for _ in cutoffs], labels, len(cutoffs))
    h, pb = _cell_args(cell)
    c = np.ascontiguousarray(np.array(cutoffs, dtype=np.float64))
    r = _Result()
    _check(lib.nl_search_single_multi(pos.ctypes.data_as(_lib.dp), n, h, pb,
                                      c.ctypes.data_as(_lib.dp), c.size, int(bool(disjoint)),
                                      C.byref(r.ptr)))
    _log("build_neighborlists_multi", n)
    return _keyed([r.list(q) for q in range(len(cutoffs))], labels, len(cutoffs))


def _batch_cells(cells, n_systems):
    """Python batch cell convention (src/python_api.rs:374-407): (B, 3, 3)
    ASE rows; an all-zero matrix is an isolated system, any other matrix is
    transposed into H with pbc = (true, true, true)."""
    if cells is None:
        return 0, None, None, None
    cv = cells if isinstance(cells, np.ndarray) else np.asarray(cells)
    if cv.dtype != np.float64:
        raise TypeError("argument 'cells': array must be float64")
    if cv.ndim != 3:
        raise TypeError("argument 'cells': expected a 3-d array")
    if cv.shape[0] < n_systems:
        raise ValueError(f"Expected at least {n_systems} cells, but got {cv.shape[0]}")
    H = np.zeros((n_systems, 9))
    has = np.zeros(n_systems, np.uint8)
    pb = np.ones((n_systems, 3), np.uint8)
    for i in range(n_systems):
        m = np.asarray(cv[i], dtype=np.float64)
        if np.any(m != 0.0):
            H[i] = m.T.reshape(9)
            has[i] = 1
    return n_systems, H, has, pb


def _count_systems(b):
    """Monotonic-id check and system count (src/python_api.rs:360-372)."""
    if b.size == 0:
        return 0
    d = np.diff(b.astype(np.int64))
    if np.any(d < 0):
        raise ValueError("Batch IDs must be monotonic (non-decreasing)")
    return int(np.count_nonzero(d)) + 1


def build_neighborlists_batch(positions, batch, cells=None, cutoff=5.0, parallel=True):
    """src/python_api.rs:315-423."""
    cutoff = float(cutoff)
    if not np.isfinite(cutoff) or cutoff <= 0.0:
        raise ValueError("Cutoff must be positive and finite")
    pos = _positions(positions)
    b = _batch_ids(batch)
    n = pos.shape[0]
    if b.size != n:
        raise ValueError("positions and batch must have the same length")
    if n == 0:
        return _empty()
    ns = _count_systems(b)
    n_cells, H, has, pb = _batch_cells(cells, ns)
    r = _Result()
    _check(lib.nl_search_batch(
        pos.ctypes.data_as(_lib.dp), n, b.ctypes.data_as(_lib.i32p), n_cells,
        None if H is None else H.ctypes.data_as(_lib.dp),
        None if has is None else has.ctypes.data_as(_lib.u8p),
        None if pb is None else pb.ctypes.data_as(_lib.u8p), cutoff, int(parallel),
        C.byref(r.ptr)))
    _log("build_neighborlists_batch", n)
    return r.list(0)


def build_neighborlists_batch_multi(positions, batch, cells=None, cutoffs=[5.0], labels=None,
                                    disjoint=False):
    """src/python_api.rs:425-546."""
    pos = _positions(positions)
    b = _batch_ids(batch)
    n = pos.shape[0]
    if b.size != n:
        raise ValueError("positions and batch must have the same length")
    cutoffs = _cutoff_list(cutoffs)
    if n == 0 or len(cutoffs) == 0:
        return {}
    if any((not np.isfinite(c)) or c <= 0.0 for c in cutoffs):
        raise ValueError("All cutoffs must be positive and finite")
    if labels is not None:
        labels = [str(l) for l in labels]
        if len(labels) != len(cutoffs):
            raise ValueError("Length of labels must match length of cutoffs")
    ns = _count_systems(b)
    n_cells, H, has, pb = _batch_cells(cells, ns)
    c = np.ascontiguousarray(np.array(cutoffs, dtype=np.float64))
    r = _Result()
    _check(lib.nl_search_batch_multi(
        pos.ctypes.data_as(_lib.dp), n, b.ctypes.data_as(_lib.i32p), n_cells,
        None if H is None else H.ctypes.data_as(_lib.dp),
        None if has is None else has.ctypes.data_as(_lib.u8p),
        None if pb is None else pb.ctypes.data_as(_lib.u8p), c.ctypes.data_as(_lib.dp), c.size,
        int(bool(disjoint)), C.byref(r.ptr)))
    _log("build_neighborlists_batch_multi", n)
    return _keyed([r.list(q) for q in range(len(cutoffs))], labels, len(cutoffs))


# ---------------------------------------------------------------------------
def _ase_data(atoms):
    """extract_ase_data (src/python_api.rs:188-212)."""
    pos = np.asarray(atoms.get_positions(), dtype=np.float64)
    pbc = [bool(v) for v in np.asarray(atoms.get_pbc()).reshape(3)]
    cell = PyCell.from_ase(atoms) if any(pbc) else None
    return pos, cell


def build_from_ase(atoms, cutoff):
    """src/python_api.rs:214-222."""
    pos, cell = _ase_data(atoms)
    return build_neighborlists(cell, pos, cutoff, True)


def build_multi_from_ase(atoms, cutoffs, labels=None, disjoint=False):
    """src/python_api.rs:224-233."""
    pos, cell = _ase_data(atoms)
    return build_neighborlists_multi(cell, pos, cutoffs, labels, disjoint)


# ---------------------------------------------------------------------------
def get_num_threads():
    """The GPU engine has no host thread pool; this keeps the reference's
    accessor (src/python_api.rs:548-552) and reports the host CPU count
    (or the value fixed by the first set_num_threads call)."""
    return _num_threads if _num_threads is not None else (os.cpu_count() or 1)


def set_num_threads(n):
    """Only the first call takes effect, as with the reference's rayon
    global pool (src/python_api.rs:554-561)."""
    global _num_threads
    if _num_threads is None and int(n) > 0:
        _num_threads = int(n)


def init_logging(level=None):
    """At "debug"/"trace" level every call prints its hipEvent stage timings
    (build, count, fill) to stderr; other levels are silent.  RUST_LOG, when
    set, takes precedence as in the reference (src/python_api.rs:593-605)."""
    global _log_level
    env = os.environ.get("RUST_LOG")
    lvl = (env or level or "info").lower()
    _log_level = lvl
    lib.nl_set_timing(1 if lvl in ("debug", "trace") else 0)


def _log(name, n):
    if _log_level in ("debug", "trace"):
        import sys
        b, c, f, t = C.c_float(), C.c_float(), C.c_float(), C.c_float()
        e, nb = C.c_int64(), C.c_int64()
        lib.nl_last_timing(C.byref(b), C.byref(c), C.byref(f), C.byref(t), C.byref(e),
                           C.byref(nb))
        print(f"[neighborlist_rs] {name} n={n} edges={e.value} build={b.value:.3f}ms "
              f"count={c.value:.3f}ms fill={f.value:.3f}ms total={t.value:.3f}ms",
              file=sys.stderr)


def get_brute_force_threshold():
    return int(lib.nl_get_brute_force_threshold())


def set_brute_force_threshold(val):
    lib.nl_set_brute_force_threshold(int(val))


def get_parallel_threshold():
    return int(lib.nl_get_parallel_threshold())


def set_parallel_threshold(val):
    lib.nl_set_parallel_threshold(int(val))


def get_stack_threshold():
    return int(lib.nl_get_stack_threshold())


def set_stack_threshold(val):
    lib.nl_set_stack_threshold(int(val))


from .serving import Server  # noqa: E402  (continuous-batching server)
from .tensors import (build_neighborlists_torch, build_neighborlists_multi_torch,  # noqa: E402
                      build_neighborlists_batch_torch, build_neighborlists_batch_multi_torch)
from .md import DynamicNeighborManager  # noqa: E402  (skin lists for MD)
from .plans import NeighborPlan  # noqa: E402  (graph-captured searches)
