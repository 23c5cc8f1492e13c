"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  Results come back in the reference's EdgeResult shape
(search.rs:14): per cutoff a tuple (edge_i int64[M], edge_j int64[M],
shifts int32[M, 3]).
"""
import ctypes as C
import os

import numpy as np

from . import build_oracle

_LIB = None


class _List(C.Structure):
    _fields_ = [("n", C.c_int64), ("ei", C.POINTER(C.c_int64)),
                ("ej", C.POINTER(C.c_int64)), ("s", C.POINTER(C.c_int32))]


class _Result(C.Structure):
    _fields_ = [("k", C.c_int64), ("lists", C.POINTER(_List))]


def lib():
    global _LIB
    if _LIB is None:
        path = build_oracle.OUT
        if not os.path.exists(path):
            build_oracle.build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.nlo_search_single.argtypes = [dp, C.c_int64, dp, ip, C.c_double, C.c_int,
                                        C.POINTER(_Result)]
        L.nlo_search_single_multi.argtypes = [dp, C.c_int64, dp, ip, dp, C.c_int, C.c_int,
                                              C.POINTER(_Result)]
        L.nlo_search_batch.argtypes = [dp, C.c_int64, C.POINTER(C.c_int32), C.c_int64, dp, ip,
                                       ip, C.c_double, C.c_int, C.POINTER(_Result)]
        L.nlo_search_batch_multi.argtypes = [dp, C.c_int64, C.POINTER(C.c_int32), C.c_int64, dp,
                                             ip, ip, dp, C.c_int, C.c_int, C.POINTER(_Result)]
        for f in ("nlo_search_single", "nlo_search_single_multi", "nlo_search_batch",
                  "nlo_search_batch_multi"):
            getattr(L, f).restype = C.c_int
        L.nlo_result_free.argtypes = [C.POINTER(_Result)]
        L.nlo_last_error.restype = C.c_char_p
        L.nlo_cell_inverse.argtypes = [dp, dp]
        L.nlo_cell_inverse.restype = C.c_int
        L.nlo_perpendicular_widths.argtypes = [dp, dp]
        L.nlo_cell_wrap.argtypes = [dp, ip, dp, dp]
        L.nlo_cell_wrap.restype = C.c_int
        L.nlo_cell_mic.argtypes = [dp, ip, dp, dp, C.POINTER(C.c_int32), dp]
        L.nlo_cell_mic.restype = C.c_int
        L.nlo_z_order.argtypes = [dp]
        L.nlo_z_order.restype = C.c_uint64
        L.nlo_celllist_bin_atoms.argtypes = [dp, C.c_int64, dp, ip, C.c_double, C.c_int64,
                                             C.c_int64, C.c_int64, C.POINTER(C.c_int64),
                                             C.POINTER(C.c_int64), C.c_int64]
        L.nlo_celllist_bin_atoms.restype = C.c_int64
        for g in ("brute_force", "parallel", "stack"):
            getattr(L, f"nlo_get_{g}_threshold").restype = C.c_size_t
            getattr(L, f"nlo_set_{g}_threshold").argtypes = [C.c_size_t]
        L.nlo_max_threads.restype = C.c_int
        L.nlo_set_threads.argtypes = [C.c_int]
        _LIB = L
    return _LIB


class OracleError(ValueError):
    pass


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _pos(positions):
    p = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 3))
    return p


def _cell(h, pbc):
    if h is None:
        return None, None
    H = np.ascontiguousarray(np.asarray(h, dtype=np.float64).reshape(3, 3))
    P = np.ascontiguousarray(np.asarray([True] * 3 if pbc is None else pbc,
                                        dtype=np.int32).reshape(3))
    return H, P


def _take(res):
    out = []
    for k in range(res.k):
        l = res.lists[k]
        n = int(l.n)
        if n:
            ei = np.ctypeslib.as_array(l.ei, shape=(n,)).copy()
            ej = np.ctypeslib.as_array(l.ej, shape=(n,)).copy()
            s = np.ctypeslib.as_array(l.s, shape=(3 * n,)).copy().reshape(n, 3)
        else:
            ei = np.zeros(0, np.int64)
            ej = np.zeros(0, np.int64)
            s = np.zeros((0, 3), np.int32)
        out.append((ei, ej, s))
    lib().nlo_result_free(C.byref(res))
    return out


def _check(rc):
    if rc != 0:
        raise OracleError(lib().nlo_last_error().decode())


def search_single(positions, h=None, pbc=None, cutoff=5.0, parallel=True):
    """single::search_single (single.rs:8-90); h is H (lattice vectors as columns)."""
    p = _pos(positions)
    H, P = _cell(h, pbc)
    res = _Result()
    _check(lib().nlo_search_single(_dp(p), p.shape[0], None if H is None else _dp(H),
                                   None if P is None else _ip(P), float(cutoff), int(parallel),
                                   C.byref(res)))
    return _take(res)[0]


def search_single_multi(positions, h=None, pbc=None, cutoffs=(5.0,), disjoint=False):
    """single::search_single_multi (single.rs:92-164)."""
    p = _pos(positions)
    H, P = _cell(h, pbc)
    c = np.ascontiguousarray(np.asarray(cutoffs, dtype=np.float64).reshape(-1))
    res = _Result()
    _check(lib().nlo_search_single_multi(_dp(p), p.shape[0], None if H is None else _dp(H),
                                         None if P is None else _ip(P), _dp(c), c.size,
                                         int(disjoint), C.byref(res)))
    return _take(res)


def _cells(cells, n_sys_hint):
    """cells: None or list of (H, pbc) / None per system."""
    if cells is None:
        return None, None, None, 0
    n = len(cells)
    H = np.zeros((max(n, 1), 9))
    has = np.zeros(max(n, 1), np.int32)
    P = np.zeros((max(n, 1), 3), np.int32)
    for q, c in enumerate(cells):
        if c is not None:
            H[q] = np.asarray(c[0], dtype=np.float64).reshape(9)
            P[q] = np.asarray(c[1], dtype=np.int32).reshape(3)
            has[q] = 1
    return H, has, P, n


def search_batch(positions, batch, cells=None, cutoff=5.0, parallel=True):
    """batch::search_batch (batch.rs:10-147); cells: list of (H, pbc) or None per system."""
    p = _pos(positions)
    b = np.ascontiguousarray(np.asarray(batch, dtype=np.int32).reshape(-1))
    if b.size != p.shape[0]:
        raise OracleError("positions and batch must have the same length")
    H, has, P, n = _cells(cells, None)
    res = _Result()
    _check(lib().nlo_search_batch(_dp(p), p.shape[0], b.ctypes.data_as(C.POINTER(C.c_int32)),
                                  n, None if H is None else _dp(H),
                                  None if has is None else _ip(has),
                                  None if P is None else _ip(P), float(cutoff), int(parallel),
                                  C.byref(res)))
    return _take(res)[0]


def search_batch_multi(positions, batch, cells=None, cutoffs=(5.0,), disjoint=False):
    """batch::search_batch_multi (batch.rs:149-290)."""
    p = _pos(positions)
    b = np.ascontiguousarray(np.asarray(batch, dtype=np.int32).reshape(-1))
    if b.size != p.shape[0]:
        raise OracleError("positions and batch must have the same length")
    H, has, P, n = _cells(cells, None)
    c = np.ascontiguousarray(np.asarray(cutoffs, dtype=np.float64).reshape(-1))
    res = _Result()
    _check(lib().nlo_search_batch_multi(_dp(p), p.shape[0],
                                        b.ctypes.data_as(C.POINTER(C.c_int32)), n,
                                        None if H is None else _dp(H),
                                        None if has is None else _ip(has),
                                        None if P is None else _ip(P), _dp(c), c.size,
                                        int(disjoint), C.byref(res)))
    return _take(res)


def cell_inverse(h):
    H = np.ascontiguousarray(np.asarray(h, dtype=np.float64).reshape(9))
    out = np.zeros(9)
    if lib().nlo_cell_inverse(_dp(H), _dp(out)) != 0:
        raise OracleError("Cell matrix is not invertible")
    return out.reshape(3, 3)


def perpendicular_widths(h):
    inv = np.ascontiguousarray(cell_inverse(h).reshape(9))
    w = np.zeros(3)
    lib().nlo_perpendicular_widths(_dp(inv), _dp(w))
    return w


def cell_wrap(h, pbc, pos):
    H, P = _cell(h, pbc)
    p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(3))
    out = np.zeros(3)
    _check(lib().nlo_cell_wrap(_dp(H), _ip(P), _dp(p), _dp(out)))
    return out


def cell_mic(h, pbc, ri, rj):
    H, P = _cell(h, pbc)
    a = np.ascontiguousarray(np.asarray(ri, dtype=np.float64).reshape(3))
    b = np.ascontiguousarray(np.asarray(rj, dtype=np.float64).reshape(3))
    s = np.zeros(3, np.int32)
    d = np.zeros(3)
    _check(lib().nlo_cell_mic(_dp(H), _ip(P), _dp(a), _dp(b),
                              s.ctypes.data_as(C.POINTER(C.c_int32)), _dp(d)))
    return s, d


def z_order(frac):
    f = np.ascontiguousarray(np.asarray(frac, dtype=np.float64).reshape(3))
    return int(lib().nlo_z_order(_dp(f)))


def celllist_bin_atoms(positions, h, pbc, cutoff, bx, by, bz):
    p = _pos(positions)
    H, P = _cell(h, pbc)
    nb = np.zeros(3, np.int64)
    out = np.zeros(max(1, p.shape[0]), np.int64)
    cnt = lib().nlo_celllist_bin_atoms(_dp(p), p.shape[0], _dp(H), _ip(P), float(cutoff), bx,
                                       by, bz, nb.ctypes.data_as(C.POINTER(C.c_int64)),
                                       out.ctypes.data_as(C.POINTER(C.c_int64)), out.size)
    if cnt < 0:
        raise OracleError("Cell matrix is not invertible")
    return tuple(int(x) for x in nb), sorted(out[:cnt].tolist())


def set_threshold(name, val):
    getattr(lib(), f"nlo_set_{name}_threshold")(int(val))


def get_threshold(name):
    return int(getattr(lib(), f"nlo_get_{name}_threshold")())


def max_threads():
    return int(lib().nlo_max_threads())


def set_threads(n):
    lib().nlo_set_threads(int(n))
