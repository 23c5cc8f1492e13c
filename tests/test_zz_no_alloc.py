"""nl_search_noalloc (no allocation) against nl_search_ex.

Two calls in a row (the second with a smaller input, so the shared workspace
holds the first call's contents when the second starts), each checked
against nl_search_ex on the same device positions: the same edges, bit for
bit after the canonical sort, and the same arrays in the same order (both
run the engine's pipeline with the same tunables).  Too small a capacity
writes nothing and reports the sizes; an oversized request is refused
before any device work.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import assert_same_edges

NL_DEVICE_POSITIONS = 1


def _lib():
    from neighborlist_rs import _lib
    return _lib


def _arrays(sizes, cells):
    S = len(sizes)
    H9 = np.zeros((S, 9))
    has = np.zeros(S, np.uint8)
    pbc = np.ones((S, 3), np.uint8)
    for s, H in enumerate(cells):
        if H is not None:
            H9[s] = np.asarray(H).reshape(9)
            has[s] = 1
    return H9, has, pbc


class Dev:
    def __init__(self, L, nbytes):
        self.L = L
        self.ptr = L.lib.nl_device_malloc(max(nbytes, 16))
        assert self.ptr
        self.nbytes = nbytes

    def put(self, host):
        assert host.nbytes <= self.nbytes
        assert self.L.lib.nl_memcpy_h2d(C.c_void_p(self.ptr), host.ctypes.data_as(C.c_void_p),
                                        host.nbytes) == 0

    def get(self, host, offset=0):
        assert offset + host.nbytes <= self.nbytes
        assert self.L.lib.nl_memcpy_d2h(host.ctypes.data_as(C.c_void_p),
                                        C.c_void_p(self.ptr + offset), host.nbytes) == 0
        return host

    def free(self):
        self.L.lib.nl_device_free(C.c_void_p(self.ptr))


def _search_ex(L, dpos, n, sizes, H9, has, pbc, cut, multi):
    r = C.c_void_p()
    rc = L.lib.nl_search_ex(C.c_void_p(dpos), n, sizes.ctypes.data_as(L.i64p), sizes.size,
                            H9.ctypes.data_as(L.dp), has.ctypes.data_as(L.u8p),
                            pbc.ctypes.data_as(L.u8p), cut.ctypes.data_as(L.dp), cut.size, 0,
                            int(multi), NL_DEVICE_POSITIONS, None, C.byref(r))
    assert rc == 0, L.last_error()
    out = []
    for q in range(cut.size):
        m = int(L.lib.nl_result_num_edges(r, q))
        ei, ej, s = np.empty(m, np.int64), np.empty(m, np.int64), np.empty((m, 3), np.int32)
        assert L.lib.nl_result_copy(r, q, ei.ctypes.data_as(L.i64p), ej.ctypes.data_as(L.i64p),
                                    s.ctypes.data_as(L.i32p)) == 0
        out.append((ei, ej, s))
    L.lib.nl_result_free(r)
    return out


def _noalloc(L, dpos, n, sizes, H9, has, pbc, cut, multi, edges, shifts, capacity):
    base = np.full(cut.size, -7, np.int64)
    length = np.full(cut.size, -7, np.int64)
    rc = L.lib.nl_search_noalloc(
        C.c_void_p(dpos), n, sizes.ctypes.data_as(L.i64p), sizes.size, H9.ctypes.data_as(L.dp),
        has.ctypes.data_as(L.u8p), pbc.ctypes.data_as(L.u8p), cut.ctypes.data_as(L.dp),
        cut.size, 0, int(multi), 0, C.c_void_p(edges.ptr), C.c_void_p(shifts.ptr), capacity,
        base.ctypes.data_as(L.i64p), length.ctypes.data_as(L.i64p), None)
    return rc, base, length


def _lists(edges, shifts, base, length):
    out = []
    for b, m in zip(base.tolist(), length.tolist()):
        ei = edges.get(np.empty(m, np.int64), 16 * b)
        ej = edges.get(np.empty(m, np.int64), 16 * b + 8 * m)
        s = shifts.get(np.empty((m, 3), np.int32), 12 * b)
        out.append((ei, ej, s))
    return out


@pytest.mark.gpu
def test_noalloc_matches_search_ex_twice():
    from neighborlist_rs import workloads as W
    L = _lib()
    cases = [
        # a ragged batch with three lists (multi rules), then a smaller one
        # with one list: the workspace still holds the first call's contents
        dict(n_systems=24, nmax=1500, seed=41, cut=[5.0, 6.0, 14.0], multi=True),
        dict(n_systems=6, nmax=400, seed=42, cut=[6.0], multi=False),
    ]
    cap = 0
    edges = shifts = None
    for k, c in enumerate(cases):
        p, sizes, cells, _ = W.c4_ragged_batch(n_systems=c["n_systems"], nmax=c["nmax"],
                                               seed=c["seed"])
        print(f"[case] noalloc call {k}: {sizes.size} systems, {p.shape[0]} atoms, "
              f"cutoffs {c['cut']}", flush=True)
        n = p.shape[0]
        H9, has, pbc = _arrays(sizes, cells)
        cut = np.array(c["cut"])
        dpos = Dev(L, p.nbytes)
        dpos.put(np.ascontiguousarray(p))
        want = _search_ex(L, dpos.ptr, n, sizes, H9, has, pbc, cut, c["multi"])
        total = sum(w[0].size for w in want)
        if k == 0:
            # buffers sized for the first (larger) call, reused by the second
            cap = total + 1000
            edges, shifts = Dev(L, 16 * cap), Dev(L, 12 * cap)
        rc, base, length = _noalloc(L, dpos.ptr, n, sizes, H9, has, pbc, cut, c["multi"],
                                    edges, shifts, cap)
        assert rc == 0, L.last_error()
        assert length.tolist() == [w[0].size for w in want]
        got = _lists(edges, shifts, base, length)
        for q in range(cut.size):
            assert_same_edges(got[q], want[q], f"call {k} list {q}")
            for a, b in zip(got[q], want[q]):
                np.testing.assert_array_equal(a, b)  # same pipeline: same order
        dpos.free()
    edges.free()
    shifts.free()


@pytest.mark.gpu
def test_noalloc_capacity_too_small_writes_nothing():
    from neighborlist_rs import workloads as W
    L = _lib()
    p, sizes, cells, _ = W.c4_ragged_batch(n_systems=5, nmax=800, seed=43)
    n = p.shape[0]
    H9, has, pbc = _arrays(sizes, cells)
    cut = np.array([6.0, 4.0])
    dpos = Dev(L, p.nbytes)
    dpos.put(np.ascontiguousarray(p))
    want = _search_ex(L, dpos.ptr, n, sizes, H9, has, pbc, cut, True)
    total = sum(w[0].size for w in want)
    cap = total - 1
    edges, shifts = Dev(L, 16 * cap), Dev(L, 12 * cap)
    sentinel = np.full(2 * cap, -5, np.int64)
    edges.put(sentinel)
    rc, base, length = _noalloc(L, dpos.ptr, n, sizes, H9, has, pbc, cut, True, edges, shifts,
                                cap)
    assert rc == 2
    assert "capacity" in L.last_error()
    assert length.tolist() == [w[0].size for w in want]
    np.testing.assert_array_equal(edges.get(np.empty(2 * cap, np.int64)), sentinel)
    # with room the same call succeeds
    edges2, shifts2 = Dev(L, 16 * total), Dev(L, 12 * total)
    rc, base, length = _noalloc(L, dpos.ptr, n, sizes, H9, has, pbc, cut, True, edges2, shifts2,
                                total)
    assert rc == 0, L.last_error()
    got = _lists(edges2, shifts2, base, length)
    for q in range(cut.size):
        assert_same_edges(got[q], want[q], f"list {q}")
    for d in (dpos, edges, shifts, edges2, shifts2):
        d.free()


def test_noalloc_refuses_an_oversized_request():
    """Refused from the sizes alone, before any device work (runs without a
    GPU too): 2,000,000 atoms need far more than the 64 MiB workspace."""
    L = _lib()
    n = 2_000_000
    sizes = np.array([n], np.int64)
    cut = np.array([5.0, 6.0, 14.0])
    base = np.zeros(3, np.int64)
    length = np.zeros(3, np.int64)
    rc = L.lib.nl_search_noalloc(C.c_void_p(16), n, sizes.ctypes.data_as(L.i64p), 1, None, None,
                                 None, cut.ctypes.data_as(L.dp), 3, 0, 1, 0, None, None, 0,
                                 base.ctypes.data_as(L.i64p), length.ctypes.data_as(L.i64p),
                                 None)
    assert rc == 1
    assert "workspace" in L.last_error() and "nl_search_ex" in L.last_error()
