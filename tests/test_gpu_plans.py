"""Graph-captured steady state (nl_plan_*): every replay must equal the
oracle, with eager engine work and position updates between replays, and
the capacity-overflow path must grow, re-capture and still be exact."""
import ctypes as C

import numpy as np
import pytest

from conftest import assert_same_edges

pytestmark = pytest.mark.gpu


def _dev(lib, a):
    a = np.ascontiguousarray(a)
    ptr = lib.nl_device_malloc(a.nbytes)
    assert ptr
    assert lib.nl_memcpy_h2d(C.c_void_p(ptr), a.ctypes.data_as(C.c_void_p), a.nbytes) == 0
    return ptr


def _plan_lists(L, plan, k):
    out = []
    for q in range(k):
        m = L.lib.nl_plan_num_edges(plan, q)
        ei = np.empty(m, np.int64)
        ej = np.empty(m, np.int64)
        s = np.empty((m, 3), np.int32)
        assert L.lib.nl_plan_copy(plan, q, ei.ctypes.data_as(L.i64p), ej.ctypes.data_as(L.i64p),
                                  s.ctypes.data_as(L.i32p)) == 0
        out.append((ei, ej, s))
    return out


def _cells(cells):
    S = len(cells)
    H9 = np.zeros((S, 9))
    has = np.zeros(S, np.uint8)
    pbc = np.ones((S, 3), np.uint8)
    for s, H in enumerate(cells):
        if H is not None:
            H9[s] = np.asarray(H).reshape(9)
            has[s] = 1
    return H9, has, pbc


@pytest.mark.parametrize("capacity", [0, 16])
def test_plan_replays_match_oracle(nr, oracle, capacity):
    from neighborlist_rs import _lib as L, workloads as W
    lib = L.lib
    pos, sizes, cells, cut = W.c4_ragged_batch(n_systems=32, nmax=1500, seed=21)
    cut = np.array(cut)
    H9, has, pbc = _cells(cells)
    batch = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)
    ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
    dpos = _dev(lib, pos)
    plan = C.c_void_p()
    print("[case] plan capacity", capacity, flush=True)
    assert lib.nl_plan_create(C.c_void_p(dpos), pos.shape[0], sizes.ctypes.data_as(L.i64p),
                              len(sizes), H9.ctypes.data_as(L.dp), has.ctypes.data_as(L.u8p),
                              pbc.ctypes.data_as(L.u8p), cut.ctypes.data_as(L.dp), cut.size, 0,
                              1, 1, capacity, C.byref(plan)) == 0, L.last_error()
    try:
        rng = np.random.default_rng(5)
        cur = pos.copy()
        for rep in range(4):
            if rep >= 2:  # molecular-dynamics-like step: move atoms in place
                cur = cur + rng.normal(scale=0.05, size=cur.shape)
                assert lib.nl_memcpy_h2d(C.c_void_p(dpos), cur.ctypes.data_as(C.c_void_p),
                                         cur.nbytes) == 0
            assert lib.nl_plan_run(plan) == 0, L.last_error()
            got = _plan_lists(L, plan, cut.size)
            want = oracle.search_batch_multi(cur, batch, ocells, cut.tolist(), False)
            for q in range(cut.size):
                assert_same_edges(got[q], want[q], f"replay {rep} list {q}")
            # eager engine work between replays
            eager = nr.build_neighborlists_batch_multi(cur, batch, W.batch_arrays(sizes, cells)[1],
                                                       list(cut))
            for q in range(cut.size):
                e = eager[q]
                assert_same_edges((e["edge_index"][0], e["edge_index"][1], e["shift"]), want[q])
        rp, rcap, cp = C.c_int64(), C.c_int64(), C.c_int64()
        lib.nl_plan_stats(plan, C.byref(rp), C.byref(rcap), C.byref(cp))
        assert rp.value >= 4
        if capacity == 16:
            assert rcap.value >= 2  # grew and re-captured
    finally:
        lib.nl_plan_free(plan)
        lib.nl_device_free(C.c_void_p(dpos))
