"""Lists delivered to numpy arrays through page-locked blocks (nl_host_alloc /
nl_host_free and the Python results that use them).

Large lists land in one pinned block per list (edge rows and shifts back to
back); the arrays must equal those copied into ordinary memory, stay valid
after the result, the dict and later searches are gone or reuse the pool,
and freed blocks must be handed out again."""
import ctypes as C
import gc

import numpy as np
import pytest

from conftest import as_tuple, assert_same_edges
from xyzio import read_xyz

pytestmark = pytest.mark.gpu


def _print_case(*a):
    print("[case]", *a, flush=True)


def _pinned(a):
    """True when the array views a pool block (its base is the ctypes buffer)."""
    b = a
    while isinstance(b, np.ndarray) and b.base is not None:
        b = b.base
    return hasattr(b, "_owner")


def test_pool_reuses_blocks(nr):
    from neighborlist_rs import _lib
    lib = _lib.lib
    _print_case("pinned pool")
    p = lib.nl_host_alloc(3 << 20)
    assert p
    lib.nl_host_free(C.c_void_p(p))
    q = lib.nl_host_alloc((3 << 20) + 12345)  # same 4 MiB class
    assert q == p
    r = lib.nl_host_alloc(3 << 20)            # the block is taken: a new one
    assert r and r != q
    lib.nl_host_free(C.c_void_p(q))
    lib.nl_host_free(C.c_void_p(r))
    assert not lib.nl_host_alloc(0)


def test_pinned_lists_equal_pageable_and_outlive(nr, oracle, monkeypatch):
    import neighborlist_rs as pkg
    pos, rows, pbc = read_xyz("ethanol_1000.xyz")
    cell = nr.PyCell(rows.T.tolist(), pbc)
    _print_case("pinned delivery ethanol_1000 rc 14")
    big = nr.build_neighborlists(cell, pos, 14.0)
    assert big["edge_index"].shape[1] > 100000
    assert _pinned(big["edge_index"]) and _pinned(big["shift"])
    assert big["edge_index"].flags.writeable and big["edge_index"].flags.c_contiguous
    keep_ei, keep_sh = big["edge_index"], big["shift"]
    ref = (keep_ei.copy(), keep_sh.copy())
    del big
    gc.collect()
    # more searches, some of which reuse freed blocks
    for rc in (14.0, 12.0, 14.0):
        r = nr.build_neighborlists(cell, pos, rc)
        del r
        gc.collect()
    np.testing.assert_array_equal(keep_ei, ref[0])
    np.testing.assert_array_equal(keep_sh, ref[1])
    # the same lists through ordinary memory
    monkeypatch.setattr(pkg, "_PINNED_MIN", 1 << 62)
    plain = nr.build_neighborlists(cell, pos, 14.0)
    assert not _pinned(plain["edge_index"])
    np.testing.assert_array_equal(plain["edge_index"], ref[0])
    np.testing.assert_array_equal(plain["shift"], ref[1])
    assert_same_edges(as_tuple(plain), oracle.search_single(pos, rows.T, pbc, 14.0),
                      "ethanol_1000 rc 14")


def test_pinned_multi_and_serving(nr, oracle):
    pos, rows, pbc = read_xyz("ethanol_1000.xyz")
    cell = nr.PyCell(rows.T.tolist(), pbc)
    _print_case("pinned multi + serving")
    res = nr.build_neighborlists_multi(cell, pos, [6.0, 14.0])
    want = oracle.search_single_multi(pos, rows.T, pbc, [6.0, 14.0])
    for q in range(2):
        assert_same_edges(as_tuple(res[q]), want[q], f"multi list {q}")
    assert _pinned(res[1]["edge_index"])
    with nr.Server(slots=1, distances=True, displacements=True) as srv:
        r = srv.result(srv.submit(cell, pos, 14.0), 60.0)
    assert _pinned(r["edge_index"]) and _pinned(r["distance"]) and _pinned(r["shift"])
    for k in ("distance", "displacement"):
        assert r[k].flags.aligned
    assert_same_edges(as_tuple(r), want[1], "served rc 14")
    v = r["displacement"]
    np.testing.assert_array_equal(r["distance"], np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1])
                                                         + v[:, 2] * v[:, 2]))
