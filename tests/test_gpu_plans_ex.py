"""Graph-captured plans with options (nl_plan_create_ex, nl_plan_array,
nl_plan_system_ranges, and the torch NeighborPlan).

Every replay, with positions moved in place between replays, must give:

* the oracle's lists (mirrored for full lists), bit-exact after the
  canonical sort;
* the eager search with the same options (nl_search_ex, itself checked
  against the oracle elsewhere): the same edges with the same distance and
  displacement per edge, bit for bit, and the same per-system ranges and
  modes;
* distances equal to sqrt((x^2 + y^2) + z^2) of the displacements exactly,
  and for full lists the mirror layout (slot 2e + 1 = slot 2e reversed, its
  displacement negated).

The capacity-overflow path (capacity 16) must grow every array it emits.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import assert_same_edges

pytestmark = pytest.mark.gpu

NL_DEVICE_POSITIONS, NL_RANGES, NL_DIST, NL_FULL, NL_DISP = 1, 4, 8, 16, 32


def _print_case(*a):
    print("[case]", *a, flush=True)


def _sorted(ei, ej, s, *extra):
    order = np.lexsort((s[:, 2], s[:, 1], s[:, 0], ej, ei))
    return [ei[order], ej[order], s[order]] + [None if x is None else x[order] for x in extra]


def _d2h(lib, ptr, shape, dtype):
    a = np.empty(shape, dtype)
    if a.size:
        assert lib.nl_memcpy_d2h(a.ctypes.data_as(C.c_void_p), ptr, a.nbytes) == 0
    return a


def _plan_lists(lib, plan, k, opts):
    out = []
    for q in range(k):
        m = int(lib.nl_plan_num_edges(plan, q))
        arrs = []
        for what, shape, dt, flag in ((0, (2, m), np.int64, 0), (1, (m, 3), np.int32, 0),
                                      (2, (m,), np.float64, NL_DIST),
                                      (3, (m, 3), np.float64, NL_DISP)):
            if flag and not opts & flag:
                arrs.append(None)
                continue
            p = C.c_void_p()
            assert lib.nl_plan_array(plan, q, what, C.byref(p)) == 0
            arrs.append(_d2h(lib, p, shape, dt))
        ei, sh, d, r = arrs
        out.append((ei[0], ei[1], sh, d, r))
    return out


def _eager(lib, L, dpos, n, sizes, H9, has, pb, cut, opts):
    r = C.c_void_p()
    assert lib.nl_search_ex(C.c_void_p(dpos), n, sizes.ctypes.data_as(L.i64p), sizes.size,
                            H9.ctypes.data_as(L.dp), has.ctypes.data_as(L.u8p),
                            pb.ctypes.data_as(L.u8p), cut.ctypes.data_as(L.dp), cut.size, 0, 1,
                            opts | NL_DEVICE_POSITIONS, None, C.byref(r)) == 0, L.last_error()
    out, ranges, modes = [], None, None
    try:
        for q in range(cut.size):
            m = int(lib.nl_result_num_edges(r, q))
            ei, ej = np.empty(m, np.int64), np.empty(m, np.int64)
            sh = np.empty((m, 3), np.int32)
            d = np.empty(m) if opts & NL_DIST else None
            v = np.empty((m, 3)) if opts & NL_DISP else None
            if m:
                assert lib.nl_result_copy(r, q, ei.ctypes.data_as(L.i64p),
                                          ej.ctypes.data_as(L.i64p),
                                          sh.ctypes.data_as(L.i32p)) == 0
                if d is not None:
                    assert lib.nl_result_copy_dist(r, q, d.ctypes.data_as(L.dp)) == 0
                if v is not None:
                    assert lib.nl_result_copy_disp(r, q, v.ctypes.data_as(L.dp)) == 0
            out.append((ei, ej, sh, d, v))
        if opts & NL_RANGES:
            ranges = np.zeros((sizes.size, cut.size, 2), np.int64)
            modes = np.zeros(sizes.size, np.int32)
            assert lib.nl_result_system_ranges(r, ranges.ctypes.data_as(L.i64p),
                                               modes.ctypes.data_as(L.i32p)) == 0
    finally:
        lib.nl_result_free(r)
    return out, ranges, modes


def _mirror(ei, ej, s):
    keep = ei != ej
    return (np.concatenate([ei, ej[keep]]), np.concatenate([ej, ei[keep]]),
            np.concatenate([s, -s[keep]]))


def _check_geometry(t, full, what):
    ei, ej, sh, d, v = t
    if d is not None and v is not None:
        np.testing.assert_array_equal(d, np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) +
                                                 v[:, 2] * v[:, 2]), err_msg=what)
    if full:
        assert ei.size % 2 == 0, what
        np.testing.assert_array_equal(ei[0::2], ej[1::2], err_msg=what)
        np.testing.assert_array_equal(ej[0::2], ei[1::2], err_msg=what)
        np.testing.assert_array_equal(sh[0::2], -sh[1::2], err_msg=what)
        if d is not None:
            np.testing.assert_array_equal(d[0::2], d[1::2], err_msg=what)
        if v is not None:
            np.testing.assert_array_equal(v[0::2], -v[1::2], err_msg=what)


@pytest.mark.parametrize("opts,capacity", [(NL_FULL | NL_DIST | NL_DISP | NL_RANGES, 0),
                                           (NL_DIST | NL_DISP | NL_RANGES, 16),
                                           (NL_FULL | NL_DISP, 16)])
def test_plan_options_replays(nr, oracle, opts, capacity):
    from neighborlist_rs import _lib as L, workloads as W
    lib = L.lib
    pos, sizes, cells, cut = W.c4_ragged_batch(n_systems=24, nmax=1200, seed=33)
    sizes = np.ascontiguousarray(sizes, dtype=np.int64)
    cut = np.array(cut)
    S = sizes.size
    H9 = np.zeros((S, 9))
    has = np.zeros(S, np.uint8)
    pb = np.ones((S, 3), np.uint8)
    for s, H in enumerate(cells):
        if H is not None:
            H9[s] = np.asarray(H).reshape(9)
            has[s] = 1
    batch = np.repeat(np.arange(S, dtype=np.int32), sizes)
    ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
    full = bool(opts & NL_FULL)
    n = pos.shape[0]
    dpos = lib.nl_device_malloc(pos.nbytes)
    assert dpos
    plan = C.c_void_p()
    _print_case("plan options", opts, "capacity", capacity)
    try:
        cur = np.ascontiguousarray(pos.copy())
        assert lib.nl_memcpy_h2d(C.c_void_p(dpos), cur.ctypes.data_as(C.c_void_p),
                                 cur.nbytes) == 0
        assert lib.nl_plan_create_ex(C.c_void_p(dpos), n, sizes.ctypes.data_as(L.i64p), S,
                                     H9.ctypes.data_as(L.dp), has.ctypes.data_as(L.u8p),
                                     pb.ctypes.data_as(L.u8p), cut.ctypes.data_as(L.dp),
                                     cut.size, 0, 1, opts, capacity,
                                     C.byref(plan)) == 0, L.last_error()
        if not opts & NL_RANGES:
            assert lib.nl_plan_system_ranges(plan, None, None) != 0
            assert "NL_RANGES" in L.last_error()
        if not opts & NL_DIST:
            p = C.c_void_p()
            assert lib.nl_plan_array(plan, 0, 2, C.byref(p)) != 0
            assert "NL_DIST" in L.last_error()
        rng = np.random.default_rng(9)
        for rep in range(3):
            if rep:
                cur = np.ascontiguousarray(cur + rng.normal(scale=0.05, size=cur.shape))
                assert lib.nl_memcpy_h2d(C.c_void_p(dpos), cur.ctypes.data_as(C.c_void_p),
                                         cur.nbytes) == 0
            assert lib.nl_plan_run(plan) == 0, L.last_error()
            got = _plan_lists(lib, plan, cut.size, opts)
            want, ranges, modes = _eager(lib, L, dpos, n, sizes, H9, has, pb, cut, opts)
            ref = oracle.search_batch_multi(cur, batch, ocells, cut.tolist(), False)
            for q in range(cut.size):
                what = f"replay {rep} list {q}"
                g, w = got[q], want[q]
                _check_geometry(g, full, what)
                o = _mirror(*ref[q]) if full else ref[q]
                assert_same_edges(g[:3], o, what + " vs oracle")
                gs, ws = _sorted(*g), _sorted(*w)
                for a, b in zip(gs, ws):
                    if a is None:
                        assert b is None
                    else:
                        np.testing.assert_array_equal(a, b, err_msg=what + " vs eager")
            if opts & NL_RANGES:
                r = np.zeros((S, cut.size, 2), np.int64)
                m = np.zeros(S, np.int32)
                assert lib.nl_plan_system_ranges(plan, r.ctypes.data_as(L.i64p),
                                                 m.ctypes.data_as(L.i32p)) == 0
                np.testing.assert_array_equal(r, ranges)
                np.testing.assert_array_equal(m, modes)
                # ranges tile each list in system order
                for q in range(cut.size):
                    assert r[0, q, 0] == 0
                    np.testing.assert_array_equal(r[1:, q, 0], r[:-1, q, 1])
                    assert r[-1, q, 1] == got[q][0].size
        rp, rc, cp = C.c_int64(), C.c_int64(), C.c_int64()
        assert lib.nl_plan_stats(plan, C.byref(rp), C.byref(rc), C.byref(cp)) == 0
        assert rp.value >= 3
        if capacity == 16:
            assert rc.value >= 2 and cp.value > 16  # grew and re-captured
    finally:
        lib.nl_plan_free(plan)
        lib.nl_device_free(C.c_void_p(dpos))


def _t(x):
    return x.detach().cpu().numpy()


def _torch_tuple(d):
    return (_t(d["edge_index"][0]), _t(d["edge_index"][1]), _t(d["shift"]),
            _t(d["distance"]) if "distance" in d else None,
            _t(d["displacement"]) if "displacement" in d else None)


def test_neighbor_plan_torch(nr, oracle):
    import torch
    from neighborlist_rs import workloads as W
    rng = np.random.default_rng(17)
    Hc = np.array([[17.0, 0.0, 0.0], [1.5, 16.0, 0.0], [0.5, 1.0, 18.0]]).T  # columns
    pos = rng.uniform(0, 1, (700, 3)) @ Hc.T
    cell = nr.PyCell(Hc.tolist(), [True] * 3)
    pt = torch.from_numpy(pos).cuda()
    _print_case("torch plan single")
    with nr.NeighborPlan(pt, 5.0, cell, full=True, distances=True, displacements=True) as p:
        for rep in range(3):
            if rep:
                pos = pos + rng.normal(0.0, 0.05, pos.shape)
                pt.copy_(torch.from_numpy(pos))
            g = p.run()
            assert g["edge_index"].device == pt.device
            want = nr.build_neighborlists_torch(cell, pt, 5.0, full=True, distances=True,
                                                displacements=True)
            gt, wt = _torch_tuple(g), _torch_tuple(want)
            _check_geometry(gt, True, f"torch plan rep {rep}")
            o = _mirror(*oracle.search_single(pos, Hc, [True] * 3, 5.0))
            assert_same_edges(gt[:3], o, f"torch plan rep {rep} vs oracle")
            for a, b in zip(_sorted(*gt), _sorted(*wt)):
                np.testing.assert_array_equal(a, b, err_msg=f"torch plan rep {rep} vs search")
        assert p.stats()["replays"] >= 3

    # ragged batch, two labelled lists, per-system ranges
    bpos, sizes, cells, _ = W.c4_ragged_batch(n_systems=10, nmax=500, seed=44)
    batch = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)
    rows = W.batch_arrays(sizes, cells)[1]
    bt = torch.from_numpy(np.ascontiguousarray(bpos)).cuda()
    bb = torch.from_numpy(batch).cuda()
    _print_case("torch plan batch")
    with nr.NeighborPlan(bt, [4.0, 6.0], batch=bb, cells=rows, labels=["a", "b"],
                         distances=True, ranges=True) as p:
        g = p.run()
        want = nr.build_neighborlists_batch_multi_torch(bt, bb, rows, [4.0, 6.0],
                                                        labels=["a", "b"], distances=True)
        ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
        ref = oracle.search_batch_multi(bpos, batch, ocells, [4.0, 6.0], False)
        for q, key in enumerate(("a", "b")):
            gt, wt = _torch_tuple(g[key]), _torch_tuple(want[key])
            assert_same_edges(gt[:3], ref[q], f"torch plan batch list {key} vs oracle")
            for a, b in zip(_sorted(*gt[:4]), _sorted(*wt[:4])):
                np.testing.assert_array_equal(a, b, err_msg=f"batch list {key} vs search")
        r, m = p.system_ranges()
        assert r.shape == (len(sizes), 2, 2) and m.shape == (len(sizes),)
        assert int(r[-1, 1, 1]) == g["b"]["edge_index"].shape[1]
        np.testing.assert_array_equal(r[1:, 0, 0], r[:-1, 0, 1])
    with pytest.raises(ValueError, match="closed"):
        p.run()
