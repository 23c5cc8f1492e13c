"""GPU serving: the continuous-batching server against the plain path and the oracle.

Each request of a Poisson trace (random isolated, orthorhombic and triclinic
systems, each with its own cutoff) must give the oracle's edges bit-exactly
after the canonical sort, and exactly the same arrays (order included)
whether it ran alone or batched with other requests.
"""
import threading

import numpy as np
import pytest

from conftest import as_tuple, assert_same_edge_sets, assert_same_edges

pytestmark = pytest.mark.gpu


def _print_case(*a):
    print("[case]", *a, flush=True)


def _cell(nr, H):
    return None if H is None else nr.PyCell(np.asarray(H).tolist(), [True, True, True])


def _trace(seed=11, n=40, nmax=3000):
    from neighborlist_rs.workloads import c5_serving_trace
    return c5_serving_trace(seed=seed, n_requests=n, nmin=8, nmax=nmax)


def _serve(nr, reqs, copy=True, **kw):
    with nr.Server(**kw) as srv:
        ids = [srv.submit(_cell(nr, H), p, rc, copy=copy) for (_, p, H, rc) in reqs]
        info = []
        out = []
        for rid in ids:
            assert srv.wait(rid, timeout=60.0)
            info.append(srv.info(rid))
            out.append(srv.result(rid))
        stats = srv.stats()
    return out, info, stats


def test_serving_mix_equals_alone_and_oracle(nr, oracle):
    reqs = _trace()
    _print_case("serving alone", len(reqs))
    alone, info_a, st_a = _serve(nr, reqs, slots=1, max_systems=1)
    assert st_a["max_batch"] == 1 and st_a["completed"] == len(reqs)
    assert all(b == 1 for (_, _, b) in info_a)
    _print_case("serving mix")
    mix, info_m, st_m = _serve(nr, reqs, slots=3, max_atoms=20000, max_systems=64)
    assert st_m["completed"] == len(reqs)
    assert st_m["max_batch"] > 1, st_m
    for k, ((_, p, H, rc), a, m) in enumerate(zip(reqs, alone, mix)):
        what = f"request {k}: n={p.shape[0]} rc={rc:.3f} cell={H is not None}"
        np.testing.assert_array_equal(a["edge_index"], m["edge_index"], err_msg=what)
        np.testing.assert_array_equal(a["shift"], m["shift"], err_msg=what)
        want = oracle.search_single(p, H, None if H is None else [True] * 3, rc)
        assert_same_edges(as_tuple(m), want, what)
        assert info_m[k][0] == m["shift"].shape[0]
        plain = as_tuple(nr.build_neighborlists(_cell(nr, H), p, rc))
        assert_same_edges(as_tuple(m), plain, what + " vs plain path")


def test_serving_full_size_trace(nr, oracle):
    """The serving trace at its stated size: N log-uniform in [16, 50000],
    rc uniform in [4, 12], a third each isolated / orthorhombic / triclinic.
    Seven of the sixteen requests exceed 4096 atoms, so their batches take the
    global radix sort, and they share batches with smaller requests, each
    with its own cutoff.  Every request: the same arrays (order included)
    served alone and in the mix, and the oracle's edge set."""
    from neighborlist_rs.workloads import c5_serving_trace
    reqs = c5_serving_trace(seed=31, n_requests=16)
    sizes = [p.shape[0] for (_, p, _, _) in reqs]
    assert sum(n > 4096 for n in sizes) == 7 and max(sizes) > 30000
    _print_case("serving full size alone", sizes)
    alone, _, st_a = _serve(nr, reqs, slots=1, max_systems=1)
    assert st_a["max_batch"] == 1
    _print_case("serving full size mix")
    mix, info, st = _serve(nr, reqs, slots=3)
    assert st["completed"] == len(reqs)
    # a batch holding a request over 4096 atoms together with other requests
    assert any(b > 1 and n > 4096 for n, (_, _, b) in zip(sizes, info)), info
    for k, ((_, p, H, rc), a, m) in enumerate(zip(reqs, alone, mix)):
        what = f"request {k}: n={p.shape[0]} rc={rc:.3f} cell={H is not None}"
        _print_case(what)
        np.testing.assert_array_equal(a["edge_index"], m["edge_index"], err_msg=what)
        np.testing.assert_array_equal(a["shift"], m["shift"], err_msg=what)
        want = oracle.search_single(p, H, None if H is None else [True] * 3, rc)
        assert_same_edge_sets(as_tuple(m), want, what)
        assert info[k][0] == m["shift"].shape[0]


def test_serving_threads_and_budget(nr, oracle):
    reqs = _trace(seed=12, n=32, nmax=2000)
    want = [oracle.search_single(p, H, None if H is None else [True] * 3, rc)
            for (_, p, H, rc) in reqs]
    _print_case("serving threads, tight budget")
    # a budget of one edge: every batch holds one request and only one batch is
    # on the device at a time (a request alone is always admitted)
    with nr.Server(slots=3, edge_budget=1) as srv:
        got = [None] * len(reqs)

        def client(lo):
            for k in range(lo, len(reqs), 4):
                _, p, H, rc = reqs[k]
                got[k] = srv.result(srv.submit(_cell(nr, H), p, rc), timeout=60.0)

        ts = [threading.Thread(target=client, args=(i,)) for i in range(4)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["max_batch"] == 1 and st["max_inflight"] == 1, st
    for k in range(len(reqs)):
        assert_same_edges(as_tuple(got[k]), want[k], f"request {k}")


def test_serving_refill_when_estimate_is_short(nr, oracle):
    # a dense cluster plus one far atom: the edge estimate from the bounding
    # box is far too small, and the cluster's ~2.3 M edges overflow the
    # smallest output block (2^20 edges), so the fill pass must run again
    # into a larger block
    rng = np.random.default_rng(3)
    p = np.vstack([rng.uniform(0.0, 6.0, size=(2500, 3)), [[900.0, 900.0, 900.0]]])
    _print_case("serving refill", p.shape)
    with nr.Server(slots=2) as srv:
        a = srv.submit(None, p, 5.0)
        b = srv.submit(None, p[:100], 5.0)
        ra, rb = srv.result(a, 60.0), srv.result(b, 60.0)
        st = srv.stats()
    assert st["refills"] >= 1, st
    assert_same_edges(as_tuple(ra), oracle.search_single(p, None, None, 5.0), "cluster")
    assert_same_edges(as_tuple(rb), oracle.search_single(p[:100], None, None, 5.0), "part")


def test_serving_errors(nr):
    _print_case("serving errors")
    with nr.Server(slots=1) as srv:
        p = np.zeros((4, 3))
        with pytest.raises(ValueError, match="Cutoff must be positive and finite"):
            srv.submit(None, p, -1.0)
        with pytest.raises(ValueError, match="Cell matrix is not invertible"):
            srv.submit(nr.PyCell([[1, 0, 0], [0, 1, 0], [0, 0, 0]], [True] * 3), p, 2.0)
        r = srv.result(srv.submit(None, np.zeros((1, 3)), 2.0), 60.0)
        assert r["edge_index"].shape == (2, 0)


def test_serving_device_results(nr, oracle):
    import ctypes as C
    from neighborlist_rs import _lib
    reqs = _trace(seed=13, n=12, nmax=2500)
    _print_case("serving device results", len(reqs))
    with nr.Server(slots=2, max_atoms=6000) as srv:
        ids = [srv.submit(_cell(nr, H), p, rc) for (_, p, H, rc) in reqs]
        for k, rid in enumerate(ids):
            _, p, H, rc = reqs[k]
            m, pi, pj, ps = srv.result_device(rid, 60.0)
            ei = np.empty((2, m), np.int64)
            sh = np.empty((m, 3), np.int32)
            if m:
                for dst, src, nb in ((ei[0], pi, 8 * m), (ei[1], pj, 8 * m), (sh, ps, 12 * m)):
                    assert _lib.lib.nl_memcpy_d2h(dst.ctypes.data_as(C.c_void_p),
                                                  C.c_void_p(src), nb) == 0
            srv.release(rid)
            with pytest.raises(RuntimeError, match="not complete"):
                srv.release(rid)
            want = oracle.search_single(p, H, None if H is None else [True] * 3, rc)
            assert_same_edges((ei[0], ei[1], sh), want, f"request {k}")


def _mirror(ei, ej, s):
    keep = ei != ej
    return (np.concatenate([ei, ej[keep]]), np.concatenate([ej, ei[keep]]),
            np.concatenate([s, -s[keep]]))


def _sorted_all(ei, ej, s, d, v):
    o = np.lexsort((s[:, 2], s[:, 1], s[:, 0], ej, ei))
    return ei[o], ej[o], s[o], d[o], v[o]


def _torch_search(nr, p, H, rc):
    import torch
    t = nr.build_neighborlists_torch(_cell(nr, H), torch.from_numpy(np.ascontiguousarray(p)).cuda(),
                                     rc, full=True, distances=True, displacements=True)
    c = lambda x: x.detach().cpu().numpy()
    return (c(t["edge_index"][0]), c(t["edge_index"][1]), c(t["shift"]), c(t["distance"]),
            c(t["displacement"]))


def test_serving_full_lists_distances_displacements(nr, oracle):
    """A server created with full lists, distances and displacements: every
    request equals the oracle mirrored, and the torch search with the same
    options edge for edge (distance and displacement bit for bit); alone and
    batched give the same arrays, order included; a refill (short estimate)
    rewrites every array."""
    reqs = _trace(seed=19, n=20, nmax=2500)
    opts = dict(full=True, distances=True, displacements=True)
    _print_case("serving full+dist+disp alone", len(reqs))
    alone, _, _ = _serve(nr, reqs, slots=1, max_systems=1, **opts)
    _print_case("serving full+dist+disp mix")
    mix, info, st = _serve(nr, reqs, slots=3, max_atoms=20000, max_systems=64, **opts)
    assert st["max_batch"] > 1, st
    for k, ((_, p, H, rc), a, m) in enumerate(zip(reqs, alone, mix)):
        what = f"request {k}: n={p.shape[0]} rc={rc:.3f} cell={H is not None}"
        for key in ("edge_index", "shift", "distance", "displacement"):
            np.testing.assert_array_equal(a[key], m[key], err_msg=what + " " + key)
        ei, ej = m["edge_index"]
        sh, d, v = m["shift"], m["distance"], m["displacement"]
        assert info[k][0] == ei.size
        # mirror layout and geometry
        np.testing.assert_array_equal(ei[0::2], ej[1::2], err_msg=what)
        np.testing.assert_array_equal(sh[0::2], -sh[1::2], err_msg=what)
        np.testing.assert_array_equal(v[0::2], -v[1::2], err_msg=what)
        np.testing.assert_array_equal(d, np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) +
                                                 v[:, 2] * v[:, 2]), err_msg=what)
        want = oracle.search_single(p, H, None if H is None else [True] * 3, rc)
        assert_same_edges((ei, ej, sh), _mirror(*want), what)
        for x, y in zip(_sorted_all(ei, ej, sh, d, v), _sorted_all(*_torch_search(nr, p, H, rc))):
            np.testing.assert_array_equal(x, y, err_msg=what + " vs torch search")

    # ~1.6 M directed edges: more than the smallest output block holds
    rng = np.random.default_rng(4)
    p = np.vstack([rng.uniform(0.0, 6.0, size=(1500, 3)), [[800.0, 800.0, 800.0]]])
    _print_case("serving full+dist+disp refill", p.shape)
    with nr.Server(slots=1, **opts) as srv:
        r = srv.result(srv.submit(None, p, 5.0), 60.0)
        assert srv.stats()["refills"] >= 1
    ei, ej = r["edge_index"]
    assert_same_edges((ei, ej, r["shift"]), _mirror(*oracle.search_single(p, None, None, 5.0)),
                      "refill")
    for x, y in zip(_sorted_all(ei, ej, r["shift"], r["distance"], r["displacement"]),
                    _sorted_all(*_torch_search(nr, p, None, 5.0))):
        np.testing.assert_array_equal(x, y, err_msg="refill vs torch search")


def test_serving_result_arrays(nr, oracle):
    import ctypes as C
    from neighborlist_rs import _lib
    lib = _lib.lib
    reqs = _trace(seed=23, n=6, nmax=1500)
    _print_case("serving result arrays", len(reqs))
    with nr.Server(slots=2, distances=True) as srv:
        ids = [srv.submit(_cell(nr, H), p, rc) for (_, p, H, rc) in reqs]
        for k, rid in enumerate(ids):
            _, p, H, rc = reqs[k]
            assert srv.wait(rid, 60.0)
            m, _, _ = srv.info(rid)
            ptr = [C.c_void_p() for _ in range(5)]
            # displacements were not requested
            assert lib.nl_server_result_arrays(srv._p, rid, *[C.byref(x) for x in ptr]) != 0
            assert "NL_DISP" in _lib.last_error()
            assert lib.nl_server_result_arrays(srv._p, rid, *[C.byref(x) for x in ptr[:4]],
                                               None) == 0
            ei = np.empty((2, m), np.int64)
            sh = np.empty((m, 3), np.int32)
            d = np.empty(m)
            if m:
                for dst, src, nb in ((ei[0], ptr[0], 8 * m), (ei[1], ptr[1], 8 * m),
                                     (sh, ptr[2], 12 * m), (d, ptr[3], 8 * m)):
                    assert lib.nl_memcpy_d2h(dst.ctypes.data_as(C.c_void_p), src, nb) == 0
            srv.release(rid)
            want = oracle.search_single(p, H, None if H is None else [True] * 3, rc)
            assert_same_edges((ei[0], ei[1], sh), want, f"request {k}")
            import torch
            t = nr.build_neighborlists_torch(_cell(nr, H), torch.from_numpy(p).cuda(), rc,
                                             distances=True)
            c = lambda x: x.detach().cpu().numpy()
            o1 = np.lexsort((sh[:, 2], sh[:, 1], sh[:, 0], ei[1], ei[0]))
            ts = c(t["shift"])
            te = c(t["edge_index"])
            o2 = np.lexsort((ts[:, 2], ts[:, 1], ts[:, 0], te[1], te[0]))
            np.testing.assert_array_equal(d[o1], c(t["distance"])[o2], err_msg=f"request {k}")


def test_serving_borrowed_positions(nr, oracle):
    """submit(copy=False) (NL_SUBMIT_BORROW): the server reads the caller's
    positions when it stages the batch instead of copying them at submit;
    the lists are the same arrays as with copied positions, and the server
    holds no reference once a request is done."""
    reqs = _trace(seed=23, n=30)
    _print_case("serving borrowed positions")
    copied, _, _ = _serve(nr, reqs, slots=3, max_atoms=20000, max_systems=64)
    with nr.Server(slots=3, max_atoms=20000, max_systems=64) as srv:
        ids = [srv.submit(_cell(nr, H), p, rc, copy=False) for (_, p, H, rc) in reqs]
        borrowed = [srv.result(rid, timeout=60.0) for rid in ids]
        assert not srv._borrowed
    for k, (a, b) in enumerate(zip(copied, borrowed)):
        np.testing.assert_array_equal(a["edge_index"], b["edge_index"], err_msg=f"request {k}")
        np.testing.assert_array_equal(a["shift"], b["shift"], err_msg=f"request {k}")
    _, p, H, rc = reqs[0]
    want = oracle.search_single(p, H, None if H is None else [True] * 3, rc)
    assert_same_edges(as_tuple(borrowed[0]), want, "request 0")


def test_serving_warm_server_stops_allocating(nr):
    """Slot buffers grow towards the largest admitted batch and output blocks
    come in capacity classes, so repeating a trace reaches a pass with no
    new device or page-locked allocation (a freed device buffer would
    synchronise the device); every pass returns the same arrays."""
    reqs = _trace(seed=31, n=30)
    _print_case("serving warm server", len(reqs))
    with nr.Server(slots=2, max_atoms=20000, max_systems=64) as srv:
        first, counts = None, [srv.stats()["allocations"]]
        for _ in range(6):
            ids = [srv.submit(_cell(nr, H), p, rc) for (_, p, H, rc) in reqs]
            got = [srv.result(rid, timeout=60.0) for rid in ids]
            counts.append(srv.stats()["allocations"])
            if first is None:
                first = got
            else:
                for k, (a, b) in enumerate(zip(first, got)):
                    np.testing.assert_array_equal(a["edge_index"], b["edge_index"],
                                                  err_msg=f"request {k}")
                    np.testing.assert_array_equal(a["shift"], b["shift"], err_msg=f"request {k}")
    assert counts[1] > counts[0], counts
    assert any(b == a for a, b in zip(counts[1:], counts[2:])), counts
