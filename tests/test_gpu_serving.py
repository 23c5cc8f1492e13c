"""GPU serving: the continuous-batching server against the plain path and the oracle.

Each request of a Poisson trace (random isolated, orthorhombic and triclinic
systems, each with its own cutoff) must give the oracle's edges bit-exactly
after the canonical sort, and exactly the same arrays (order included)
whether it ran alone or batched with other requests.
"""
import threading

import numpy as np
import pytest

from conftest import as_tuple, assert_same_edges

pytestmark = pytest.mark.gpu


def _print_case(*a):
    print("[case]", *a, flush=True)


def _cell(nr, H):
    return None if H is None else nr.PyCell(np.asarray(H).tolist(), [True, True, True])


def _trace(seed=11, n=40, nmax=3000):
    from neighborlist_rs.workloads import c5_serving_trace
    return c5_serving_trace(seed=seed, n_requests=n, nmin=8, nmax=nmax)


def _serve(nr, reqs, **kw):
    with nr.Server(**kw) as srv:
        ids = [srv.submit(_cell(nr, H), p, rc) for (_, p, H, rc) in reqs]
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
    # box is far too small, so the fill pass must run again into larger buffers
    rng = np.random.default_rng(3)
    p = np.vstack([rng.uniform(0.0, 6.0, size=(600, 3)), [[900.0, 900.0, 900.0]]])
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
