"""GPU skin lists for molecular dynamics (nl_skin_*, DynamicNeighborManager).

Along random-walk trajectories, against the oracle on the current positions:

* the measured largest displacement equals the host one (positions minus
  the positions of the last build) and a rebuild happens exactly when it
  exceeds half the smallest skin (or the cell changes);
* after every build the lists equal the oracle's at rc + skin, bit-exactly
  after the canonical sort (half lists, or the mirrored half list for full
  lists);
* at every step each list holds every pair the oracle finds within rc for
  the current positions (the property the skin exists for).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import assert_same_edges, canon

pytestmark = pytest.mark.gpu

NL_FULL = 16


def _print_case(*a):
    print("[case]", *a, flush=True)


class _Dev:
    def __init__(self, lib, host):
        self.lib = lib
        self.nbytes = host.nbytes
        self.ptr = lib.nl_device_malloc(max(self.nbytes, 1))
        assert self.ptr
        self.set(host)

    def set(self, host):
        assert self.lib.nl_memcpy_h2d(C.c_void_p(self.ptr), host.ctypes.data_as(C.c_void_p),
                                      self.nbytes) == 0

    def free(self):
        self.lib.nl_device_free(C.c_void_p(self.ptr))


def _cell_arrays(cells):
    S = len(cells)
    H = np.zeros((S, 9))
    has = np.zeros(S, np.uint8)
    pb = np.ones((S, 3), np.uint8)
    for s, c in enumerate(cells):
        if c is not None:
            H[s] = np.asarray(c, dtype=np.float64).reshape(9)
            has[s] = 1
    return H, has, pb


def _lists(lib, L, k, nq):
    out = []
    for q in range(nq):
        m = int(lib.nl_skin_num_edges(k, q))
        pe, ps = C.c_void_p(), C.c_void_p()
        assert lib.nl_skin_device(k, q, C.byref(pe), C.byref(ps)) == 0
        ei = np.empty((2, m), np.int64)
        sh = np.empty((m, 3), np.int32)
        if m:
            assert lib.nl_memcpy_d2h(ei.ctypes.data_as(C.c_void_p), pe, 16 * m) == 0
            assert lib.nl_memcpy_d2h(sh.ctypes.data_as(C.c_void_p), ps, 12 * m) == 0
        out.append((ei[0], ei[1], sh))
    return out


def _mirror(ei, ej, s):
    keep = ei != ej
    return (np.concatenate([ei, ej[keep]]), np.concatenate([ej, ei[keep]]),
            np.concatenate([s, -s[keep]]))


def _contains(big, small, what):
    """Every (i, j, S) of `small` is in `big`."""
    def keys(t):
        ei, ej, s = canon(*t)
        return np.stack([ei, ej, s[:, 0], s[:, 1], s[:, 2]], axis=1)
    kb, ks = keys(big), keys(small)
    if ks.shape[0] == 0:
        return
    allk = np.concatenate([kb, ks])
    _, inv = np.unique(allk, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    missing = np.setdiff1d(inv[kb.shape[0]:], inv[:kb.shape[0]])
    assert missing.size == 0, f"{what}: {missing.size} pairs within rc missing from the skin list"


def _oracle_batch(oracle, pos, sizes, cells, cutoffs, multi):
    batch = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)
    oc = [None if c is None else (np.asarray(c), [True] * 3) for c in cells]
    if multi:
        return oracle.search_batch_multi(pos, batch, oc, cutoffs, False)
    return [oracle.search_batch(pos, batch, oc, cutoffs[0])]


def _run_trajectory(nr, oracle, pos0, sizes, cells, cutoffs, skins, steps, sigma, seed,
                    full=False, multi=True, new_cells=None):
    from neighborlist_rs import _lib as L
    lib = L.lib
    rng = np.random.default_rng(seed)
    H, has, pb = _cell_arrays(cells)
    sz = np.ascontiguousarray(sizes, dtype=np.int64)
    cut = np.ascontiguousarray(cutoffs, dtype=np.float64)
    sk = np.ascontiguousarray(skins, dtype=np.float64)
    radii = list(cut + sk)
    trigger = 0.5 * sk.min()
    pos = np.ascontiguousarray(pos0.copy())
    d = _Dev(lib, pos)
    k = C.c_void_p()
    assert lib.nl_skin_create(C.c_void_p(d.ptr), pos.shape[0], sz.ctypes.data_as(L.i64p),
                              sz.size, H.ctypes.data_as(L.dp), has.ctypes.data_as(L.u8p),
                              pb.ctypes.data_as(L.u8p), cut.ctypes.data_as(L.dp),
                              sk.ctypes.data_as(L.dp), cut.size, 0, int(multi),
                              NL_FULL if full else 0, None, C.byref(k)) == 0, L.last_error()
    ref = pos.copy()
    builds = 1
    cur_cells = list(cells)
    try:
        def check_built(what):
            want = _oracle_batch(oracle, pos, sizes, cur_cells, radii, multi)
            got = _lists(lib, L, k, cut.size)
            for q in range(cut.size):
                w = _mirror(*want[q]) if full else want[q]
                assert_same_edges(got[q], w, f"{what} list {q} at r_list")

        check_built("initial build")
        for step in range(steps):
            pos = np.ascontiguousarray(pos + rng.normal(0.0, sigma, pos.shape))
            d.set(pos)
            change = new_cells is not None and step == steps // 2
            disp = C.c_double(0.0)
            needs = C.c_int(0)
            assert lib.nl_skin_check(k, C.c_void_p(d.ptr), None, C.byref(disp),
                                     C.byref(needs)) == 0
            host = float(np.sqrt(((pos - ref) ** 2).sum(axis=1).max()))
            assert abs(disp.value - host) <= 1e-12 * max(1.0, host), (disp.value, host)
            assert bool(needs.value) == (host > trigger)
            rebuilt = C.c_int(0)
            if change:
                cur_cells = list(new_cells)
                H2, _, _ = _cell_arrays(cur_cells)
                assert lib.nl_skin_update(k, C.c_void_p(d.ptr), H2.ctypes.data_as(L.dp), None,
                                          C.byref(rebuilt)) == 0, L.last_error()
                assert rebuilt.value == 1
            else:
                assert lib.nl_skin_update(k, C.c_void_p(d.ptr), None, None,
                                          C.byref(rebuilt)) == 0, L.last_error()
                assert rebuilt.value == int(host > trigger), (step, host, trigger)
            what = f"step {step}"
            if rebuilt.value:
                ref = pos.copy()
                builds += 1
                check_built(what)
            # every pair within rc now is in the skin list
            now = _oracle_batch(oracle, pos, sizes, cur_cells, list(cut), multi)
            got = _lists(lib, L, k, cut.size)
            for q in range(cut.size):
                _contains(got[q], _mirror(*now[q]) if full else now[q], f"{what} list {q}")
        u, b, dl = C.c_int64(), C.c_int64(), C.c_double()
        assert lib.nl_skin_stats(k, C.byref(u), C.byref(b), C.byref(dl)) == 0
        assert u.value == steps and b.value == builds
        return builds
    finally:
        lib.nl_skin_free(k)
        d.free()


def test_skin_single_triclinic(nr, oracle):
    rng = np.random.default_rng(31)
    Hc = np.array([[22.0, 3.0, 1.5], [0.0, 21.0, 2.0], [0.0, 0.0, 23.0]])
    pos = rng.uniform(0, 1, (1500, 3)) @ Hc.T
    _print_case("skin single triclinic")
    builds = _run_trajectory(nr, oracle, pos, [1500], [Hc], [5.0], [1.0], steps=16, sigma=0.04,
                             seed=1, multi=False)
    assert 2 <= builds < 16  # some steps rebuilt, most reused the lists


def test_skin_batch_multi_full_and_cell_change(nr, oracle):
    from neighborlist_rs.workloads import c4_ragged_batch
    pos, sizes, cells, _ = c4_ragged_batch(seed=8, n_systems=12, nmin=40, nmax=600)
    new_cells = [None if c is None else np.asarray(c) * 1.01 for c in cells]
    _print_case("skin batch multi full")
    builds = _run_trajectory(nr, oracle, pos, sizes, cells, [3.0, 5.0], [0.6, 1.0], steps=10,
                             sigma=0.03, seed=2, full=True, multi=True, new_cells=new_cells)
    assert builds >= 2


def test_skin_errors(nr):
    from neighborlist_rs import _lib as L
    lib = L.lib
    pos = np.random.default_rng(0).uniform(0, 10, (50, 3))
    d = _Dev(lib, pos)
    try:
        sz = np.array([50], np.int64)
        cut = np.array([3.0])
        k = C.c_void_p()
        for bad in (-0.5, float("nan")):
            sk = np.array([bad])
            assert lib.nl_skin_create(C.c_void_p(d.ptr), 50, sz.ctypes.data_as(L.i64p), 1, None,
                                      None, None, cut.ctypes.data_as(L.dp),
                                      sk.ctypes.data_as(L.dp), 1, 0, 0, 0, None,
                                      C.byref(k)) != 0
            assert "Skins must be non-negative and finite" in L.last_error()
        sk = np.array([0.5])
        assert lib.nl_skin_create(C.c_void_p(d.ptr), 50, sz.ctypes.data_as(L.i64p), 1, None, None,
                                  None, cut.ctypes.data_as(L.dp), sk.ctypes.data_as(L.dp), 1, 0,
                                  0, 8, None, C.byref(k)) != 0  # distances are not offered
        assert lib.nl_skin_create(C.c_void_p(d.ptr), 50, sz.ctypes.data_as(L.i64p), 1, None, None,
                                  None, cut.ctypes.data_as(L.dp), sk.ctypes.data_as(L.dp), 1, 0,
                                  0, 0, None, C.byref(k)) == 0
        # a NaN position is an unbounded displacement: rebuild
        bad = pos.copy()
        bad[7, 1] = np.nan
        d.set(bad)
        needs = C.c_int(0)
        assert lib.nl_skin_check(k, C.c_void_p(d.ptr), None, None, C.byref(needs)) == 0
        assert needs.value == 1
        lib.nl_skin_free(k)
    finally:
        d.free()


def test_dynamic_neighbor_manager_torch(nr, oracle):
    import torch
    rng = np.random.default_rng(41)
    Hc = np.array([[18.0, 0.0, 0.0], [1.0, 17.0, 0.0], [0.5, 1.0, 19.0]]).T  # columns = vectors
    pos = rng.uniform(0, 1, (900, 3)) @ Hc.T
    cell = nr.PyCell(Hc.tolist(), [True] * 3)
    _print_case("torch dynamic manager")
    pt = torch.from_numpy(pos).cuda()
    with nr.DynamicNeighborManager(pt, 4.5, 0.8, cell) as m:
        g = m.get()
        want = oracle.search_single(pos, Hc, [True] * 3, 5.3)
        assert_same_edges((g["edge_index"][0].cpu().numpy(), g["edge_index"][1].cpu().numpy(),
                           g["shift"].cpu().numpy()), want, "initial")
        ref = pos.copy()
        for step in range(8):
            pos = pos + rng.normal(0.0, 0.05, pos.shape)
            pt = torch.from_numpy(pos).cuda()
            host = float(np.sqrt(((pos - ref) ** 2).sum(axis=1).max()))
            assert abs(m.max_displacement(pt) - host) <= 1e-12 * max(1.0, host)
            assert m.needs_rebuild(pt) == (host > 0.4)
            r = m.update(pt)
            assert r == (host > 0.4)
            if r:
                ref = pos.copy()
            g = m.get()
            now = oracle.search_single(pos, Hc, [True] * 3, 4.5)
            _contains((g["edge_index"][0].cpu().numpy(), g["edge_index"][1].cpu().numpy(),
                       g["shift"].cpu().numpy()), now, f"torch step {step}")
        st = m.stats()
        assert st["updates"] == 8 and st["builds"] >= 2
