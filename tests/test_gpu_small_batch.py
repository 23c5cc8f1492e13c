"""Batches of small systems that all take a brute-force path run without the
pipeline: a count launch (16 rows per block, each block planning its own
system, the scan of the [list][atom] counts in the last block) and a fill
launch.

The lists must be byte-identical to the pipeline's (same order, not only the
same set): every case runs with the small paths on (nl_set_small_path 1, the
default; 2 skips the host's minimum-image screen so that the kernels' plan
decides) and off (0), and the arrays are compared as they come.  Some cases
are also checked against the oracle.  Each case prints its name before it
runs."""
import ctypes as C

import numpy as np
import pytest

from conftest import as_tuple, assert_same_edges

pytestmark = pytest.mark.gpu


@pytest.fixture
def small_mode(nr):
    lib = nr._lib.lib
    old = lib.nl_get_small_path()
    old_thr = nr.get_brute_force_threshold()
    yield lib.nl_set_small_path
    lib.nl_set_small_path(old)
    nr.set_brute_force_threshold(old_thr)


def _both(set_mode, mode, fn):
    """fn() with the small paths in `mode`, then off: numpy copies."""
    import torch

    def grab(out):
        if isinstance(out, dict) and "edge_index" not in out:
            return {k: grab(v) for k, v in out.items()}
        return {k: (v.detach().cpu().numpy() if torch.is_tensor(v) else np.asarray(v))
                for k, v in out.items()}

    set_mode(mode)
    a = grab(fn())
    set_mode(0)
    b = grab(fn())
    return a, b


def _assert_identical(a, b, what):
    if "edge_index" not in a:
        assert a.keys() == b.keys(), what
        for key in a:
            _assert_identical(a[key], b[key], f"{what} list {key}")
        return
    assert a.keys() == b.keys(), what
    for key in a:
        assert a[key].dtype == b[key].dtype and a[key].shape == b[key].shape, (what, key)
        np.testing.assert_array_equal(a[key], b[key], err_msg=f"{what}: {key}")


def _clusters(sizes, seed, density=0.1, gap=40.0):
    """Isolated clusters at `density`, one per size, placed `gap` apart."""
    rng = np.random.default_rng(seed)
    out = []
    for s, n in enumerate(sizes):
        side = max(1.0, (n / density) ** (1.0 / 3.0))
        out.append(rng.uniform(0.0, side, (n, 3)) + np.array([gap * s, 0.0, 0.0]))
    pos = np.concatenate(out) if out else np.zeros((0, 3))
    batch = np.repeat(np.arange(len(sizes)), sizes).astype(np.int32)
    return np.ascontiguousarray(pos), batch


def _periodic(sizes, seed, kinds):
    """Periodic systems large enough for the minimum image at cutoffs <= 5."""
    rng = np.random.default_rng(seed)
    pos, cells = [], []
    for s, n in enumerate(sizes):
        L = max(12.0, (n / 0.02) ** (1.0 / 3.0))
        H = np.diag([L, 1.05 * L, 0.97 * L])
        pbc = [True, True, True]
        kind = kinds[s % len(kinds)]
        if kind == "triclinic":
            H[0, 1] = 0.15 * L
            H[1, 2] = -0.1 * L
        elif kind == "mixed":
            pbc = [True, False, True]
        pos.append(rng.uniform(-0.3, 1.3, (n, 3)) @ H.T)
        cells.append((H, pbc))
    batch = np.repeat(np.arange(len(sizes)), sizes).astype(np.int32)
    return np.ascontiguousarray(np.concatenate(pos)), batch, cells


def _cell_rows(cells):
    """The Python batch convention: (B, 3, 3) rows of H (all periodic)."""
    return np.ascontiguousarray(np.stack([H.T for H, _ in cells]))


def _search_ex(pos, sizes, cells, cut, disjoint=False, multi=True, options=0):
    """nl_search_ex on host positions, per-system (H, pbc) or None: numpy
    lists {q: {"edge_index", "shift"}} (and "dist" with NL_DIST)."""
    from neighborlist_rs import _lib as L
    lib = L.lib
    sizes = np.ascontiguousarray(np.asarray(sizes, np.int64))
    S = sizes.size
    H9 = np.zeros((S, 9))
    has = np.zeros(S, np.uint8)
    pb = np.zeros((S, 3), np.uint8)
    for s, c in enumerate(cells if cells is not None else [None] * S):
        if c is not None:
            H9[s] = np.asarray(c[0], np.float64).reshape(9)
            has[s] = 1
            pb[s] = [1 if v else 0 for v in c[1]]
    cut = np.ascontiguousarray(np.asarray(cut, np.float64))
    r = C.c_void_p()
    assert lib.nl_search_ex(
        pos.ctypes.data_as(L.dp), pos.shape[0], sizes.ctypes.data_as(L.i64p), S,
        H9.ctypes.data_as(L.dp), has.ctypes.data_as(L.u8p), pb.ctypes.data_as(L.u8p),
        cut.ctypes.data_as(L.dp), cut.size, int(disjoint), int(multi), options, None,
        C.byref(r)) == 0, L.last_error()
    out = {}
    for q in range(cut.size):
        m = int(lib.nl_result_num_edges(r, q))
        ei = np.zeros((2, m), np.int64)
        sh = np.zeros((m, 3), np.int32)
        d = {"edge_index": ei, "shift": sh}
        if m:
            assert lib.nl_result_copy(r, q, ei[0].ctypes.data_as(L.i64p),
                                      ei[1].ctypes.data_as(L.i64p),
                                      sh.ctypes.data_as(L.i32p)) == 0
        if options & 8:  # NL_DIST
            dist = np.zeros(m)
            if m:
                assert lib.nl_result_copy_dist(r, q, dist.ctypes.data_as(L.dp)) == 0
            d["dist"] = dist
        out[q] = d
    lib.nl_result_free(r)
    return out


CASES = [
    # name, sizes, cutoffs, options
    ("two atoms and one", [2, 1], [50.0], {}),
    ("gnn-sized molecules", [9, 21, 42, 60, 33, 17, 42, 12] * 4, [6.0], {}),
    ("ragged up to 1000 atoms, three lists", [1, 7, 64, 300, 999, 150, 2], [5.0, 6.0, 14.0], {}),
    ("disjoint shells", [40, 80, 120, 5], [3.0, 5.0, 4.0], {"disjoint": True}),
    ("six lists", [30, 90, 200], [2.0, 3.0, 4.0, 5.0, 6.0, 7.0], {}),
    ("nine lists (two kernel groups)", [50, 70], [1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 4.5, 5.0],
     {}),
    ("full lists", [100, 250, 17], [6.0], {"full": True}),
    ("distances and displacements", [100, 250, 17], [6.0],
     {"distances": True, "displacements": True}),
    ("full, distances, displacements, two lists", [64, 128, 300], [4.0, 6.0],
     {"full": True, "distances": True, "displacements": True}),
    ("256 systems of up to 100 atoms", list(np.random.default_rng(3).integers(1, 100, 256)),
     [6.0], {}),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_small_batch_identical_to_pipeline(nr, small_mode, case):
    import torch
    name, sizes, cut, opts = case
    print("[case] small batch", name, flush=True)
    pos, batch = _clusters([int(s) for s in sizes], seed=len(sizes))
    tpos = torch.from_numpy(pos).cuda()
    opts = dict(opts)
    disjoint = opts.pop("disjoint", False)

    def run():
        if len(cut) == 1:
            return nr.build_neighborlists_batch_torch(tpos, batch, None, cut[0], **opts)
        return nr.build_neighborlists_batch_multi_torch(tpos, batch, None, cut,
                                                        disjoint=disjoint, **opts)

    a, b = _both(small_mode, 1, run)
    _assert_identical(a, b, name)
    if not opts:
        def run_np():
            if len(cut) == 1:
                return nr.build_neighborlists_batch(pos, batch, None, cut[0])
            return nr.build_neighborlists_batch_multi(pos, batch, None, cut, disjoint=disjoint)
        a, b = _both(small_mode, 1, run_np)
        _assert_identical(a, b, name + " (numpy)")


@pytest.mark.parametrize("mode", [1, 2])
def test_small_batch_periodic_identical_to_pipeline(nr, small_mode, mode):
    import torch
    sizes = [60, 200, 35, 500, 90]
    pos, batch, cells = _periodic(sizes, 11, ["orthorhombic", "triclinic", "mixed"])
    print("[case] small batch periodic, mixed flags, full lists, distances, mode", mode,
          flush=True)
    a, b = _both(small_mode, mode, lambda: _search_ex(pos, sizes, cells, [4.5], multi=False,
                                                      options=16 | 8))
    _assert_identical(a, b, "periodic")
    print("[case] small batch periodic, multi with disjoint, mode", mode, flush=True)
    a, b = _both(small_mode, mode, lambda: _search_ex(pos, sizes, cells, [2.5, 4.5],
                                                      disjoint=True))
    _assert_identical(a, b, "periodic multi")
    print("[case] small batch periodic through the torch surface, mode", mode, flush=True)
    p2, b2, c2 = _periodic(sizes, 12, ["orthorhombic", "triclinic"])
    rows = _cell_rows(c2)
    tpos = torch.from_numpy(p2).cuda()
    a, b = _both(small_mode, mode, lambda: nr.build_neighborlists_batch_torch(
        tpos, b2, rows, 4.5, displacements=True))
    _assert_identical(a, b, "periodic torch")


@pytest.mark.parametrize("mode", [1, 2])
def test_small_batch_declines_cell_list_systems(nr, small_mode, mode):
    """One periodic system whose cutoff exceeds half a width: in mode 1 the
    host sends the batch to the pipeline; in mode 2 the kernels' plan takes
    the cell-list path for that system, the count launch writes bounds = -1
    and the pipeline runs.  Both give the pipeline's lists."""
    sizes = [60, 200, 35, 40]
    pos, batch, cells = _periodic(sizes[:3], 5, ["orthorhombic"])
    L = 9.0
    small = np.random.default_rng(1).uniform(0.0, L, (40, 3))
    pos = np.ascontiguousarray(np.concatenate([pos, small]))
    cells = cells + [(np.diag([L, L, L]), [True, True, True])]
    print("[case] small batch with a cell-list system, mode", mode, flush=True)
    a, b = _both(small_mode, mode, lambda: _search_ex(pos, sizes, cells, [5.0], multi=False))
    _assert_identical(a, b, "declined")
    assert a[0]["edge_index"].shape[1] > 0
    # and again after the declined call: the ticket and flag were re-armed
    print("[case] small batch after a declined one", flush=True)
    p2, b2 = _clusters([30, 40, 50], 2)
    a, b = _both(small_mode, mode, lambda: nr.build_neighborlists_batch(p2, b2, None, 6.0))
    _assert_identical(a, b, "after declined")


def test_small_batch_grows_its_output(nr, small_mode):
    """Calls of one shape reuse the last total as the output estimate; a
    larger list overflows it and the fill launch alone is repeated with
    exact room."""
    import torch
    pos, batch = _clusters([200, 700, 300], 9)
    tpos = torch.from_numpy(pos).cuda()
    sizes = []
    for rc in [2.0, 9.0, 3.0, 12.0]:
        print("[case] small batch growth rc", rc, flush=True)
        a, b = _both(small_mode, 1, lambda: nr.build_neighborlists_batch_torch(
            tpos, batch, None, rc, distances=True))
        _assert_identical(a, b, f"rc {rc}")
        sizes.append(a["edge_index"].shape[1])
    assert sizes[1] > sizes[0] + sizes[0] // 8 + 4096


def test_small_batch_limits(nr, small_mode):
    """Systems of 1024 atoms (threshold raised) take the batched form, 1025
    and the threshold itself send the batch to the pipeline, and so does a
    batch of more than 65,536 atoms x lists.  All agree with the pipeline."""
    nr.set_brute_force_threshold(5000)
    for sizes, cut in [([1024, 10], [5.0]), ([1025, 10], [5.0]),
                       ([500] * 131 + [37], [3.0]), ([500] * 40, [2.0, 3.0, 4.0, 4.5])]:
        print("[case] small batch limits", len(sizes), "systems", max(sizes), "max atoms",
              len(cut), "lists", flush=True)
        pos, batch = _clusters(sizes, len(sizes))
        a, b = _both(small_mode, 1, lambda: nr.build_neighborlists_batch_multi(
            pos, batch, None, cut))
        _assert_identical(a, b, f"{len(sizes)} systems")
    nr.set_brute_force_threshold(100)
    pos, batch = _clusters([99, 100, 20], 4)
    print("[case] small batch at the threshold", flush=True)
    a, b = _both(small_mode, 1, lambda: nr.build_neighborlists_batch(pos, batch, None, 5.0))
    assert_same_edges(as_tuple(a), as_tuple(b), "at threshold 100")


def test_small_batch_local_indices(nr, small_mode):
    """nl_search_ex with per-system sizes and local indices, against the
    pipeline."""
    sizes = [30, 45, 12, 1]
    pos, _ = _clusters(sizes, 6)
    print("[case] small batch, local indices", flush=True)
    a, b = _both(small_mode, 1, lambda: _search_ex(pos, sizes, None, [4.0, 6.0],
                                                   options=2))  # NL_LOCAL_INDEX
    _assert_identical(a, b, "local indices")
    assert a[1]["edge_index"].shape[1] > 0
    assert a[1]["edge_index"].max() < 45


def test_small_batch_against_oracle(nr, oracle, small_mode):
    small_mode(1)
    pos, batch = _clusters([9, 21, 42, 60, 1, 300], 21)
    print("[case] small batch vs oracle, isolated", flush=True)
    got = nr.build_neighborlists_batch(pos, batch, None, 6.0)
    assert_same_edges(as_tuple(got), oracle.search_batch(pos, batch, None, 6.0), "isolated")
    got = nr.build_neighborlists_batch_multi(pos, batch, None, [6.0, 3.0], disjoint=True)
    want = oracle.search_batch_multi(pos, batch, None, [6.0, 3.0], True)
    for q in range(2):
        assert_same_edges(as_tuple(got[q]), want[q], f"isolated multi list {q}")
    print("[case] small batch vs oracle, periodic", flush=True)
    sizes = [50, 120, 80]
    pos, batch, cells = _periodic(sizes, 8, ["orthorhombic", "triclinic", "mixed"])
    got = _search_ex(pos, sizes, cells, [4.0], multi=False)
    assert_same_edges(as_tuple(got[0]), oracle.search_batch(pos, batch, cells, 4.0),
                      "periodic")
