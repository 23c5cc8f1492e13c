"""Single small systems run without the nine-launch pipeline: one launch
(k_small: plan, per-row counts, scan and fill in one block) or two
(k_brute_count over many blocks with the scan in the last block, then
k_brute_fill).

The lists must be byte-identical to the pipeline's (same order, not only the
same set): every case runs with the small paths on (nl_set_small_path 1, the
default: one launch up to 128 atoms, two up to 1024; 2: one launch and 3: two
launches for every single system up to 1024 atoms, even where the host
expects the cell-list path) and off (0), and the arrays are compared as they
come.  A few cases are also
checked against the oracle.  Each case prints its name before it runs."""
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


def _cluster(n, seed, density=0.1):
    rng = np.random.default_rng(seed)
    side = max(1.0, (n / density) ** (1.0 / 3.0))
    return rng.uniform(0.0, side, (n, 3))


def _both(set_mode, mode, fn):
    """fn() with the one-launch path in `mode`, then off: numpy copies."""
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


CASES = [
    # name, n, cutoffs, cell kind, options
    ("one atom", 1, [3.0], None, {}),
    ("two atoms", 2, [50.0], None, {}),
    ("ethanol-sized", 9, [6.0], None, {}),
    ("100 atoms", 100, [6.0], None, {}),
    ("500 atoms, three lists", 500, [5.0, 6.0, 14.0], None, {}),
    ("999 atoms", 999, [6.0], None, {}),
    ("300 atoms, disjoint", 300, [3.0, 5.0, 4.0], None, {"disjoint": True}),
    ("200 atoms, six lists", 200, [2.0, 3.0, 4.0, 5.0, 6.0, 7.0], None, {}),
    ("400 atoms, full lists", 400, [6.0], None, {"full": True}),
    ("400 atoms, distances and displacements", 400, [6.0], None,
     {"distances": True, "displacements": True}),
    ("300 atoms, full, distances, displacements, two lists", 300, [4.0, 6.0], None,
     {"full": True, "distances": True, "displacements": True}),
    ("periodic, minimum image", 300, [4.0], "orthorhombic", {}),
    ("triclinic, minimum image, two lists", 250, [3.0, 4.5], "triclinic", {}),
    ("periodic, mixed flags", 200, [4.0], "mixed", {"full": True, "distances": True}),
    ("periodic, minimum image, 100 atoms", 100, [4.0], "orthorhombic", {}),
    ("triclinic, 120 atoms, full, distances, displacements", 120, [3.0, 4.5], "triclinic",
     {"full": True, "distances": True, "displacements": True}),
]


def _cell_for(nr, kind, n):
    if kind is None:
        return None, None, None
    L = max(20.0, (n / 0.02) ** (1.0 / 3.0))
    H = np.diag([L, 1.1 * L, 0.95 * L])
    if kind == "triclinic":
        H[0, 1] = 0.2 * L
        H[0, 2] = -0.1 * L
        H[1, 2] = 0.15 * L
    pbc = [True, False, True] if kind == "mixed" else [True, True, True]
    return nr.PyCell(H.tolist(), pbc), H, pbc


@pytest.mark.parametrize("mode", [1, 2, 3])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_small_path_identical_to_pipeline(nr, small_mode, case, mode):
    import torch
    name, n, cut, kind, opts = case
    print("[case] small path", name, "mode", mode, flush=True)
    pos = _cluster(n, seed=n)
    cell, H, pbc = _cell_for(nr, kind, n)
    if H is not None:
        pos = (np.random.default_rng(n + 1).uniform(-0.5, 1.5, (n, 3)) @ H.T).copy()
    tpos = torch.from_numpy(pos).cuda()
    opts = dict(opts)
    disjoint = opts.pop("disjoint", False)

    def run():
        if len(cut) == 1:
            return nr.build_neighborlists_torch(cell, tpos, cut[0], **opts)
        return nr.build_neighborlists_multi_torch(cell, tpos, cut, disjoint=disjoint, **opts)

    a, b = _both(small_mode, mode, run)
    _assert_identical(a, b, name)
    # the numpy surface (host positions) through the same path
    if not opts:
        def run_np():
            if len(cut) == 1:
                return nr.build_neighborlists(cell, pos, cut[0])
            return nr.build_neighborlists_multi(cell, pos, cut, disjoint=disjoint)
        a, b = _both(small_mode, mode, run_np)
        _assert_identical(a, b, name + " (numpy)")


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_small_path_against_oracle(nr, oracle, small_mode, mode):
    small_mode(mode)
    for n, cut, kind in [(9, 6.0, None), (700, 6.0, None), (300, 4.0, "triclinic")]:
        print("[case] small path vs oracle n", n, "kind", kind, "mode", mode, flush=True)
        pos = _cluster(n, seed=7 * n)
        cell, H, pbc = _cell_for(nr, kind, n)
        if H is not None:
            pos = (np.random.default_rng(n).uniform(-0.5, 1.5, (n, 3)) @ H.T).copy()
        got = nr.build_neighborlists(cell, pos, cut)
        assert_same_edges(as_tuple(got), oracle.search_single(pos, H, pbc, cut), f"n {n}")
        multi = nr.build_neighborlists_multi(cell, pos, [cut, 0.5 * cut], disjoint=True)
        want = oracle.search_single_multi(pos, H, pbc, [cut, 0.5 * cut], True)
        for q in range(2):
            assert_same_edges(as_tuple(multi[q]), want[q], f"n {n} disjoint list {q}")


@pytest.mark.parametrize("mode", [2, 3])
def test_small_path_declines_cell_list_systems(nr, small_mode, mode):
    """A periodic cell too small for the minimum image (cutoff beyond half a
    width): with modes 2 and 3 the kernels' plan takes the cell-list path and
    hands the system back; the pipeline's lists come out."""
    import torch
    L = 12.0
    H = np.diag([L, L, L])
    pos = np.random.default_rng(5).uniform(0.0, L, (150, 3))
    cell = nr.PyCell(H.tolist(), [True, True, True])
    print("[case] small path declined, cutoff 7 in a 12 A cube, mode", mode, flush=True)
    tpos = torch.from_numpy(pos).cuda()
    a, b = _both(small_mode, mode, lambda: nr.build_neighborlists_torch(cell, tpos, 7.0))
    _assert_identical(a, b, "declined")
    assert a["edge_index"].shape[1] > 0


@pytest.mark.parametrize("mode", [2, 3])
def test_small_path_grows_its_output(nr, small_mode, mode):
    """Calls of one shape reuse the last total as the output estimate; a
    larger list (same atoms, larger cutoff) overflows it and the launch (the
    fill launch alone, in the two-launch form) is repeated with exact room."""
    import torch
    pos = torch.from_numpy(_cluster(800, seed=11)).cuda()
    sizes = []
    for rc in [2.0, 9.0, 3.0, 12.0]:
        print("[case] small path growth rc", rc, "mode", mode, flush=True)
        a, b = _both(small_mode, mode, lambda: nr.build_neighborlists_torch(
            None, pos, rc, distances=True))
        _assert_identical(a, b, f"rc {rc}")
        sizes.append(a["edge_index"].shape[1])
    assert sizes[1] > sizes[0] + sizes[0] // 8 + 4096


def test_small_path_row_limit_and_threshold(nr, small_mode):
    """128 atoms take one launch and 129 two by default; 1024 atoms (the
    kernels' row limit, threshold raised) take either form; 1025 atoms and
    systems at the threshold take the pipeline.  All agree with the
    pipeline."""
    import torch
    nr.set_brute_force_threshold(5000)
    for n, mode in [(128, 1), (129, 1), (1024, 1), (1025, 1), (1024, 2), (1024, 3), (1, 3),
                    (17, 3)]:
        print("[case] small path rows", n, "mode", mode, flush=True)
        pos = torch.from_numpy(_cluster(n, seed=n)).cuda()
        a, b = _both(small_mode, mode, lambda: nr.build_neighborlists_torch(None, pos, 5.0))
        _assert_identical(a, b, f"n {n}")
    nr.set_brute_force_threshold(100)
    pos = torch.from_numpy(_cluster(100, seed=3)).cuda()
    print("[case] small path at the threshold (cell-list path)", flush=True)
    a, b = _both(small_mode, 1, lambda: nr.build_neighborlists_torch(None, pos, 5.0))
    assert_same_edges(as_tuple(a), as_tuple(b), "n 100 at threshold 100")


def test_small_path_host_copy_of_small_lists(nr, small_mode):
    """Numpy callers: small lists are copied to page-locked memory with the
    search and read from there.  The copy is sized by the last total of the
    same shape; lists that outgrow it are read from the device instead.
    Same atoms, cutoffs up and down, against the pipeline."""
    pos = _cluster(40, seed=4)
    for rc in [1.5, 4.0, 2.0, 6.0, 6.0, 1.0]:
        print("[case] small path host copy rc", rc, flush=True)
        a, b = _both(small_mode, 1, lambda: nr.build_neighborlists(None, pos, rc))
        _assert_identical(a, b, f"rc {rc}")
        multi_a, multi_b = _both(small_mode, 1, lambda: nr.build_neighborlists_multi(
            None, pos, [rc, 0.5 * rc], disjoint=True))
        _assert_identical(multi_a, multi_b, f"rc {rc} multi")
