"""Pins of the CPU oracle against what the reference itself fixes.

Each check here is an answer the reference's own tests or logs hold, or a
mathematical invariant that a plausible slip (dropped term, wrong sign or
index, transposed operand, wrong rounding) would break.  None of them
re-types the oracle's formulas.
"""
import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

from conftest import assert_same_edges, canon
from xyzio import read_xyz

SQ2 = 2.0 ** 0.5


# ---- cell.rs:102-239 -------------------------------------------------------
def test_cell_fractional_and_triclinic(oracle):
    inv = oracle.cell_inverse(np.eye(3) * 10.0)
    np.testing.assert_allclose(inv @ [5.0, 2.0, 8.0], [0.5, 0.2, 0.8])
    H = np.array([[10.0, 2.0, 1.0], [0.0, 10.0, 0.5], [0.0, 0.0, 10.0]])
    np.testing.assert_allclose(oracle.cell_inverse(H) @ [13.0, 10.5, 10.0], [1, 1, 1])
    np.testing.assert_allclose(oracle.cell_inverse(H) @ H, np.eye(3), atol=1e-15)


def test_cell_invalid(oracle):
    with pytest.raises(ValueError, match="Cell matrix is not invertible"):
        oracle.cell_inverse(np.zeros((3, 3)))


def test_cell_wrap(oracle):
    np.testing.assert_allclose(oracle.cell_wrap(np.eye(3) * 10, [1, 1, 1], [15, -2, 8]), [5, 8, 8])
    np.testing.assert_allclose(oracle.cell_wrap(np.eye(3) * 10, [1, 0, 0], [15, 15, 15]),
                               [5, 15, 15])


def test_cell_minimum_image(oracle):
    s, d = oracle.cell_mic(np.eye(3) * 10, [1, 1, 1], [1, 1, 1], [9, 9, 9])
    assert s.tolist() == [-1, -1, -1]
    np.testing.assert_allclose(d, [-2, -2, -2])
    s, d = oracle.cell_mic(np.eye(3) * 10, [1, 0, 0], [1, 1, 1], [9, 9, 9])
    assert s.tolist() == [-1, 0, 0]
    np.testing.assert_allclose(d, [-2, 8, 8])


def test_round_half_away_from_zero(oracle):
    # f64::round: 0.5 -> 1, -0.5 -> -1 (a banker's rounding slip gives 0)
    s, _ = oracle.cell_mic(np.eye(3) * 10, [1, 1, 1], [0, 0, 0], [5.0, -5.0, 15.0])
    assert s.tolist() == [-1, 1, -2]


def test_perpendicular_widths(oracle):
    np.testing.assert_allclose(oracle.perpendicular_widths(np.diag([10.0, 5.0, 2.0])),
                               [10, 5, 2])
    H = np.array([[1.0, 1.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    assert oracle.perpendicular_widths(H)[0] == pytest.approx(1.0 / SQ2)


def test_inverse_identity_and_triclinic_exact(oracle):
    # cell.rs test_getters: the inverse of the identity IS the identity
    # (assert_eq!, exact); cell.rs test_triclinic_cell: (13, 10.5, 10) maps to
    # (1, 1, 1) within approx's default epsilon (f64::EPSILON, relative and
    # absolute), i.e. to within one ulp of 1
    assert np.array_equal(oracle.cell_inverse(np.eye(3)), np.eye(3))
    H = np.array([[10.0, 2.0, 1.0], [0.0, 10.0, 0.5], [0.0, 0.0, 10.0]])
    inv = oracle.cell_inverse(H)
    x = np.array([13.0, 10.5, 10.0])
    frac = [(inv[r, 0] * x[0] + inv[r, 1] * x[1]) + inv[r, 2] * x[2] for r in range(3)]
    eps = np.finfo(np.float64).eps
    for f in frac:
        assert abs(f - 1.0) <= eps, frac


def test_inverse_ill_conditioned_cells(oracle):
    # H^-1 H = I to within a few condition numbers of rounding, on skewed,
    # nearly singular, non-symmetric cells: a transposed cofactor matrix, a
    # swapped minor or a wrong determinant sign all leave residuals of O(1)
    rng = np.random.default_rng(7)
    eps = np.finfo(np.float64).eps
    for t in range(200):
        a = rng.normal(size=(3, 3)) * 10.0
        a[2] = a[0] + a[1] + rng.normal(size=3) * 10.0 ** rng.uniform(-6, 0)
        inv = oracle.cell_inverse(a)
        cond = np.linalg.cond(a)
        res = np.abs(inv @ a - np.eye(3)).max()
        assert res <= 64 * cond * eps, (t, res, cond)
        res2 = np.abs(a @ inv - np.eye(3)).max()
        assert res2 <= 64 * cond * eps, (t, res2, cond)


# ---- search.rs:1413-1910 ---------------------------------------------------
def _lcg_positions(n=100, seed=42):
    """search.rs:1413-1432: x <- x*6364136223846793005 + 1, value x / u64::MAX."""
    x = seed
    out = []
    for _ in range(3 * n):
        x = (x * 6364136223846793005 + 1) % (1 << 64)
        out.append(float(x) / float((1 << 64) - 1) * 10.0)
    return np.array(out).reshape(n, 3)


def test_bruteforce_lcg_vs_plain_loop(oracle):
    p = _lcg_positions()
    ei, ej, s = oracle.search_single(p, None, None, 2.5)
    want = [(i, j) for i in range(100) for j in range(i + 1, 100)
            if ((p[i] - p[j]) ** 2).sum() < 6.25]
    assert list(zip(ei.tolist(), ej.tolist())) == want  # row-major, i < j
    assert np.all(s == 0)


def test_z_order(oracle):
    z1 = oracle.z_order([0.1, 0.1, 0.1])
    z2 = oracle.z_order([0.1, 0.1, 0.11])
    z3 = oracle.z_order([0.9, 0.9, 0.9])
    assert z1 < z2 < z3
    assert oracle.z_order([1.1, -0.1, 0.5]) > 0


def test_cell_list_structure(oracle):
    p = np.array([[1.0, 1.0, 1.0], [9.0, 9.0, 9.0], [1.1, 1.1, 1.1]])
    nb, b0 = oracle.celllist_bin_atoms(p, np.eye(3) * 10, [1, 1, 1], 3.0, 0, 0, 0)
    assert nb == (3, 3, 3) and b0 == [0, 2]
    assert oracle.celllist_bin_atoms(p, np.eye(3) * 10, [1, 1, 1], 3.0, 2, 2, 2)[1] == [1]
    assert oracle.celllist_bin_atoms(p, np.eye(3) * 10, [1, 1, 1], 3.0, 1, 1, 1)[1] == []
    # spatial reordering (search.rs:1477-1520): 0 and 2 share bin (0,0,0) at rc 2
    assert oracle.celllist_bin_atoms(p, np.eye(3) * 10, [1, 1, 1], 2.0, 0, 0, 0)[1] == [0, 2]


def test_brute_force_reference(oracle):
    p = np.array([[1.0, 1.0, 1.0], [1.0, 3.5, 1.0], [8.5, 1.0, 1.0]])
    ei, ej, _ = oracle.search_single(p, np.eye(3) * 10, [1, 1, 1], 3.0)
    assert sorted(zip(ei.tolist(), ej.tolist())) == [(0, 1), (0, 2)]


def test_cell_list_vs_brute_force(oracle):
    p = np.array([[1.0, 1.0, 1.0], [1.2, 1.2, 1.2], [9.8, 9.8, 9.8], [5.0, 5.0, 5.0]])
    oracle.set_threshold("brute_force", 0)
    try:
        cl = oracle.search_single(p, np.eye(3) * 10, [1, 1, 1], 2.0)
    finally:
        oracle.set_threshold("brute_force", 1000)
    bf = oracle.search_single(p, np.eye(3) * 10, [1, 1, 1], 2.0)
    assert sorted(zip(cl[0].tolist(), cl[1].tolist())) == sorted(zip(bf[0].tolist(),
                                                                       bf[1].tolist()))


def test_parallel_vs_serial_grid(oracle):
    g = np.array([(i, j, k) for i in range(5) for j in range(5) for k in range(5)], float)
    a = oracle.search_single(g, np.eye(3) * 10, [1, 1, 1], 1.5, parallel=True)
    b = oracle.search_single(g, np.eye(3) * 10, [1, 1, 1], 1.5, parallel=False)
    assert_same_edges(a, b)


def test_dense_bin(oracle):
    p = np.array([[0.1 * i, 0.1, 0.1] for i in range(5)])
    ei, _, _ = oracle.search_single(p, np.eye(3) * 10, [1, 1, 1], 2.0)
    assert len(ei) == 10


def test_large_search_radius(oracle):
    ei, ej, s = oracle.search_single(np.array([[0.5, 0.5, 0.5]]), np.eye(3), [1, 1, 1], 10.0)
    assert len(ei) > 0
    # self images of a unit cube within 10: lattice points 0 < |S| < 10, both signs
    g = np.arange(-10, 11)
    S = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    r2 = (S ** 2).sum(1)
    assert len(ei) == int(((r2 > 0) & (r2 < 100)).sum())


# ---- single.rs:177-284, batch.rs:297-564 -----------------------------------
def test_single_basic_and_multi(oracle):
    p2 = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    ei, ej, _ = oracle.search_single(p2, None, None, 1.5, parallel=False)
    assert ei.tolist() == [0] and ej.tolist() == [1]
    p3 = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.5, 0, 0]])
    r = oracle.search_single_multi(p3, None, None, [1.5, 3.0], False)
    assert len(r[0][0]) == 1 and len(r[1][0]) == 3
    p3d = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.1, 0, 0]])
    r = oracle.search_single_multi(p3d, None, None, [1.5, 3.0], True)
    assert len(r[0][0]) == 2 and len(r[1][0]) == 1


def test_single_force_cell_list(oracle):
    p = np.array([[1.0, 1, 1], [2.0, 2, 2]])
    ei, _, _ = oracle.search_single(p, np.eye(3) * 10, [1, 1, 1], 6.0, parallel=False)
    assert len(ei) == 1
    r = oracle.search_single_multi(p, np.eye(3) * 10, [1, 1, 1], [6.0], False)
    assert len(r[0][0]) == 1


def test_single_empty(oracle):
    ei, _, _ = oracle.search_single(np.zeros((0, 3)), None, None, 1.0)
    assert ei.size == 0
    assert oracle.search_single_multi(np.zeros((0, 3)), None, None, [1.0], False)[0][0].size == 0
    assert oracle.search_single_multi(np.zeros((1, 3)), None, None, [], False) == []


def test_batch_basic(oracle):
    p = np.array([[1.0, 1, 1], [1.2, 1, 1], [5.0, 5, 5], [5.2, 5, 5]])
    c = (np.eye(3) * 10, [1, 1, 1])
    ei, ej, _ = oracle.search_batch(p, [0, 0, 1, 1], [c, c], 0.5)
    assert sorted(zip(ei.tolist(), ej.tolist())) == [(0, 1), (2, 3)]
    p2 = p.copy()
    p2[3] = [5.5, 5, 5]
    r = oracle.search_batch_multi(p2, [0, 0, 1, 1], [c, c], [0.3, 0.6], False)
    assert len(r[0][0]) == 1 and len(r[1][0]) == 2
    pa = np.array([[1.0, 1, 1], [1.2, 1, 1], [10.0, 10, 10], [10.2, 10, 10]])
    assert len(oracle.search_batch(pa, [0, 0, 1, 1], [None, None], 0.5)[0]) == 2


def test_batch_errors(oracle):
    with pytest.raises(ValueError, match="Expected at least 1 cells, but got 0"):
        oracle.search_batch(np.zeros((1, 3)), [0], [], 1.0)
    with pytest.raises(ValueError, match="Expected at least 1 cells, but got 0"):
        oracle.search_batch_multi(np.zeros((1, 3)), [0], [], [1.0], False)
    with pytest.raises(ValueError, match="positions and batch must have the same length"):
        oracle.search_batch(np.zeros((1, 3)), [0, 1], [None], 1.0)


def test_batch_is_concatenation_of_singles(oracle):
    rng = np.random.default_rng(0)
    parts, cells = [], []
    for s in range(6):
        n = int(rng.integers(2, 400))
        parts.append(rng.uniform(0, 12, size=(n, 3)))
        cells.append(None if s % 2 else (np.diag(rng.uniform(9, 14, 3)), [1, 1, 1]))
    p = np.concatenate(parts)
    b = np.repeat(np.arange(6), [len(x) for x in parts])
    got = oracle.search_batch(p, b, cells, 4.0)
    off = 0
    ei, ej, ss = [], [], []
    for x, c in zip(parts, cells):
        r = oracle.search_single(x, None if c is None else c[0], None if c is None else c[1], 4.0)
        ei.append(r[0] + off)
        ej.append(r[1] + off)
        ss.append(r[2])
        off += len(x)
    want = (np.concatenate(ei), np.concatenate(ej), np.concatenate(ss))
    for a, w in zip(got, want):
        np.testing.assert_array_equal(a, w)  # same order too: systems concatenated


# ---- python tests (tests/*.py) ---------------------------------------------
def test_basic_pbc_shift(oracle):
    p = np.array([[1.0, 1.0, 1.0], [1.5, 1.0, 1.0], [19.5, 1.0, 1.0]])
    ei, ej, s = oracle.search_single(p, np.eye(3) * 20, [1, 1, 1], 2.0)
    d = {(i, j): tuple(x) for i, j, x in zip(ei.tolist(), ej.tolist(), s.tolist())}
    assert d == {(0, 1): (0, 0, 0), (0, 2): (-1, 0, 0)}


def _si():
    pos, rows, pbc = read_xyz("si_bulk.xyz")
    return pos, rows.T


def test_silicon_16_edges(oracle):
    pos, H = _si()
    assert len(oracle.search_single(pos, H, [1, 1, 1], 2.5)[0]) == 16


def test_single_atom_six_self_images(oracle):
    ei, ej, s = oracle.search_single(np.zeros((1, 3)), np.eye(3) * 5, [1, 1, 1], 6.0)
    assert len(ei) == 6 and not np.any(np.all(s == 0, axis=1))
    assert sorted(map(tuple, s.tolist())) == sorted(
        [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)])


def test_mixed_pbc(oracle):
    p = np.array([[5.0, 5.0, 0.5], [5.0, 5.0, 9.5]])
    assert len(oracle.search_single(p, np.eye(3) * 10, [1, 1, 0], 2.0)[0]) == 0
    ei, ej, s = oracle.search_single(p, np.eye(3) * 10, [1, 1, 1], 2.0)
    assert len(ei) == 1 and s[0].tolist() == [0, 0, -1]


def _cu_fcc(rep):
    a = 3.6
    prim = np.array([[0, a / 2, a / 2], [a / 2, 0, a / 2], [a / 2, a / 2, 0]])
    pts = np.array([(i, j, k) for i in range(rep) for j in range(rep) for k in range(rep)], float)
    return pts @ prim, prim * rep  # rows = lattice vectors (symmetric here)


def test_disjoint_shells(oracle):
    pos, rows = _cu_fcc(3)
    cum = oracle.search_single_multi(pos, rows, [1, 1, 1], [3.0, 4.0, 5.0], False)
    dis = oracle.search_single_multi(pos, rows, [1, 1, 1], [3.0, 4.0, 5.0], True)
    n = [len(r[0]) for r in cum]
    d = [len(r[0]) for r in dis]
    assert n[0] < n[1] < n[2]
    assert d == [n[0], n[1] - n[0], n[2] - n[1]]
    # each (i, j, S) edge of the cumulative list lies in exactly one shell
    trip = [set(zip(*(a.tolist() for a in canon(*r)[:2]), map(tuple, canon(*r)[2].tolist())))
            for r in dis]
    for a in range(3):
        for b in range(a + 1, 3):
            assert trip[a].isdisjoint(trip[b]), (a, b)
    full = set(zip(*(a.tolist() for a in canon(*cum[2])[:2]),
                   map(tuple, canon(*cum[2])[2].tolist())))
    assert trip[0] | trip[1] | trip[2] == full
    pos2, rows2 = _cu_fcc(2)
    res = oracle.search_single_multi(pos2, rows2, [1, 1, 1], [5.0, 3.0], True)
    small = len(oracle.search_single(pos2, rows2, [1, 1, 1], 3.0)[0])
    total = len(oracle.search_single(pos2, rows2, [1, 1, 1], 5.0)[0])
    assert len(res[1][0]) == small and len(res[0][0]) == total - small


def test_twenty_cutoffs(oracle):
    p = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    res = oracle.search_single_multi(p, None, None, [0.1 * i for i in range(1, 21)], False)
    assert len(res) == 20 and len(res[10][0]) == 1 and len(res[8][0]) == 0


def test_hundred_cutoffs_match_singles(oracle):
    # any number of cutoffs (search.rs:344-361): every cumulative list of a
    # 100-cutoff call is the single-cutoff list at that cutoff, and the
    # disjoint shells partition the largest one
    pos, rows = _cu_fcc(2)
    rng = np.random.default_rng(3)
    cut = [float(c) for c in rng.uniform(2.0, 6.0, size=97)] + [3.0, 3.0, 5.5]
    cum = oracle.search_single_multi(pos, rows, [1, 1, 1], cut, False)
    dis = oracle.search_single_multi(pos, rows, [1, 1, 1], cut, True)
    assert len(cum) == len(dis) == 100
    for q in (0, 17, 64, 65, 97, 98, 99):
        assert_same_edges(cum[q], oracle.search_single(pos, rows, [1, 1, 1], cut[q]), str(q))
    assert sum(len(r[0]) for r in dis) == len(cum[int(np.argmax(cut))][0])


# ---- logged counts for the reference's fixtures ----------------------------
# benchmarks/benchmark_EPYC7R13.txt lines 7-98 (self-image edges excluded,
# comprehensive_benchmark.py:96-100)
LOGGED = [
    ("isolated_100.xyz", 6.0, 2269, 2269), ("isolated_100.xyz", 14.0, 4950, 4950),
    ("isolated_two_clusters.xyz", 6.0, 90, 90), ("isolated_two_clusters.xyz", 14.0, 190, 190),
    ("isolated_1000.xyz", 6.0, 39219, 39219), ("isolated_1000.xyz", 14.0, 274621, 274621),
    ("ethanol_1000.xyz", 6.0, 57840, 57840), ("ethanol_1000.xyz", 14.0, 680982, 480535),
    ("si_bulk.xyz", 6.0, 160, 28), ("si_bulk.xyz", 14.0, 2064, 28),
    ("ethanol_10000.xyz", 6.0, 568606, None), ("isolated_10000.xyz", 6.0, 482454, None),
]


@pytest.mark.parametrize("name,rc,edges,unique", LOGGED)
def test_logged_counts(oracle, name, rc, edges, unique):
    pos, rows, pbc = read_xyz(name)
    H = None if rows is None else rows.T
    ei, ej, s = oracle.search_single(pos, H, pbc if rows is not None else None, rc)
    ns = ei != ej
    assert int(ns.sum()) == edges
    if unique is not None:
        assert len(set(zip(ei[ns].tolist(), ej[ns].tolist()))) == unique


def test_si_self_images(oracle):
    """Si at rc 6 / 14 also emits 48 / 640 self-image edges (both +-S)."""
    pos, H = _si()
    for rc, n_self in ((6.0, 48), (14.0, 640)):
        ei, ej, s = oracle.search_single(pos, H, [1, 1, 1], rc)
        self_s = s[ei == ej]
        assert len(self_s) == n_self
        assert set(map(tuple, self_s.tolist())) == set(map(tuple, (-self_s).tolist()))


# ---- invariants on random inputs -------------------------------------------
def _multi_image_bruteforce(pos, H, rc):
    """Textbook all-images search: every (i, j, S) with |r_j + H S - r_i| < rc,
    i < j or (i == j, S != 0)."""
    inv = np.linalg.inv(H)
    w = np.min(1.0 / np.linalg.norm(inv, axis=1))
    # positions need not lie in the cell: widen the image range by their
    # fractional spread
    frac = pos @ inv.T
    spread = float((frac.max(0) - frac.min(0)).max()) if len(pos) else 0.0
    m = int(np.ceil(rc / w + spread)) + 1
    g = np.arange(-m, m + 1)
    S = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    off = S @ H.T
    out = []
    n = len(pos)
    for i in range(n):
        d = pos[None, :, :] + off[:, None, :] - pos[i]
        r2 = (d ** 2).sum(-1)
        si, jj = np.nonzero(r2 < rc * rc)
        for a, j in zip(si, jj):
            if i < j or (i == j and np.any(S[a] != 0)):
                out.append((i, j, *S[a]))
    out = np.array(out, dtype=np.int64).reshape(-1, 5)
    return out[:, 0], out[:, 1], out[:, 2:].astype(np.int32)


@settings(max_examples=25, deadline=None)
@given(box=st.floats(10.0, 20.0), rc=st.floats(1.0, 3.0), n=st.integers(2, 50),
       seed=st.integers(0, 2 ** 31))
def test_property_pbc_vs_bruteforce(oracle, box, rc, n, seed):
    """search.rs:1758-1800: cubic PBC cell vs brute force (pairs only)."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 20.0, size=(n, 3)) % box
    oracle.set_threshold("brute_force", 0)
    try:
        ei, ej, _ = oracle.search_single(pos, np.eye(3) * box, [1, 1, 1], rc)
    finally:
        oracle.set_threshold("brute_force", 1000)
    want = [(i, j) for i in range(n) for j in range(i + 1, n)
            if (((pos[j] - pos[i]) - box * np.round((pos[j] - pos[i]) / box)) ** 2).sum()
            < rc * rc]
    assert sorted(zip(ei.tolist(), ej.tolist())) == sorted(want)


@settings(max_examples=15, deadline=None)
@given(n=st.integers(1, 40), rc=st.floats(2.0, 9.0), seed=st.integers(0, 2 ** 31),
       shear=st.floats(-0.4, 0.4))
def test_property_triclinic_all_images(oracle, n, rc, seed, shear):
    """Cell-list path (all images, self images) vs a textbook multi-image search."""
    rng = np.random.default_rng(seed)
    H = np.array([[6.0, shear * 6.0, 0.3], [0.0, 5.5, shear * 2.0], [0.0, 0.0, 6.5]])
    pos = rng.uniform(-3, 9, size=(n, 3))
    got = oracle.search_single(pos, H, [1, 1, 1], rc)
    want = _multi_image_bruteforce(pos, H, rc)
    G = set(zip(got[0].tolist(), got[1].tolist(), map(tuple, got[2].tolist())))
    Wt = set(zip(want[0].tolist(), want[1].tolist(), map(tuple, want[2].tolist())))
    # identical except for pairs within 1e-9 of the cutoff, where the textbook
    # expression and the reference's wrapped-coordinate one may round apart
    for i, j, S in G ^ Wt:
        r = pos[j] + np.array(S) @ H.T - pos[i]
        assert abs(np.sqrt((r ** 2).sum()) - rc) < 1e-9, (i, j, S)
    # every emitted edge is within the cutoff
    rij = pos[got[1]] + got[2] @ H.T - pos[got[0]]
    assert np.all((rij ** 2).sum(1) < rc * rc * (1 + 1e-12))


@settings(max_examples=20, deadline=None)
@given(n=st.integers(2, 100), rc=st.floats(1.0, 5.0), seed=st.integers(0, 2 ** 31))
def test_property_isolated_vs_plain(oracle, n, rc, seed):
    """search.rs:1802-1856: isolated brute force in a 1000 A non-PBC cell."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-10, 10, size=(n, 3))
    ei, ej, s = oracle.search_single(pos, np.eye(3) * 1000, [0, 0, 0], rc)
    want = [(i, j) for i in range(n) for j in range(i + 1, n)
            if np.sqrt(((pos[i] - pos[j]) ** 2).sum()) < rc]
    assert sorted(zip(ei.tolist(), ej.tolist())) == sorted(want)
    assert np.all(s == 0)


def test_packed_edge_set_comparison_catches_a_changed_edge(oracle):
    # the fast comparison used for the multi-million-edge serving lists must
    # see a single changed shift component and a swapped (i, j)
    from conftest import assert_same_edge_sets
    pos, rows = _cu_fcc(2)
    r = oracle.search_single(pos, rows, [1, 1, 1], 6.0)
    assert_same_edge_sets(r, r, "same")
    s2 = r[2].copy()
    s2[7, 2] += 1
    with pytest.raises(AssertionError):
        assert_same_edge_sets((r[0], r[1], s2), r, "shift")
    i2, j2 = r[0].copy(), r[1].copy()
    k = int(np.nonzero(i2 != j2)[0][0])
    i2[k], j2[k] = j2[k], i2[k]
    with pytest.raises(AssertionError):
        assert_same_edge_sets((i2, j2, r[2]), r, "swap")
