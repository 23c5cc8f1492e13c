"""GPU parity: the HIP engine against the CPU oracle, through the C ABI.

Bit-exact bar: after the canonical sort on (i, j, S), edge_index and shift
arrays are identical to the oracle's (which restates the reference and is
pinned by the reference's own known answers, see test_oracle_reference.py).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import as_tuple, assert_same_edges
from xyzio import read_xyz

pytestmark = pytest.mark.gpu


def _print_case(*a):
    print("[case]", *a, flush=True)


def _gpu_single(nr, pos, H, pbc, rc):
    cell = None if H is None else nr.PyCell(np.asarray(H).tolist(), pbc)
    return as_tuple(nr.build_neighborlists(cell, pos, rc))


FIXTURES = [
    ("si_bulk.xyz", 6.0), ("si_bulk.xyz", 14.0), ("isolated_100.xyz", 6.0),
    ("isolated_100.xyz", 14.0), ("isolated_two_clusters.xyz", 6.0),
    ("isolated_1000.xyz", 6.0), ("isolated_1000.xyz", 14.0), ("ethanol_1000.xyz", 6.0),
    ("ethanol_1000.xyz", 14.0),
]


@pytest.mark.parametrize("name,rc", FIXTURES)
def test_fixture_vs_oracle(nr, oracle, name, rc):
    _print_case("fixture", name, rc)
    pos, rows, pbc = read_xyz(name)
    H = None if rows is None else rows.T
    want = oracle.search_single(pos, H, pbc if rows is not None else None, rc)
    got = _gpu_single(nr, pos, H, pbc, rc)
    assert_same_edges(got, want, f"{name} rc={rc}")


# counts the reference logged for its fixtures (benchmarks/benchmark_EPYC7R13.txt),
# self-image (i == j) edges excluded as in comprehensive_benchmark.py:96-100
LOGGED = [("ethanol_10000.xyz", 6.0, 568606), ("ethanol_10000.xyz", 14.0, 6807370),
          ("isolated_10000.xyz", 6.0, 482454)]


@pytest.mark.parametrize("name,rc,count", LOGGED)
def test_large_fixture_logged_counts(nr, name, rc, count):
    _print_case("logged", name, rc)
    pos, rows, pbc = read_xyz(name)
    H = None if rows is None else rows.T
    ei, ej, s = _gpu_single(nr, pos, H, pbc, rc)
    assert int((ei != ej).sum()) == count
    assert np.all((ei < ej) | ((ei == ej) & np.any(s != 0, axis=1)))


@pytest.mark.parametrize("cell_div", [0, 1, 2, 3])
@pytest.mark.parametrize("fp32", [0, 1])
def test_configs_vs_oracle(nr, oracle, cell_div, fp32):
    from neighborlist_rs import workloads as W
    from neighborlist_rs._lib import lib
    was = (lib.nl_get_autotune(), lib.nl_get_cell_div(), lib.nl_get_fp32_filter())
    lib.nl_set_autotune(0)  # the settings under test reach every request
    lib.nl_set_cell_div(cell_div)
    lib.nl_set_fp32_filter(fp32)
    try:
        cases = [("c1", *W.c1_isolated())]
        p, H, rcs = W.c2_ethanol()
        cases += [("c2", p, H, rc) for rc in rcs]
        p, H, rcs = W.c3_triclinic()
        cases += [("c3", p, H, 6.0)]
        for name, p, H, rc in cases:
            _print_case("config", name, rc, "div", cell_div, "fp32", fp32)
            want = oracle.search_single(p, H, None if H is None else [1, 1, 1], rc)
            got = _gpu_single(nr, p, H, [True] * 3, rc)
            assert_same_edges(got, want, f"{name} rc={rc} div={cell_div} fp32={fp32}")
    finally:
        lib.nl_set_autotune(was[0])
        lib.nl_set_cell_div(was[1])
        lib.nl_set_fp32_filter(was[2])


def test_c3_rc14_vs_oracle(nr, oracle):
    from neighborlist_rs import workloads as W
    p, H, _ = W.c3_triclinic()
    _print_case("config c3 rc14")
    want = oracle.search_single(p, H, [1, 1, 1], 14.0)
    got = _gpu_single(nr, p, H, [True] * 3, 14.0)
    assert_same_edges(got, want, "c3 rc=14")


@pytest.mark.parametrize("disjoint", [False, True])
def test_multi_vs_oracle(nr, oracle, disjoint):
    from neighborlist_rs import workloads as W
    p, H, _ = W.c2_ethanol()
    cut = [14.0, 5.0, 6.0]
    _print_case("multi c2", disjoint)
    want = oracle.search_single_multi(p, H, [1, 1, 1], cut, disjoint)
    got = nr.build_neighborlists_multi(nr.PyCell(H.tolist()), p, cut, disjoint=disjoint)
    for k in range(len(cut)):
        assert_same_edges(as_tuple(got[k]), want[k], f"list {k}")


@pytest.mark.parametrize("disjoint", [False, True])
def test_ragged_batch_multi_vs_oracle(nr, oracle, disjoint):
    from neighborlist_rs import workloads as W
    p, sizes, cells, cut = W.c4_ragged_batch(n_systems=64, nmax=1200)
    batch, rows = W.batch_arrays(sizes, cells)
    _print_case("batch multi", len(sizes), disjoint)
    ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
    want = oracle.search_batch_multi(p, batch, ocells, cut, disjoint)
    got = nr.build_neighborlists_batch_multi(p, batch, rows, list(cut), disjoint=disjoint)
    for k in range(len(cut)):
        assert_same_edges(as_tuple(got[k]), want[k], f"list {k}")


def test_ragged_batch_single_vs_oracle(nr, oracle):
    from neighborlist_rs import workloads as W
    p, sizes, cells, _ = W.c4_ragged_batch(n_systems=48, nmax=1500, seed=11)
    batch, rows = W.batch_arrays(sizes, cells)
    _print_case("batch single", len(sizes))
    ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
    want = oracle.search_batch(p, batch, ocells, 6.0)
    got = as_tuple(nr.build_neighborlists_batch(p, batch, rows, 6.0))
    assert_same_edges(got, want, "batch rc=6")


def test_bruteforce_row_order_matches_reference(nr, oracle):
    """Brute-force systems come out in the reference's row-major order."""
    rng = np.random.default_rng(7)
    p = rng.uniform(0, 12, size=(300, 3))
    _print_case("bf order")
    want = oracle.search_single(p, None, None, 3.0)
    got = _gpu_single(nr, p, None, None, 3.0)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("thr", [0, 1000, 100000])
def test_dispatch_threshold(nr, oracle, thr):
    """brute_force_threshold keeps its dispatch meaning (config.rs:8, single.rs:64)."""
    rng = np.random.default_rng(8)
    H = np.diag([14.0, 15.0, 16.0])
    p = rng.uniform(0, 14, size=(500, 3))
    old = nr.get_brute_force_threshold()
    nr.set_brute_force_threshold(thr)
    oracle.set_threshold("brute_force", thr)
    try:
        _print_case("threshold", thr)
        for pbc in ([True, True, True], [True, False, True], [False, False, False]):
            want = oracle.search_single(p, H, pbc, 5.0)
            got = _gpu_single(nr, p, H, pbc, 5.0)
            assert_same_edges(got, want, f"thr={thr} pbc={pbc}")
    finally:
        nr.set_brute_force_threshold(old)
        oracle.set_threshold("brute_force", old)


def test_small_box_many_images(nr, oracle):
    """rc far beyond the box: many periodic images (search.rs:1862-1882)."""
    _print_case("many images")
    for H, p, rc in [(np.eye(3), np.array([[0.5, 0.5, 0.5]]), 10.0),
                     (np.eye(3) * 5.0, np.array([[0.0, 0.0, 0.0]]), 6.0),
                     (np.diag([4.1, 4.4, 4.7]), np.random.default_rng(3).uniform(0, 4, (8, 3)),
                      14.0)]:
        want = oracle.search_single(p, H, [1, 1, 1], rc)
        got = _gpu_single(nr, p, H, [True] * 3, rc)
        assert_same_edges(got, want, f"rc={rc}")


def test_distances_fp64(nr, oracle):
    """FP64 |r_ij| emitted with the edges equals a float64 recomputation."""
    from neighborlist_rs import workloads as W
    from neighborlist_rs._lib import lib, dp, i64p, i32p
    p, H, _ = W.c2_ethanol()
    sizes = np.array([p.shape[0]], np.int64)
    cut = np.array([6.0])
    has = np.ones(1, np.uint8)
    pb = np.ones(3, np.uint8)
    Hf = np.ascontiguousarray(H.reshape(9))
    r = C.c_void_p()
    assert lib.nl_search_systems(p.ctypes.data_as(dp), p.shape[0], sizes.ctypes.data_as(i64p), 1,
                                 Hf.ctypes.data_as(dp), has.ctypes.data_as(C.POINTER(C.c_uint8)),
                                 pb.ctypes.data_as(C.POINTER(C.c_uint8)),
                                 cut.ctypes.data_as(dp), 1, 0, 0, 1, 0, 1, C.byref(r)) == 0
    m = lib.nl_result_num_edges(r, 0)
    ei = np.empty(m, np.int64)
    ej = np.empty(m, np.int64)
    s = np.empty((m, 3), np.int32)
    d = np.empty(m)
    assert lib.nl_result_copy(r, 0, ei.ctypes.data_as(i64p), ej.ctypes.data_as(i64p),
                              s.ctypes.data_as(i32p)) == 0
    assert lib.nl_result_copy_dist(r, 0, d.ctypes.data_as(dp)) == 0
    lib.nl_result_free(r)
    rij = p[ej] + s @ H.T - p[ei]
    ref = np.sqrt((rij ** 2).sum(1))
    np.testing.assert_allclose(d, ref, rtol=0, atol=1e-12)
    assert np.all(d < 6.0)


@pytest.mark.parametrize("group,parts", [(1, 1), (3, 2), (16, 4), (64, 3)])
def test_work_split_does_not_change_results(nr, oracle, group, parts):
    from neighborlist_rs import workloads as W
    from neighborlist_rs._lib import lib
    g0, p0 = C.c_int(), C.c_int()
    lib.nl_get_work_split(C.byref(g0), C.byref(p0))
    was = lib.nl_get_autotune()
    lib.nl_set_autotune(0)  # the split under test reaches every request
    lib.nl_set_work_split(group, parts)
    try:
        p, H, _ = W.c2_ethanol()
        for rc in (6.0, 14.0):
            _print_case("split", group, parts, rc)
            want = oracle.search_single(p, H, [1, 1, 1], rc)
            got = _gpu_single(nr, p, H, [True] * 3, rc)
            assert_same_edges(got, want, f"G={group} P={parts} rc={rc}")
        p, sizes, cells, cut = W.c4_ragged_batch(n_systems=24, nmax=900, seed=12)
        batch, rows = W.batch_arrays(sizes, cells)
        ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
        want = oracle.search_batch_multi(p, batch, ocells, cut, False)
        got = nr.build_neighborlists_batch_multi(p, batch, rows, list(cut))
        for k in range(len(cut)):
            assert_same_edges(as_tuple(got[k]), want[k], f"G={group} P={parts} list {k}")
    finally:
        lib.nl_set_work_split(g0.value, p0.value)
        lib.nl_set_autotune(was)


@pytest.mark.parametrize("group,parts", [(1, 1), (16, 2), (64, 4)])
def test_grid_bound_dense_and_dilute_cells(nr, oracle, group, parts):
    # The pair-kernel grids stop at a bound on the work items: per system
    # ceil(n / group) + min(n, bins) home groups.  A very dense periodic box
    # (about 30 atoms per bin: several groups per bin) and a dilute periodic
    # gas (many more bins than atoms: one group per atom) sit at either end
    # of that bound, beside an isolated cluster; every list must match the
    # oracle whatever the split.
    from neighborlist_rs import workloads as W
    from neighborlist_rs._lib import lib
    rng = np.random.default_rng(77)
    dense = rng.uniform(0.0, 6.0, size=(800, 3))
    gas = rng.uniform(0.0, 40.0, size=(60, 3))
    cluster = rng.normal(scale=4.0, size=(500, 3))
    pos = np.vstack([dense, gas, cluster])
    sizes = np.array([800, 60, 500], dtype=np.int64)
    cells = [np.eye(3) * 6.0, np.eye(3) * 40.0, None]
    cut = [3.0, 5.0]
    g0, p0 = C.c_int(), C.c_int()
    lib.nl_get_work_split(C.byref(g0), C.byref(p0))
    was = lib.nl_get_autotune()
    lib.nl_set_autotune(0)
    lib.nl_set_work_split(group, parts)
    try:
        _print_case("grid bound", group, parts)
        batch, rows = W.batch_arrays(sizes, cells)
        ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
        want = oracle.search_batch_multi(pos, batch, ocells, cut, False)
        got = nr.build_neighborlists_batch_multi(pos, batch, rows, list(cut))
        for k in range(len(cut)):
            assert_same_edges(as_tuple(got[k]), want[k], f"G={group} P={parts} list {k}")
    finally:
        lib.nl_set_work_split(g0.value, p0.value)
        lib.nl_set_autotune(was)


@pytest.mark.parametrize("disjoint", [False, True])
def test_twenty_cutoffs_vs_oracle(nr, oracle, disjoint):
    """More lists than one pair-kernel launch holds (8): several launches per pass,
    unsorted input order, duplicates (search.rs:354-360 stable sort)."""
    from neighborlist_rs import workloads as W
    rng = np.random.default_rng(20)
    cut = [float(c) for c in rng.uniform(2.0, 9.0, size=18)] + [4.0, 4.0]
    p, H, _ = W.c2_ethanol()
    _print_case("twenty cutoffs single", disjoint)
    want = oracle.search_single_multi(p, H, [1, 1, 1], cut, disjoint)
    got = nr.build_neighborlists_multi(nr.PyCell(H.tolist()), p, cut, disjoint=disjoint)
    for k in range(len(cut)):
        assert_same_edges(as_tuple(got[k]), want[k], f"single list {k} rc={cut[k]:.3f}")
    p, sizes, cells, _ = W.c4_ragged_batch(n_systems=16, nmax=1100, seed=21)
    batch, rows = W.batch_arrays(sizes, cells)
    _print_case("twenty cutoffs batch", disjoint)
    ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
    want = oracle.search_batch_multi(p, batch, ocells, cut, disjoint)
    got = nr.build_neighborlists_batch_multi(p, batch, rows, cut, disjoint=disjoint)
    for k in range(len(cut)):
        assert_same_edges(as_tuple(got[k]), want[k], f"batch list {k} rc={cut[k]:.3f}")


@pytest.mark.gpu
@pytest.mark.parametrize("disjoint", [False, True])
def test_hundred_cutoffs_vs_oracle(nr, oracle, disjoint):
    """The reference takes any number of cutoffs (search.rs:344-361,
    python_api.rs:235-260): 100 lists, unsorted with repeats, thirteen
    pair-kernel launches per pass; single system and a ragged batch."""
    from neighborlist_rs import workloads as W
    rng = np.random.default_rng(100)
    cut = [float(c) for c in rng.uniform(1.5, 7.0, size=96)] + [3.0, 3.0, 6.5, 2.0]
    p, H, _ = W.c2_ethanol()
    _print_case("hundred cutoffs single", disjoint)
    want = oracle.search_single_multi(p, H, [1, 1, 1], cut, disjoint)
    got = nr.build_neighborlists_multi(nr.PyCell(H.tolist()), p, cut, disjoint=disjoint)
    assert len(got) == len(cut)
    total = 0
    for k in range(len(cut)):
        assert_same_edges(as_tuple(got[k]), want[k], f"single list {k} rc={cut[k]:.3f}")
        total += got[k]["edge_index"].shape[1]
    assert total > 0
    p, sizes, cells, _ = W.c4_ragged_batch(n_systems=8, nmax=900, seed=101)
    batch, rows = W.batch_arrays(sizes, cells)
    _print_case("hundred cutoffs batch", disjoint)
    ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
    want = oracle.search_batch_multi(p, batch, ocells, cut, disjoint)
    got = nr.build_neighborlists_batch_multi(p, batch, rows, cut, disjoint=disjoint)
    for k in range(len(cut)):
        assert_same_edges(as_tuple(got[k]), want[k], f"batch list {k} rc={cut[k]:.3f}")


@pytest.mark.parametrize("order", [0, 1])
def test_bin_order_linear_and_morton(nr, oracle, order):
    """Bins keyed by linear index (0) or by Morton rank (1, the default,
    search.rs:126-146): the same edge sets either way (a 10k-atom triclinic
    system, the many-image C2 at rc 14, a ragged multi-cutoff batch with
    disjoint shells)."""
    from neighborlist_rs import _lib, workloads as W
    old = _lib.lib.nl_get_bin_order()
    _lib.lib.nl_set_bin_order(order)
    try:
        p3, H3, _ = W.c3_triclinic()
        _print_case("bin order", order, "c3 rc6")
        assert_same_edges(_gpu_single(nr, p3, H3, [True] * 3, 6.0),
                          oracle.search_single(p3, H3, [1, 1, 1], 6.0), f"order {order} c3")
        p2, H2, _ = W.c2_ethanol()
        _print_case("bin order", order, "c2 rc14")
        assert_same_edges(_gpu_single(nr, p2, H2, [True] * 3, 14.0),
                          oracle.search_single(p2, H2, [1, 1, 1], 14.0), f"order {order} c2")
        p, sizes, cells, cut = W.c4_ragged_batch(n_systems=24, nmax=2048, seed=77)
        batch, rows = W.batch_arrays(sizes, cells)
        _print_case("bin order", order, "ragged batch")
        ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
        want = oracle.search_batch_multi(p, batch, ocells, list(cut), True)
        got = nr.build_neighborlists_batch_multi(p, batch, rows, list(cut), disjoint=True)
        for k in range(len(cut)):
            assert_same_edges(as_tuple(got[k]), want[k], f"order {order} batch list {k}")
    finally:
        _lib.lib.nl_set_bin_order(old)


def test_parallel_flag_does_not_change_results(nr, oracle):
    """parallel=False selects the reference's serial search (search.rs:275-289);
    the edge set is the same."""
    from neighborlist_rs import workloads as W
    p, H, _ = W.c2_ethanol()
    _print_case("parallel flag")
    cell = nr.PyCell(H.tolist())
    a = as_tuple(nr.build_neighborlists(cell, p, 6.0, parallel=True))
    b = as_tuple(nr.build_neighborlists(cell, p, 6.0, parallel=False))
    assert_same_edges(a, b, "parallel vs serial")
    assert_same_edges(b, oracle.search_single(p, H, [1, 1, 1], 6.0, parallel=False), "serial")


def test_full_ragged_batch_vs_oracle(nr, oracle):
    """The full 256-system batch of the benchmark (113,559 atoms, 56 radix tiles:
    the multi-tile sort path), cutoffs {5, 6, 14} in one pass."""
    from neighborlist_rs import workloads as W
    p, sizes, cells, cut = W.c4_ragged_batch()
    batch, rows = W.batch_arrays(sizes, cells)
    _print_case("full ragged batch", len(sizes), p.shape[0])
    ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
    want = oracle.search_batch_multi(p, batch, ocells, list(cut), False)
    got = nr.build_neighborlists_batch_multi(p, batch, rows, list(cut))
    for k in range(len(cut)):
        assert_same_edges(as_tuple(got[k]), want[k], f"list {k}")


def test_large_single_system_vs_oracle(nr, oracle):
    """One periodic system of 60,000 atoms (30 radix tiles), rc 5."""
    from neighborlist_rs import workloads as W
    rng = np.random.default_rng(60)
    p, H = W.random_system(rng, 60000, 2)
    _print_case("large single", p.shape[0])
    want = oracle.search_single(p, H, [1, 1, 1], 5.0)
    got = _gpu_single(nr, p, H, [True] * 3, 5.0)
    assert_same_edges(got, want, "60k triclinic rc=5")


@pytest.mark.parametrize("sizes", [(4096, 1, 4000, 37), (4096, 4097, 1, 2, 700)])
def test_segmented_sort_boundary_vs_oracle(nr, oracle, sizes):
    """Batches whose systems all have at most 4096 atoms are sorted one
    system per block (segmented sort); one larger system sends the whole
    batch through the global radix sort.  Both sides of that boundary, with
    periodic, isolated and single-atom systems."""
    from neighborlist_rs import workloads as W
    rng = np.random.default_rng(sum(sizes))
    ps, cells = [], []
    for s, n in enumerate(sizes):
        p, H = W.random_system(rng, n, s % 3)
        ps.append(p)
        cells.append(H)
    p = np.ascontiguousarray(np.concatenate(ps))
    batch, rows = W.batch_arrays(np.array(sizes), cells)
    _print_case("segmented sort boundary", sizes)
    ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
    cut = [3.5, 6.0]
    want = oracle.search_batch_multi(p, batch, ocells, cut, False)
    got = nr.build_neighborlists_batch_multi(p, batch, rows, cut)
    for k in range(len(cut)):
        assert_same_edges(as_tuple(got[k]), want[k], f"{sizes} list {k}")


def _search_ex_device(lib, L, dpos, n, sizes, H9, has, pb, cut, multi, disjoint, opts):
    """nl_search_ex on device positions: every array of every list as numpy."""
    r = C.c_void_p()
    assert lib.nl_search_ex(C.c_void_p(dpos), n, sizes.ctypes.data_as(L.i64p), sizes.size,
                            H9.ctypes.data_as(L.dp), has.ctypes.data_as(L.u8p),
                            pb.ctypes.data_as(L.u8p), cut.ctypes.data_as(L.dp), cut.size,
                            int(disjoint), int(multi), opts | 1, None, C.byref(r)) == 0, \
        L.last_error()
    out = []
    for q in range(cut.size):
        m = int(lib.nl_result_num_edges(r, q))
        ei = np.zeros((2, m), np.int64)
        sh = np.zeros((m, 3), np.int32)
        if m:
            assert lib.nl_result_copy(r, q, ei[0].ctypes.data_as(L.i64p),
                                      ei[1].ctypes.data_as(L.i64p),
                                      sh.ctypes.data_as(L.i32p)) == 0
        d = np.zeros(m)
        if m and opts & 8:
            assert lib.nl_result_copy_dist(r, q, d.ctypes.data_as(L.dp)) == 0
        rng = np.zeros(2 * sizes.size * cut.size, np.int64)
        modes = np.zeros(sizes.size, np.int32)
        if opts & 4:
            assert lib.nl_result_system_ranges(r, rng.ctypes.data_as(L.i64p),
                                               modes.ctypes.data_as(L.i32p)) == 0
        out.append((ei, sh, d, rng, modes))
    lib.nl_result_free(r)
    return out


@pytest.mark.parametrize("tuning", [(0, 8, 2), (3, 16, 4), (2, 3, 1), (1, 64, 3)])
def test_fused_build_identical_to_stages(nr, oracle, tuning):
    """Batches whose systems all have at most 12288 atoms build in one launch
    (k_build_systems: plan, counting sort by bin, bin offsets, sorted
    positions, work list; taken for every case here with
    nl_set_fused_build(2)); the separate launches of the general path
    (nl_set_fused_build(0)) must give the same arrays, order included:
    the 256-system ragged batch with three lists, a batch at the 4096-atom
    boundary of the segmented sort with isolated, orthorhombic and triclinic
    systems and single atoms (full lists, distances, ranges), a batch up to
    the 12288-atom limit, and a dense cluster with one far atom (its auto box
    is huge, so a few bins hold every atom); per-system cutoffs are covered
    by the serving tests.  Two cases are also checked against the oracle."""
    from neighborlist_rs import _lib as L, workloads as W
    import bench
    lib = L.lib
    cd, g, parts = tuning
    olds = (lib.nl_get_autotune(), lib.nl_get_cell_div(), lib.nl_get_fused_build())
    og, op = C.c_int(), C.c_int()
    lib.nl_get_work_split(C.byref(og), C.byref(op))
    lib.nl_set_autotune(0)
    lib.nl_set_cell_div(cd)
    lib.nl_set_work_split(g, parts)
    try:
        p4, s4, c4, cut4 = W.c4_ragged_batch()
        rng = np.random.default_rng(4096)
        sizes_b = [4096, 1, 4000, 37, 1, 900]
        ps, cb = [], []
        for s, m in enumerate(sizes_b):
            p, H = W.random_system(rng, m, s % 3)
            ps.append(p)
            cb.append(H)
        pb_ = np.ascontiguousarray(np.concatenate(ps))
        sizes_l = [12288, 5000, 4097, 1, 8000]  # one launch only when forced (mode 2)
        pl, cl = [], []
        for s, m in enumerate(sizes_l):
            p, H = W.random_system(rng, m, (s + 1) % 3)
            pl.append(p)
            cl.append(H)
        pl_ = np.ascontiguousarray(np.concatenate(pl))
        pc = np.ascontiguousarray(np.vstack([rng.uniform(0.0, 12.0, size=(6000, 3)),
                                             [[500.0, 500.0, 500.0]]]))
        cases = [("ragged batch", p4, np.asarray(s4, np.int64), c4, np.array(cut4), True, False,
                  0),
                 ("boundary batch, full lists, distances, ranges", pb_,
                  np.array(sizes_b, np.int64), cb, np.array([3.5, 6.0]), True, True,
                  16 | 8 | 4),
                 ("large systems", pl_, np.array(sizes_l, np.int64), cl, np.array([5.0]), False,
                  False, 4),
                 ("cluster with a far atom", pc, np.array([pc.shape[0]], np.int64), [None],
                  np.array([3.0]), False, False, 0)]
        for name, pos, sizes, cells, cut, multi, disjoint, opts in cases:
            _print_case("fused build", tuning, name)
            H9, has, pbc = bench.cell_arrays(cells)
            d = bench.DevArray(lib, pos)
            try:
                lib.nl_set_fused_build(2)
                a = _search_ex_device(lib, L, d.ptr, pos.shape[0], sizes, H9, has, pbc, cut,
                                      multi, disjoint, opts)
                lib.nl_set_fused_build(0)
                b = _search_ex_device(lib, L, d.ptr, pos.shape[0], sizes, H9, has, pbc, cut,
                                      multi, disjoint, opts)
            finally:
                d.free()
            for q in range(cut.size):
                for x, y in zip(a[q], b[q]):
                    np.testing.assert_array_equal(x, y, err_msg=f"{name} list {q}")
            if name in ("ragged batch", "large systems") and tuning[0] == 0:
                batch, _ = W.batch_arrays(sizes, cells)
                ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
                want = oracle.search_batch_multi(pos, batch, ocells, list(cut), False)
                for q in range(cut.size):
                    ei, sh = a[q][0], a[q][1]
                    assert_same_edges((ei[0], ei[1], sh), want[q], f"{name}: fused list {q}")
    finally:
        lib.nl_set_autotune(olds[0])
        lib.nl_set_cell_div(olds[1])
        lib.nl_set_fused_build(olds[2])
        lib.nl_set_work_split(og.value, op.value)
