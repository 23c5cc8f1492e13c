"""The reference's Python test suite restated against this package on the GPU.

The reference's tests (tests/test_basic.py, test_batch.py, test_disjoint.py,
test_issue_1.py, test_bug_002_fix.py, test_memory_layout.py, test_ase.py,
test_api_extensions.py) call neighborlist_rs with small hand-made systems and
assert exact answers.  The same inputs and the same assertions run here through
the drop-in package (`import neighborlist_rs`), whose searches run on the GPU.
ASE is not installed in this image: its structures (fcc copper, diamond
silicon) are built from their lattice definitions, and `Atoms` objects are a
duck type with get_positions / get_cell / get_pbc, which is all the package
reads (src/python_api.rs:70-98, 214-233)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Atoms:
    """The parts of ase.Atoms the package reads."""

    def __init__(self, positions, cell=None, pbc=False):
        self.positions = np.asarray(positions, dtype=np.float64)
        self.cell = np.zeros((3, 3)) if cell is None else np.asarray(cell, dtype=np.float64)
        self.pbc = np.array([bool(pbc)] * 3 if np.ndim(pbc) == 0 else pbc, dtype=bool)

    def get_positions(self):
        return self.positions.copy()

    def get_cell(self):
        return self.cell.copy()

    def get_pbc(self):
        return self.pbc.copy()


def fcc(a, reps):
    """Conventional fcc cell (4 atoms) repeated reps times: ase.build.bulk(.., 'fcc',
    a=a, cubic=True) * reps, rows are the lattice vectors."""
    basis = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    cells = np.array([[i, j, k] for i in range(reps[0]) for j in range(reps[1])
                      for k in range(reps[2])], dtype=np.float64)
    frac = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    return frac * a, np.diag([a * reps[0], a * reps[1], a * reps[2]])


def diamond(a):
    """Cubic diamond cell (8 atoms)."""
    pos, cell = fcc(a, (1, 1, 1))
    return np.concatenate([pos, pos + a / 4.0]), cell


def pairs(ei):
    return sorted((int(i), int(j)) for i, j in np.asarray(ei).T)


# ---- tests/test_basic.py ----------------------------------------------------
def test_basic_pbc_pairs_and_shifts(nr):
    cell = nr.PyCell([[20.0, 0, 0], [0, 20.0, 0], [0, 0, 20.0]])
    pos = np.array([[1.0, 1.0, 1.0], [1.5, 1.0, 1.0], [19.5, 1.0, 1.0]])
    r = nr.build_neighborlists(cell, pos, 2.0)
    got = {(int(i), int(j)): list(s) for (i, j), s in zip(r["edge_index"].T, r["shift"])}
    assert got == {(0, 1): [0, 0, 0], (0, 2): [-1, 0, 0]}


def test_basic_large_system_half_list(nr):
    cell = nr.PyCell([[20.0, 0, 0], [0, 20.0, 0], [0, 0, 20.0]])
    pos = np.random.default_rng(42).random((1000, 3)) * 20.0
    ei = nr.build_neighborlists(cell, pos, 3.5)["edge_index"]
    assert ei.shape[1] > 0 and np.all(ei[0] < ei[1])


def test_basic_silicon_16_edges(nr):
    pos, rows = diamond(5.43)
    r = nr.build_neighborlists(nr.PyCell(rows.T.tolist()), pos, 2.5)
    assert r["edge_index"].shape[1] == 16  # 8 atoms x 4 neighbours / 2


def test_basic_no_cell_auto_box(nr):
    r = nr.build_neighborlists(None, np.array([[0.0, 0, 0], [1.0, 0, 0]]), 2.0)
    assert pairs(r["edge_index"]) == [(0, 1)] and np.all(r["shift"] == 0)


def test_basic_multi_equals_singles(nr):
    cell = nr.PyCell([[10.0, 0, 0], [0, 10.0, 0], [0, 0, 10.0]])
    pos = np.array([[1.0, 1.0, 1.0], [2.0, 1.0, 1.0], [5.0, 1.0, 1.0]])
    res = nr.build_neighborlists_multi(cell, pos, [2.0, 5.0])
    assert set(res) == {0, 1}
    assert pairs(res[0]["edge_index"]) == [(0, 1)]
    assert res[1]["edge_index"].shape == (2, 3)
    for q, rc in enumerate([2.0, 5.0]):
        assert pairs(res[q]["edge_index"]) == pairs(nr.build_neighborlists(cell, pos, rc)["edge_index"])


# ---- tests/test_batch.py ----------------------------------------------------
def test_batch_basic(nr):
    pos = np.array([[0.0, 0, 0], [1.0, 0, 0], [10.0, 0, 0], [11.0, 0, 0], [12.0, 0, 0]])
    batch = np.array([0, 0, 1, 1, 1], dtype=np.int32)
    r = nr.build_neighborlists_batch(pos, batch, cutoff=2.5)
    assert pairs(r["edge_index"]) == [(0, 1), (2, 3), (2, 4), (3, 4)]


def test_batch_multi(nr):
    pos = np.array([[0.0, 0, 0], [1.0, 0, 0], [10.0, 0, 0], [13.0, 0, 0]])
    batch = np.array([0, 0, 1, 1], dtype=np.int32)
    res = nr.build_neighborlists_batch_multi(pos, batch, cutoffs=[2.0, 4.0])
    assert pairs(res[0]["edge_index"]) == [(0, 1)]
    assert pairs(res[1]["edge_index"]) == [(0, 1), (2, 3)]


def test_batch_multi_pbc_equals_singles(nr):
    pos1, rows1 = fcc(3.6, (2, 2, 2))
    pos2 = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    pos = np.concatenate([pos1, pos2])
    batch = np.array([0] * len(pos1) + [1] * len(pos2), dtype=np.int32)
    cells = np.stack([rows1, np.zeros((3, 3))])
    cut = [2.5, 6.0]
    res = nr.build_neighborlists_batch_multi(pos, batch, cells=cells, cutoffs=cut)
    for q, rc in enumerate(cut):
        a = nr.build_neighborlists(nr.PyCell(rows1.T.tolist()), pos1, rc)
        b = nr.build_neighborlists(None, pos2, rc)
        want = sorted([(int(i), int(j), *map(int, s)) for (i, j), s in zip(a["edge_index"].T, a["shift"])] +
                      [(int(i) + len(pos1), int(j) + len(pos1), *map(int, s))
                       for (i, j), s in zip(b["edge_index"].T, b["shift"])])
        got = sorted((int(i), int(j), *map(int, s))
                     for (i, j), s in zip(res[q]["edge_index"].T, res[q]["shift"]))
        assert got == want, f"cutoff {rc}"


# ---- tests/test_disjoint.py -------------------------------------------------
def test_disjoint_shells_partition(nr):
    pos, rows = fcc(3.6, (3, 3, 3))
    cell = nr.PyCell(rows.tolist())
    cut = [3.0, 4.0, 5.0]
    cum = nr.build_neighborlists_multi(cell, pos, cut, disjoint=False)
    dis = nr.build_neighborlists_multi(cell, pos, cut, disjoint=True)
    n = [cum[q]["edge_index"].shape[1] for q in range(3)]
    d = [dis[q]["edge_index"].shape[1] for q in range(3)]
    assert n[0] < n[1] < n[2]
    assert d == [n[0], n[1] - n[0], n[2] - n[1]]
    shells = [set(tuple(sorted(p)) for p in dis[q]["edge_index"].T.tolist()) for q in range(3)]
    assert shells[0].isdisjoint(shells[1]) and shells[1].isdisjoint(shells[2])
    assert shells[0].isdisjoint(shells[2])


def test_disjoint_unsorted_cutoffs(nr):
    pos, rows = fcc(3.6, (2, 2, 2))
    cell = nr.PyCell(rows.tolist())
    res = nr.build_neighborlists_multi(cell, pos, [5.0, 3.0], disjoint=True)
    small = nr.build_neighborlists(cell, pos, 3.0)["edge_index"].shape[1]
    total = nr.build_neighborlists(cell, pos, 5.0)["edge_index"].shape[1]
    assert res[1]["edge_index"].shape[1] == small
    assert res[0]["edge_index"].shape[1] == total - small


# ---- tests/test_issue_1.py --------------------------------------------------
def test_issue_1_single_atom_six_images(nr):
    cell = nr.PyCell([[5.0, 0, 0], [0, 5.0, 0], [0, 0, 5.0]], [True, True, True])
    r = nr.build_neighborlists(cell, np.array([[0.0, 0.0, 0.0]]), 6.0)
    assert r["edge_index"].shape[1] == 6
    assert not np.any(np.all(r["shift"] == 0, axis=1))


# ---- tests/test_bug_002_fix.py ----------------------------------------------
def test_bug_002_batched_pbc_images(nr):
    pos = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    cells = np.array([np.eye(3) * 5.0])
    rb = nr.build_neighborlists_batch(positions=pos, batch=np.array([0, 0], dtype=np.int32),
                                      cells=cells, cutoff=6.0, parallel=True)
    rs = nr.build_neighborlists(cell=nr.PyCell(cells[0].T.tolist()), positions=pos, cutoff=6.0)
    key = lambda r: sorted((int(i), int(j), *map(int, s))
                           for (i, j), s in zip(r["edge_index"].T, r["shift"]))
    assert key(rb) == key(rs)
    assert np.any(rb["shift"] != 0)  # periodic images are present


# ---- tests/test_memory_layout.py --------------------------------------------
def test_memory_layout_non_contiguous(nr):
    rng = np.random.default_rng(0)
    pos = rng.random((100, 3)) * 10.0
    wide = np.zeros((100, 4))
    wide[:, :3] = pos
    nc = wide[:, :3]
    assert not nc.flags["C_CONTIGUOUS"]
    a = nr.build_neighborlists(None, np.ascontiguousarray(pos), 1.5)
    b = nr.build_neighborlists(None, nc, 1.5)
    np.testing.assert_array_equal(a["edge_index"], b["edge_index"])
    np.testing.assert_array_equal(a["shift"], b["shift"])
    batch_wide = np.zeros(100, dtype=np.int32)
    r = nr.build_neighborlists_batch(nc[:50], batch_wide[::2], None, 1.5)
    assert r["edge_index"].shape[0] == 2


# ---- tests/test_ase.py ------------------------------------------------------
def test_ase_pbc_true(nr):
    atoms = Atoms([[0.1, 0.1, 0.1], [2.9, 0.1, 0.1]], cell=np.eye(3) * 3.0, pbc=True)
    r = nr.build_from_ase(atoms, 1.0)
    assert pairs(r["edge_index"]) == [(0, 1)]
    np.testing.assert_array_equal(r["shift"][0], [-1, 0, 0])
    m = nr.build_neighborlists(nr.PyCell(atoms.get_cell().T.tolist()), atoms.get_positions(), 1.0)
    np.testing.assert_array_equal(r["edge_index"], m["edge_index"])
    np.testing.assert_array_equal(r["shift"], m["shift"])


def test_ase_mixed_pbc(nr):
    atoms = Atoms([[0.1, 0.1, 0.1], [0.1, 0.1, 9.9]], cell=np.eye(3) * 10.0,
                  pbc=[True, True, False])
    assert nr.build_from_ase(atoms, 1.5)["edge_index"].shape[1] == 0
    atoms.positions[1] = [9.9, 0.1, 0.1]
    assert nr.build_from_ase(atoms, 1.5)["edge_index"].shape[1] == 1


def test_ase_no_pbc(nr):
    r = nr.build_from_ase(Atoms([[0, 0, 0], [0, 0, 1.0]], pbc=False), 1.5)
    assert pairs(r["edge_index"]) == [(0, 1)]


def test_ase_multi_cutoff(nr):
    pos, rows = fcc(3.6, (2, 2, 2))
    res = nr.build_multi_from_ase(Atoms(pos, cell=rows, pbc=True), [2.0, 3.0])
    assert res[0]["edge_index"].shape[1] == 0  # below the 2.55 A nearest neighbour
    assert res[1]["edge_index"].shape[1] > 0


# ---- tests/test_api_extensions.py -------------------------------------------
THREE = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.5, 0, 0]])


def test_ext_labels_single_and_batch(nr):
    res = nr.build_neighborlists_multi(None, THREE, [1.6, 3.0], labels=["short", "long"])
    assert res["short"]["edge_index"].shape[1] == 2 and res["long"]["edge_index"].shape[1] == 3
    rb = nr.build_neighborlists_batch_multi(THREE, np.zeros(3, dtype=np.int32), cutoffs=[1.6, 3.0],
                                            labels=["short", "long"])
    assert rb["short"]["edge_index"].shape[1] == 2 and rb["long"]["edge_index"].shape[1] == 3
    ra = nr.build_multi_from_ase(Atoms(THREE), [1.6, 3.0], labels=["s", "l"])
    assert ra["s"]["edge_index"].shape[1] == 2 and ra["l"]["edge_index"].shape[1] == 3


def test_ext_disjoint(nr):
    res = nr.build_neighborlists_multi(None, THREE, [1.6, 3.0], disjoint=True)
    assert res[0]["edge_index"].shape[1] == 2
    assert pairs(res[1]["edge_index"]) == [(0, 2)]


def test_ext_batches(nr):
    rng = np.random.default_rng(3)
    far = rng.random((10, 3)) * 100.0
    assert nr.build_neighborlists_batch(far, np.arange(10, dtype=np.int32),
                                        cutoff=1.0)["edge_index"].shape[1] == 0
    line = np.array([[i, 0, 0] for i in range(10)], dtype=np.float64)
    assert nr.build_neighborlists_batch(line, np.zeros(10, dtype=np.int32),
                                        cutoff=1.5)["edge_index"].shape[1] == 9
    pos = rng.random((200, 3))
    pos[1::2] = pos[0::2] + [0.5, 0, 0]
    assert nr.build_neighborlists_batch(pos, np.repeat(np.arange(100, dtype=np.int32), 2),
                                        cutoff=1.0)["edge_index"].shape[1] == 100


def test_ext_mixed_cells_in_batch(nr):
    two = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    r = nr.build_neighborlists_batch(np.concatenate([two, two]), np.array([0, 0, 1, 1], dtype=np.int32),
                                     cells=np.stack([np.eye(3) * 10.0, np.zeros((3, 3))]), cutoff=1.5)
    assert r["edge_index"].shape[1] == 2 and np.all(r["shift"] == 0)
    r1 = nr.build_neighborlists_batch(two, np.array([0, 0], dtype=np.int32),
                                      cells=np.array([np.eye(3) * 10.0]), cutoff=1.5)
    assert r1["edge_index"].shape[1] == 1


def test_ext_twenty_cutoffs(nr):
    res = nr.build_neighborlists_multi(None, np.array([[0.0, 0, 0], [1.0, 0, 0]]),
                                       [0.1 * i for i in range(1, 21)])
    assert len(res) == 20 and res[10]["edge_index"].shape[1] == 1


def test_ext_mixed_pbc_logic(nr):
    h = [[10, 0, 0], [0, 10, 0], [0, 0, 10]]
    pos = np.array([[5.0, 5.0, 0.5], [5.0, 5.0, 9.5]])
    assert nr.build_neighborlists(nr.PyCell(h, pbc=[True, True, False]), pos,
                                  2.0)["edge_index"].shape[1] == 0
    r = nr.build_neighborlists(nr.PyCell(h, pbc=[True, True, True]), pos, 2.0)
    assert r["edge_index"].shape[1] == 1 and list(r["shift"][0]) == [0, 0, -1]
