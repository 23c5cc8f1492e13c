"""Host-side checks of the C ABI and the drop-in Python surface (no GPU).

The library must load and export every function include/*.h declares;
argument validation must raise the reference's errors before any device
work; with no GPU present a search must fail loudly (there is no CPU path).
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", h) for h in ("nl_api.h", "nl_engine_ext.h")]


def _declared():
    names = set()
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(nl_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol(nr):
    from neighborlist_rs import _lib
    declared = _declared()
    assert len(declared) >= 30
    for name in sorted(declared):
        assert hasattr(_lib.lib, name), name
    assert set(_lib.exported_symbols()) == declared


def test_version(nr):
    from neighborlist_rs._lib import lib
    assert b"gfx950" in lib.nl_version()


def test_no_gpu_fails_loudly(nr):
    from neighborlist_rs._lib import lib
    if lib.nl_device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(ValueError, match="no HIP device"):
        nr.build_neighborlists(None, np.random.rand(10, 3), 2.0)


# ---- validation (tests/test_validation.py, test_api_extensions.py) ---------
def test_position_shape_validation(nr):
    with pytest.raises(ValueError, match=r"Positions must be \(N, 3\)"):
        nr.build_neighborlists_batch(np.random.rand(10, 2), np.zeros(10, np.int32), None, 5.0)
    with pytest.raises(ValueError, match=r"Positions must be \(N, 3\)"):
        nr.build_neighborlists_batch_multi(np.random.rand(10, 2), np.zeros(10, np.int32), None,
                                           [5.0])
    with pytest.raises(ValueError, match=r"Positions must be \(N, 3\)"):
        nr.build_neighborlists(None, np.random.rand(10, 4), 5.0)


@pytest.mark.parametrize("bad", [-1.0, 0.0, float("nan"), float("inf")])
def test_cutoff_validation(nr, bad):
    with pytest.raises(ValueError, match="Cutoff must be positive and finite"):
        nr.build_neighborlists(None, np.random.rand(10, 3), bad)
    with pytest.raises(ValueError, match="Cutoff must be positive and finite"):
        nr.build_neighborlists_batch(np.random.rand(10, 3), np.zeros(10, np.int32), cutoff=bad)


def test_multi_cutoff_validation(nr):
    with pytest.raises(ValueError, match="All cutoffs must be positive and finite"):
        nr.build_neighborlists_multi(None, np.random.rand(10, 3), [1.0, -2.0, 3.0])


def test_unsorted_batch_validation(nr):
    b = np.array([0, 0, 1, 1, 0, 0, 1, 1, 2, 2], dtype=np.int32)
    with pytest.raises(ValueError, match=r"Batch IDs must be monotonic \(non-decreasing\)"):
        nr.build_neighborlists_batch(np.random.rand(10, 3), b, None, 5.0)


def test_batch_length_mismatch(nr):
    with pytest.raises(ValueError, match="positions and batch must have the same length"):
        nr.build_neighborlists_batch(np.random.rand(4, 3), np.zeros(3, np.int32), cutoff=1.0)


def test_batch_too_few_cells(nr):
    with pytest.raises(ValueError, match="Expected at least 2 cells, but got 1"):
        nr.build_neighborlists_batch(np.random.rand(4, 3), np.array([0, 0, 1, 1], np.int32),
                                     cells=np.zeros((1, 3, 3)), cutoff=1.0)


def test_labels_mismatch(nr):
    with pytest.raises(ValueError, match="Length of labels must match length of cutoffs"):
        nr.build_neighborlists_batch_multi(np.random.rand(10, 3), np.zeros(10, np.int32),
                                           cutoffs=[1.0, 2.0], labels=["one"])


def test_empty_inputs_need_no_device(nr):
    pos = np.zeros((0, 3))
    r = nr.build_neighborlists(None, pos, 1.0)
    assert r["edge_index"].shape == (2, 0) and r["shift"].shape == (0, 3)
    assert r["edge_index"].dtype == np.int64 and r["shift"].dtype == np.int32
    rm = nr.build_neighborlists_multi(None, pos, [1.0, 2.0])
    assert len(rm) == 2 and rm[0]["edge_index"].shape == (2, 0)
    rb = nr.build_neighborlists_batch(pos, np.zeros(0, np.int32), cutoff=1.0)
    assert rb["edge_index"].shape == (2, 0)
    assert nr.build_neighborlists_batch_multi(pos, np.zeros(0, np.int32), cutoffs=[1.0]) == {}
    assert nr.build_neighborlists_multi(None, np.random.rand(5, 3), []) == {}
    assert nr.build_neighborlists_batch_multi(np.random.rand(5, 3), np.zeros(5, np.int32),
                                              cutoffs=[]) == {}


# ---- PyCell (test_basic.py, test_api_extensions.py, test_ergonomics.py) ----
def test_pycell(nr):
    c = nr.PyCell([[10.0, 0, 0], [0, 10.0, 0], [0, 0, 10.0]])
    assert c.wrap([15.0, -2.0, 8.0]) == pytest.approx([5.0, 8.0, 8.0])
    with pytest.raises(ValueError):
        nr.PyCell([[0, 0, 0], [0, 0, 0], [0, 0, 0]])
    with pytest.raises(ValueError, match="Cell matrix must be 3x3"):
        nr.PyCell([[1, 0], [0, 1]])
    r = repr(nr.PyCell([[10, 0, 0], [0, 10, 0], [0, 0, 10]], pbc=[True, False, True]))
    assert "pbc=[true, false, true]" in r.lower()
    assert r.startswith("PyCell(h=[[10, 0, 0]")


class FakeAtoms:
    """Duck-typed stand-in for ase.Atoms (get_positions / get_cell / get_pbc)."""

    def __init__(self, positions, cell, pbc):
        self._p = np.asarray(positions, dtype=np.float64)
        self._c = np.asarray(cell, dtype=np.float64)
        self._b = np.asarray(pbc, dtype=bool)

    def get_positions(self):
        return self._p.copy()

    def get_cell(self):
        return self._c.copy()

    def get_pbc(self):
        return self._b.copy()


def test_pycell_from_ase_transposes_rows(nr):
    rows = np.array([[10.0, 0, 0], [2.0, 10.0, 0], [1.0, 0.5, 10.0]])
    c = nr.PyCell.from_ase(FakeAtoms(np.zeros((1, 3)), rows, [True, True, False]))
    np.testing.assert_array_equal(c.h, rows.T)
    assert c.pbc == [True, True, False]
    assert "PyCell" in repr(c)


def test_host_inverse_matches_oracle_bits(nr, oracle):
    from neighborlist_rs._lib import lib, dp
    rng = np.random.default_rng(1)
    for _ in range(50):
        H = np.ascontiguousarray(np.triu(rng.uniform(-3, 3, (3, 3))) + np.diag(
            rng.uniform(5, 50, 3)))
        out = np.zeros(9)
        assert lib.nl_cell_inverse(H.reshape(9).ctypes.data_as(dp), out.ctypes.data_as(dp)) == 0
        np.testing.assert_array_equal(out.reshape(3, 3), oracle.cell_inverse(H))
        for pbc in ([1, 1, 1], [1, 0, 1]):
            p = rng.uniform(-100, 100, 3)
            np.testing.assert_array_equal(nr.PyCell(H.tolist(), pbc).wrap(p),
                                          oracle.cell_wrap(H, pbc, p))


# ---- configuration (test_api_extensions.py:8-43) ---------------------------
def test_thresholds_and_threads(nr):
    for name, v in (("brute_force", 1234), ("parallel", 567), ("stack", 890)):
        old = getattr(nr, f"get_{name}_threshold")()
        getattr(nr, f"set_{name}_threshold")(v)
        assert getattr(nr, f"get_{name}_threshold")() == v
        getattr(nr, f"set_{name}_threshold")(old)
    old = nr.get_stack_threshold()
    nr.set_stack_threshold(5000)
    assert nr.get_stack_threshold() == 1024  # MAX_STACK_SIZE clamp (config.rs:36-38)
    nr.set_stack_threshold(old)
    t = nr.get_num_threads()
    assert t > 0
    nr.set_num_threads(t + 1)
    assert nr.get_num_threads() > 0
    nr.init_logging("debug")
    nr.init_logging("info")
    nr.init_logging(None)


def test_engine_setters_roundtrip(nr):
    from neighborlist_rs._lib import lib
    old = lib.nl_get_bin_order()
    for v, want in ((1, 1), (0, 0), (7, 1)):
        lib.nl_set_bin_order(v)
        assert lib.nl_get_bin_order() == want
    lib.nl_set_bin_order(old)
    old = lib.nl_get_fused_build()
    for v, want in ((0, 0), (1, 1), (2, 2), (5, 1), (-1, 1)):
        lib.nl_set_fused_build(v)
        assert lib.nl_get_fused_build() == want
    lib.nl_set_fused_build(old)


# ---- torch surface (neighborlist_rs.tensors): validation before device work --
def test_torch_surface_validation(nr):
    import torch
    from neighborlist_rs import tensors
    with pytest.raises(TypeError, match="float64"):
        nr.build_neighborlists_torch(None, torch.zeros((4, 3), dtype=torch.float32), 2.0)
    with pytest.raises(ValueError, match=r"Positions must be \(N, 3\)"):
        nr.build_neighborlists_torch(None, torch.zeros((4, 2), dtype=torch.float64), 2.0)
    with pytest.raises(ValueError, match="device"):
        nr.build_neighborlists_torch(None, torch.zeros((4, 3), dtype=torch.float64), 2.0)
    with pytest.raises(TypeError, match="torch.Tensor"):
        nr.build_neighborlists_torch(None, np.zeros((4, 3)), 2.0)
    # batch ids: maximal runs of equal ids, non-decreasing, same length
    np.testing.assert_array_equal(tensors._batch_sizes(torch.tensor([0, 0, 1, 1, 1, 4]), 6),
                                  [2, 3, 1])
    np.testing.assert_array_equal(tensors._batch_sizes(np.array([7, 7, 7]), 3), [3])
    with pytest.raises(ValueError, match="monotonic"):
        tensors._batch_sizes(torch.tensor([0, 1, 0]), 3)
    with pytest.raises(ValueError, match="same length"):
        tensors._batch_sizes(torch.tensor([0, 1]), 3)
    with pytest.raises(ValueError, match="Length of labels"):
        tensors._check_cutoffs([1.0, 2.0], ["a"])
    with pytest.raises(ValueError, match="All cutoffs must be positive and finite"):
        tensors._check_cutoffs([1.0, float("nan")], None)


def test_skin_lists_reject_disjoint(nr):
    """Skin lists are cumulative: a disjoint shell built at rc_q + skin_q
    would lose pairs that drift across a shell boundary between rebuilds, so
    both the C ABI and the torch manager refuse disjoint=True before any
    device work."""
    from neighborlist_rs import _lib
    lib = _lib.lib
    sz = np.array([4], np.int64)
    cut = np.array([2.0, 3.0])
    sk = np.array([0.5, 0.5])
    p = C.c_void_p()
    rc = lib.nl_skin_create(C.c_void_p(1), 4, sz.ctypes.data_as(_lib.i64p), 1, None, None, None,
                            cut.ctypes.data_as(_lib.dp), sk.ctypes.data_as(_lib.dp), 2, 1, 1, 0,
                            None, C.byref(p))
    assert rc != 0 and not p.value
    assert "cumulative" in _lib.last_error()
    from neighborlist_rs.md import DynamicNeighborManager
    with pytest.raises(ValueError, match="cumulative"):
        DynamicNeighborManager(np.zeros((4, 3)), [2.0, 3.0], 0.5, disjoint=True)
