"""Golden edge lists (tests/golden/vectors, written by tests/golden/make_golden.py).

CPU: the oracle still produces the committed vectors byte for byte, and the
vectors hold the pair counts the reference logged for the same fixture files
(benchmarks/benchmark_EPYC7R13.txt; self-image edges excluded as in
comprehensive_benchmark.py:96-100).  GPU: the engine reproduces them exactly
after the canonical sort.
"""
import glob
import os

import numpy as np
import pytest

from conftest import as_tuple, assert_same_edges

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "vectors", "*.npz")))
# reference-logged counts, self images excluded
LOGGED = {"si_bulk_rc6": 160, "si_bulk_rc14": 2064, "isolated_100_rc6": 2269,
          "ethanol_1000_rc6": 57840}


def _load(path):
    z = np.load(path)  # plain arrays only (allow_pickle stays False)
    H = None if z["cell"].size == 0 else z["cell"]
    pbc = None if z["pbc"].size == 0 else [bool(v) for v in z["pbc"]]
    return (os.path.basename(path)[:-4], z["positions"], H, pbc, float(z["cutoff"]),
            (z["edge_i"], z["edge_j"], z["shift"]))


def test_golden_files_present():
    names = {os.path.basename(p)[:-4] for p in FILES}
    assert set(LOGGED) <= names and "lcg_100_rc2.5" in names


@pytest.mark.parametrize("path", FILES, ids=lambda p: os.path.basename(p)[:-4])
def test_golden_counts_are_the_reference_logs(path):
    name, _, _, _, _, (ei, ej, s) = _load(path)
    if name in LOGGED:
        assert int((ei != ej).sum()) == LOGGED[name]
    assert np.all((ei < ej) | ((ei == ej) & np.any(s != 0, axis=1)))


@pytest.mark.parametrize("path", FILES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_golden(oracle, path):
    name, pos, H, pbc, rc, want = _load(path)
    got = oracle.search_single(pos, H, pbc, rc)
    assert_same_edges(got, want, name)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=lambda p: os.path.basename(p)[:-4])
def test_gpu_reproduces_golden(nr, path):
    name, pos, H, pbc, rc, want = _load(path)
    print("[case] golden", name, flush=True)
    cell = None if H is None else nr.PyCell(np.asarray(H).tolist(), pbc)
    got = as_tuple(nr.build_neighborlists(cell, pos, rc))
    assert_same_edges(got, want, name)
