import os
import sys

import numpy as np
import pytest

try:
    # loaded before the engine library, so that the engine binds to torch's
    # bundled HIP runtime and the process holds one runtime (torch's device
    # tensors and the engine's lists then live side by side)
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "neighborlist-rs_amd")
for p in (ROOT, PKG, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


def canon(ei, ej, s):
    """Canonical order of an edge list: sort on (i, j, Sx, Sy, Sz)."""
    ei = np.asarray(ei, dtype=np.int64).reshape(-1)
    ej = np.asarray(ej, dtype=np.int64).reshape(-1)
    s = np.asarray(s, dtype=np.int32).reshape(-1, 3)
    order = np.lexsort((s[:, 2], s[:, 1], s[:, 0], ej, ei))
    return ei[order], ej[order], s[order]


def assert_same_edges(got, want, what=""):
    """got/want: (ei, ej, shifts); bit-exact after the canonical sort."""
    gi, gj, gs = canon(*got)
    wi, wj, ws = canon(*want)
    assert gi.size == wi.size, f"{what}: {gi.size} edges vs {wi.size} expected"
    np.testing.assert_array_equal(gi, wi, err_msg=what)
    np.testing.assert_array_equal(gj, wj, err_msg=what)
    np.testing.assert_array_equal(gs, ws, err_msg=what)


def as_tuple(d):
    """{"edge_index", "shift"} dict -> (ei, ej, shifts)."""
    return d["edge_index"][0], d["edge_index"][1], d["shift"]


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    return O


@pytest.fixture(scope="session")
def nr():
    """The GPU engine's Python surface (needs the built library)."""
    import neighborlist_rs
    return neighborlist_rs
