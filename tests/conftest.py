import os
import sys

import numpy as np
import pytest

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


def _packed_keys(ei, ej, s, lo, span, n):
    """One int64 per edge, ordered as (i, j, Sx, Sy, Sz); None if it would
    not fit 63 bits."""
    bits_n = max(1, int(n).bit_length())
    bits_s = max(1, int(span).bit_length())
    if 2 * bits_n + 3 * bits_s > 62:
        return None
    k = ei.astype(np.int64)
    k = (k << bits_n) | ej.astype(np.int64)
    for c in range(3):
        k = (k << bits_s) | (s[:, c].astype(np.int64) - lo)
    return np.sort(k)


def assert_same_edge_sets(got, want, what=""):
    """Bit-exact (i, j, S) multisets, for lists of millions of edges: each edge
    packed into one int64 key and the keys sorted (falls back to the
    canonical lexsort when the indices and shifts do not fit 63 bits)."""
    gi, gj, gs = (np.asarray(a) for a in got)
    wi, wj, ws = (np.asarray(a) for a in want)
    gi, gj, wi, wj = (a.reshape(-1) for a in (gi, gj, wi, wj))
    gs, ws = gs.reshape(-1, 3), ws.reshape(-1, 3)
    assert gi.size == wi.size, f"{what}: {gi.size} edges vs {wi.size} expected"
    if gi.size == 0:
        return
    lo = int(min(gs.min(), ws.min()))
    span = int(max(gs.max(), ws.max())) - lo
    n = int(max(gi.max(), gj.max(), wi.max(), wj.max())) + 1
    kg = _packed_keys(gi, gj, gs, lo, span, n)
    if kg is None or min(gi.min(), gj.min(), wi.min(), wj.min()) < 0:
        assert_same_edges(got, want, what)
        return
    kw = _packed_keys(wi, wj, ws, lo, span, n)
    assert np.array_equal(kg, kw), f"{what}: edge sets differ"


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
