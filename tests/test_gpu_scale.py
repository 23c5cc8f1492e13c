"""Parity at scale: one periodic system of 2,000,000 atoms (liquid density,
box ~271 A, rc 4: ~27 M edges), bit-exact against the oracle.  Each edge is
packed into one int64 key (i, j, and the image shift, which stays in -2..2
here) so the two lists compare after a plain sort instead of a five-key
lexsort."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _keys(ei, ej, s):
    s = np.asarray(s, dtype=np.int64).reshape(-1, 3)
    assert np.all(np.abs(s) <= 2)
    code = ((s[:, 0] + 2) * 5 + (s[:, 1] + 2)) * 5 + (s[:, 2] + 2)  # 0..124
    k = (np.asarray(ei, np.int64) << 28) | (np.asarray(ej, np.int64) << 7) | code
    k.sort()
    return k


def test_two_million_atoms_vs_oracle(nr, oracle):
    n, rho, rc = 2_000_000, 0.1, 4.0
    L = (n / rho) ** (1.0 / 3.0)
    rng = np.random.default_rng(2024)
    pos = rng.uniform(0.0, L, (n, 3))
    H = np.diag([L, L, L])
    print("[case] scale", n, "atoms rc", rc, flush=True)
    got = nr.build_neighborlists(nr.PyCell(H.tolist(), [True] * 3), pos, rc)
    want = oracle.search_single(pos, H, [1, 1, 1], rc)
    g = got["edge_index"]
    assert g.shape[1] == want[0].size > 20_000_000
    np.testing.assert_array_equal(_keys(g[0], g[1], got["shift"]), _keys(*want))
