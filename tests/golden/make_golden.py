"""Writes the golden edge lists under tests/golden/vectors/ with the CPU oracle.

    python tests/golden/make_golden.py

Each file holds the inputs and the canonical (i, j, S)-sorted edge list of one
case, as .npz (plain arrays, no pickles).  The oracle itself is pinned by the
reference's known answers (tests/test_oracle_reference.py: the pair counts the
reference logged for these same fixture files, and its unit-test answers);
the vectors freeze those outputs so that the GPU path and any later change of
the oracle are compared against the same bytes.  Cases:

  si_bulk rc 6 and rc 14   periodic diamond Si (self images, many images)
  isolated_100 rc 6        isolated cluster (auto-box, brute force)
  lcg_100 rc 2.5           the reference's LCG positions (search.rs:1413-1432)
  ethanol_1000 rc 6        periodic liquid, cell-list path (57,840 edges)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402
from xyzio import read_xyz  # noqa: E402


def lcg_positions(n=100, seed=42):
    """search.rs:1413-1432: x <- x*6364136223846793005 + 1, value x / u64::MAX * 10."""
    x = seed
    out = []
    for _ in range(3 * n):
        x = (x * 6364136223846793005 + 1) % (1 << 64)
        out.append(float(x) / float((1 << 64) - 1) * 10.0)
    return np.array(out).reshape(n, 3)


def canon(ei, ej, s):
    order = np.lexsort((s[:, 2], s[:, 1], s[:, 0], ej, ei))
    return ei[order], ej[order], s[order]


def cases():
    pos, rows, pbc = read_xyz("si_bulk.xyz")
    yield "si_bulk_rc6", pos, rows.T, np.array(pbc, np.uint8), 6.0
    yield "si_bulk_rc14", pos, rows.T, np.array(pbc, np.uint8), 14.0
    pos, _, _ = read_xyz("isolated_100.xyz")
    yield "isolated_100_rc6", pos, None, None, 6.0
    yield "lcg_100_rc2.5", lcg_positions(), None, None, 2.5
    pos, rows, pbc = read_xyz("ethanol_1000.xyz")
    yield "ethanol_1000_rc6", pos, rows.T, np.array(pbc, np.uint8), 6.0


def main():
    out = os.path.join(HERE, "vectors")
    os.makedirs(out, exist_ok=True)
    for name, pos, H, pbc, rc in cases():
        ei, ej, s = canon(*O.search_single(pos, H, None if H is None else pbc.tolist(), rc))
        np.savez_compressed(os.path.join(out, name + ".npz"), positions=pos,
                            cell=np.zeros((0, 3)) if H is None else H,
                            pbc=np.zeros(0, np.uint8) if pbc is None else pbc,
                            cutoff=np.array(rc), edge_i=ei, edge_j=ej, shift=s.astype(np.int32))
        print(f"{name}: {len(ei)} edges")


if __name__ == "__main__":
    main()
