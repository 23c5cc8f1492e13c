"""Minimal (extended) XYZ reader for the reference's fixture files.

The fixtures in tests/golden/data are copied from neighborlist-rs
tests/data (MIT OR Apache-2.0, generated by its tests/generate_data.py).
ASE is not installed here, so this reader parses what those files hold:
an atom count, a comment line that may carry Lattice="ax ay az bx by bz cx cy cz"
(rows are lattice vectors, as ASE writes them) and pbc="T T T", then
`symbol x y z` lines.
"""
import gzip
import os
import re

import numpy as np

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data")


def read_xyz(name):
    path = os.path.join(DATA, name)
    if not os.path.exists(path) and os.path.exists(path + ".gz"):
        path = path + ".gz"
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt") as f:
        lines = f.read().splitlines()
    n = int(lines[0].split()[0])
    comment = lines[1]
    pos = np.array([[float(v) for v in ln.split()[1:4]] for ln in lines[2:2 + n]],
                   dtype=np.float64)
    cell_rows = None
    m = re.search(r'Lattice="([^"]+)"', comment)
    if m:
        cell_rows = np.array([float(v) for v in m.group(1).split()]).reshape(3, 3)
    pbc = [False, False, False]
    m = re.search(r'pbc="([^"]+)"', comment)
    if m:
        pbc = [t == "T" for t in m.group(1).split()]
    return pos, cell_rows, pbc
