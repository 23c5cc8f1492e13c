"""One HIP runtime per process whichever of neighborlist_rs and torch is
imported first (a GNN script that writes `import neighborlist_rs` above
`import torch`, as isort orders them).  Each case runs in a fresh child
process, since the import order of this test session is already fixed."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "neighborlist-rs_amd")

try:
    import importlib.util
    HAVE_TORCH = importlib.util.find_spec("torch") is not None
except (ImportError, ValueError):
    HAVE_TORCH = False


def run_child(code, timeout=300):
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join([PKG, os.path.join(ROOT, "tests"), ROOT,
                                         env.get("PYTHONPATH", "")])
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=timeout, env=env)
    assert r.returncode == 0, f"child failed ({r.returncode}):\n{r.stdout}\n{r.stderr}"
    return r.stdout


MAPPED = """
import os
libs = sorted({os.path.realpath(l.split()[-1]) for l in open('/proc/self/maps')
               if 'libamdhip64' in l})
print(len(libs), *libs)
"""


@pytest.mark.skipif(not HAVE_TORCH, reason="torch is not installed")
def test_library_first_maps_one_runtime():
    out = run_child("import neighborlist_rs\nimport torch\n" + MAPPED)
    n = int(out.split()[0])
    assert n == 1, out
    assert "torch" in out, out  # the runtime torch itself uses


@pytest.mark.skipif(not HAVE_TORCH, reason="torch is not installed")
def test_torch_first_maps_one_runtime():
    out = run_child("import torch\nimport neighborlist_rs\n" + MAPPED)
    assert int(out.split()[0]) == 1, out


@pytest.mark.gpu
@pytest.mark.skipif(not HAVE_TORCH, reason="torch is not installed")
def test_library_first_then_torch_device_and_search():
    code = """
import neighborlist_rs as nr
import torch
import numpy as np
from xyzio import read_xyz
x = torch.zeros(4, device="cuda")
assert float(x.sum()) == 0.0
pos, rows, pbc = read_xyz("ethanol_1000.xyz")
r = nr.build_neighborlists(nr.PyCell(rows.T.tolist(), pbc), pos, 6.0)
assert r["edge_index"].shape[1] == 57840, r["edge_index"].shape
from neighborlist_rs.tensors import build_neighborlists_torch
t = build_neighborlists_torch(nr.PyCell(rows.T.tolist(), pbc),
                                torch.from_numpy(pos).cuda(), 6.0)
assert t["edge_index"].shape[1] == 57840
assert t["edge_index"].device.type == "cuda"
print("ok")
"""
    assert run_child(code).strip().endswith("ok")
