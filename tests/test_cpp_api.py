"""The C++ mirror of the reference's Rust API (include/neighborlist.hpp).

tests/cpp/test_api.cpp restates the reference's Rust integration tests
(tests/test_single_system.rs, tests/test_batch_system.rs,
tests/test_api_structure.rs) and its error texts against the mirror.

* CPU: the header and the program compile and link (-Wall -Wextra -Werror);
  run without a GPU, the program fails loudly with the library's message
  (there is no CPU fallback), and on a machine with a GPU it passes.
* GPU: the program built by build() passes every test.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_api.cpp")
BIN = os.path.join(ROOT, "neighborlist-rs_amd", "_build", "test_api_cpp")
NAMES = ["test_build_neighborlists_simple", "test_build_neighborlists_pbc",
         "test_build_neighborlists_batch_mixed", "test_neighborlist_struct_exists",
         "test_batch_without_cells", "test_errors"]


def _builder():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "nl_build", os.path.join(ROOT, "neighborlist-rs_amd", "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_cpp_mirror_builds_and_needs_the_gpu(tmp_path):
    b = _builder()
    if not os.path.exists(b.OUT):
        pytest.fail("libnlmi355.so is not built (run __graft_entry__.build())")
    exe = b.build_cpp_test(SRC, str(tmp_path / "test_api_cpp"))
    # the rpath is relative to the build folder: point the loader at the library
    env = dict(os.environ, LD_LIBRARY_PATH=b.OUT_DIR)
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    if r.returncode != 0:
        assert "no HIP device available" in r.stderr, r.stderr
    else:
        assert r.stdout.split() == [w for n in NAMES for w in ("ok", n)], r.stdout


@pytest.mark.gpu
def test_cpp_mirror_reference_tests():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing (run __graft_entry__.build())")
    print("[case] C++ API mirror", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split() == [w for n in NAMES for w in ("ok", n)], r.stdout
