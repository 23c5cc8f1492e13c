"""Build libnlmi355.so, the gfx950 neighbour-list engine, in-tree.

    python neighborlist-rs_amd/build.py [-v]

Every translation unit is HIP built for gfx950 only:
  -ffp-contract=off   the reference's FP64 expressions are evaluated term by
                      term (Rust never fuses a*b+c), so near-cutoff decisions
                      match bit for bit
  -packed-fp32-ops    disabled: packed FP32 forms are avoided on this pool
The shared library has no PyTorch dependency; the Python package loads it
with ctypes.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
INC = os.path.join(os.path.dirname(HERE), "include")
BUILD = os.path.join(HERE, "_build")
OUT_DIR = os.path.join(HERE, "neighborlist_rs", "lib")
OUT = os.path.join(OUT_DIR, "libnlmi355.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["nl_kernels.hip", "nl_engine.cpp", "nl_capi.cpp", "nl_server.cpp", "nl_skin.cpp",
           "nl_noalloc.cpp"]
FLAGS = [
    "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I", SRC, "-I", INC,
]


def _compile(src, verbose):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(SRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if warn and verbose:
        print(warn, flush=True)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(OUT_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


CPP_TEST = os.path.join(BUILD, "test_api_cpp")


def build_cpp_test(src: str, out: str = CPP_TEST) -> str:
    """A test program over the C++ mirror of the reference's Rust API
    (include/neighborlist.hpp): host C++ only, linked to the library."""
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", INC, src,
           "-L", OUT_DIR, "-lnlmi355", "-Wl,-rpath,$ORIGIN/../neighborlist_rs/lib", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
