"""Build recipe for the CPU oracle (test infrastructure, not product code).

Compiles oracle/nl_oracle.c with gcc into oracle/lib/libnl_oracle.so.
`-ffp-contract=off` keeps every a*b+c as a separate multiply and add, as the
Rust reference does; OpenMP gives the parallel count/fill passes that stand
in for the reference's rayon pool when the oracle is timed as the CPU baseline.

The reference itself (a Rust crate) cannot be built here: no Rust toolchain
is installed, so there is no oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "nl_oracle.c")
OUT_DIR = os.path.join(HERE, "lib")
OUT = os.path.join(OUT_DIR, "libnl_oracle.so")

CFLAGS = ["-O2", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
          "-shared", "-fPIC", "-Wall", "-Wno-unused-function"]


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = ["gcc", *CFLAGS, SRC, "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
