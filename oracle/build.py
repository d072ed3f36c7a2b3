"""Builds the CPU oracle (test infrastructure) into oracle/build/liboracle.so.

gcc -O3 -ffp-contract=off -fno-fast-math: IEEE binary32 evaluated in source
order, the same arithmetic rules as the HIP library.  No -march=native: the
library is built here and run on other hosts.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "build"
LIB = OUT / "liboracle.so"


def build(verbose: bool = False) -> Path:
    OUT.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-pthread",
           "-Wall", "-Wno-unused-function", str(HERE / "oracle.c"), "-o", str(tmp), "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
