"""Builds the CPU oracle (test infrastructure) into oracle/build/liboracle.so.

gcc -O3 -ffp-contract=off -fno-fast-math: IEEE binary32 evaluated in source
order, the same arithmetic rules as the HIP library.  No -march=native for the
tests' build (built here, run on other hosts); the CPU baseline times a second,
-march=native build made on the host that runs it.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "build"
LIB = OUT / "liboracle.so"
# the timing build of the CPU baseline: the same source and FP rules, tuned for
# the host it is built on (built where it runs, never shipped)
NATIVE_LIB = OUT / "liboracle_native.so"

FLAGS = ["-O3", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-pthread", "-Wall",
         "-Wno-unused-function"]


def _compile(out: Path, extra: list[str]) -> Path:
    OUT.mkdir(exist_ok=True)
    tmp = out.with_suffix(f".so.tmp{os.getpid()}")
    cmd = ["gcc", *FLAGS, *extra, str(HERE / "oracle.c"), str(HERE / "oracle_fp64.c"), "-o", str(tmp), "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(tmp, out)
    return out


def build(verbose: bool = False) -> Path:
    _compile(LIB, [])
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_native() -> Path:
    """-march=native build for timing on this host (-ffp-contract=off keeps the
    results bit-identical to the portable build)."""
    return _compile(NATIVE_LIB, ["-march=native"])


if __name__ == "__main__":
    build(verbose=True)
