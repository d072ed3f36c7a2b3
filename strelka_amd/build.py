"""Builds libstrelka_amd.so in-tree for gfx950.

Host C++ (BVH builder, scene container, workload generators) is compiled with
g++, the HIP sources with hipcc; both with -ffp-contract=off so that host and
device arithmetic agree bit for bit.  Objects go to strelka_amd/_build/.
"""
from __future__ import annotations

import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libstrelka_amd.so"
ARCH = os.environ.get("STK_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["bvh.cpp", "bvh_api.cpp", "files.cpp", "scene.cpp", "scenegen.cpp", "buildinfo.cpp"]
HIP_SOURCES = ["kernels.hip", "pathtrace.hip", "context.hip", "cache.hip", "gpubuild.hip"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fvisibility=hidden",
          "-Wall", "-Wno-unused-function"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fhip-fp32-correctly-rounded-divide-sqrt", "-Wno-unused-command-line-argument"]


def source_hash() -> str:
    """SHA-256 over every source the library is built from (strelka_amd/csrc
    and include/stk, each file's path relative to the repository and its
    bytes, in path order) and the compile flags.  The library carries it
    (stk_build_info), so a test can tell whether the loaded binary was built
    from the sources in the tree."""
    root = PKG.parent
    files = sorted(p for d in (CSRC, root / "include" / "stk") for p in d.rglob("*")
                   if p.is_file() and p.suffix in (".hip", ".cpp", ".hpp", ".h"))
    h = hashlib.sha256()
    for f in files:
        h.update(f.relative_to(root).as_posix().encode() + b"\0")
        h.update(f.read_bytes())
        h.update(b"\0")
    h.update(" ".join(COMMON + HIP_FLAGS + [ARCH]).encode())
    return h.hexdigest()


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libstrelka_amd.so")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    ident = [f'-DSTK_SOURCE_HASH="{source_hash()}"', f'-DSTK_OFFLOAD_ARCH="{ARCH}"']
    for src in HOST_SOURCES:
        extra = ident if src == "buildinfo.cpp" else []
        jobs.append(["g++", *COMMON, *extra, "-pthread", "-c", str(CSRC / src), "-o", str(OBJ / (src + ".o"))])
    for src in HIP_SOURCES:
        if not (CSRC / src).exists():
            continue
        jobs.append([hipcc, *COMMON, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(OBJ / (src + ".o"))])
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for fut in [ex.submit(_run, j) for j in jobs]:
            fut.result()
    objs = [str(OBJ / (s + ".o")) for s in HOST_SOURCES + HIP_SOURCES if (CSRC / s).exists()]
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs, "-pthread"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
