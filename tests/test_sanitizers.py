"""The host half of the library (scene container, generators, BVH builder and
packers, two-level build, file formats, the host C-ABI entry points) and the
CPU oracle, built with AddressSanitizer and UndefinedBehaviorSanitizer and
driven by tests/cpp/sanitize_driver.cpp: out-of-bounds accesses, use after
free, leaks and undefined behaviour abort the driver.  Host code only (g++ and
gcc; no HIP source is part of this build)."""
import concurrent.futures as cf
import os
import subprocess

import pytest

from conftest import ROOT

CSRC = ROOT / "strelka_amd" / "csrc"
HOST = ["bvh.cpp", "bvh_api.cpp", "files.cpp", "scene.cpp", "scenegen.cpp"]
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]
FP = ["-ffp-contract=off", "-fno-fast-math"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    out = tmp_path_factory.mktemp("san")
    jobs = [["g++", "-std=c++17", *SAN, *FP, "-pthread", "-c", str(CSRC / f), "-o", str(out / (f + ".o"))]
            for f in HOST]
    jobs.append(["gcc", "-std=c11", *SAN, *FP, "-pthread", "-c", str(ROOT / "oracle" / "oracle.c"), "-o",
                 str(out / "oracle.o")])
    jobs.append(["g++", "-std=c++17", *SAN, *FP, "-pthread", "-c", str(ROOT / "tests" / "cpp" / "sanitize_driver.cpp"),
                 "-o", str(out / "driver.o")])
    with cf.ThreadPoolExecutor(max_workers=min(len(jobs), 8)) as ex:
        list(ex.map(run, jobs))
    exe = out / "sanitize_driver"
    run(["g++", *SAN, "-pthread", *[str(p) for p in sorted(out.glob("*.o"))], "-o", str(exe), "-lm"])
    return exe, out


def test_host_code_and_oracle_are_clean_under_asan_and_ubsan(driver):
    exe, out = driver
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", STK_SAN_TMP=str(out))
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert "all checks passed" in r.stdout
