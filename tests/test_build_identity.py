"""The loaded libstrelka_amd.so was built from the sources in this tree: the
SHA-256 over strelka_amd/csrc, include/stk and the compile flags that
strelka_amd/build.py computes now equals the one the library carries
(stk_build_info).  Run on the CPU and again in the GPU suite, so both runs
record which binary they loaded."""
import importlib.util

import pytest

import strelka_amd as sa
from conftest import ROOT


def tree_hash() -> str:
    spec = importlib.util.spec_from_file_location("_stk_build_for_hash", ROOT / "strelka_amd" / "build.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.source_hash()


def check():
    info = sa.lib().stk_build_info().decode()
    print(f"loaded {sa._abi.LIB_PATH}: {info}")
    assert info.startswith("source-sha256 ") and "offload-arch gfx950" in info
    assert tree_hash() in info, "libstrelka_amd.so is stale: rebuild it (__graft_entry__.build())"


def test_library_was_built_from_these_sources():
    check()


@pytest.mark.gpu
def test_library_on_the_gpu_box_was_built_from_these_sources():
    check()
