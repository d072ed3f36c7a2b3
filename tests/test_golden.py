"""The committed golden vectors (tests/golden/, written by make_golden.py)
against the oracle on the CPU (a change of the oracle, the scene generator or
the BVH builder shows up here) and against the GPU kernels through the C ABI
(no oracle needed on the GPU side)."""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

import strelka_amd as sa

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def gold():
    g = np.load(GOLD / "golden.npz")
    meta = json.loads((GOLD / "golden.json").read_text())
    return {k: g[k] for k in g.files}, meta


def fixture_rays(g):
    return g["rays"].view(sa.RAY).copy(), g["hits"].view(sa.HIT).copy()


def test_rays_regenerate(gold):
    g, meta = gold
    rays, _ = fixture_rays(g)
    s = sa.SceneData.generate(1)
    again = s.jittered_rays(int(meta["rays"]["seed"], 16))[::meta["rays"]["stride"]]
    assert again.tobytes() == rays.tobytes()


def test_oracle_reproduces_hits_rng_tonemap(gold):
    from oracle import Oracle, load
    from oracle.pyoracle import MODE_REF, MODE_WIDE, fptr

    g, meta = gold
    rays, hits = fixture_rays(g)
    s = sa.SceneData.generate(1)
    b = sa.HostBvh.build(s)
    assert b.nodes.size == meta["rays"]["bvh_nodes"]
    o = Oracle.from_scene(s, b.nodes, b.wide, b.wide_tris)
    for mode in (MODE_REF, MODE_WIDE):
        got, _ = o.trace(rays, mode)
        assert got.tobytes() == hits.tobytes()
    lib = load()
    for case, stream in zip(g["rng_cases"], g["rng"]):
        st = C.c_uint32(lib.orc_init_rng(*[int(x) for x in case]))
        assert st.value == stream[0]
        assert [lib.orc_xorshift(C.byref(st)) for _ in range(16)] == list(stream[1:])
    for t in range(3):
        src = np.ascontiguousarray(g["tonemap_in"])
        out = np.zeros_like(src)
        lib.orc_tonemap(fptr(src), src.shape[0], t, fptr(out))
        # the reference's Reinhard divides by the luminance: black maps to NaN
        # (shaders/tonemap.hlsl:58-63), frozen as such
        assert np.array_equal(out, g["tonemap"][t], equal_nan=True)
    assert np.all(np.isnan(g["tonemap"][0][60, :3]))


def test_oracle_reproduces_path_samples(gold):
    from oracle import Oracle

    g, meta = gold
    m = meta["pt"]
    s = sa.SceneData.generate(m["config"])
    b = sa.HostBvh.build(s)
    p = sa.render_params(s.camera, spp=m["spp"], max_depth=m["max_depth"], frame=m["frame"], sampling=m["sampling"],
                         sky=m["sky"])
    x, y, w, h = m["crop"]
    assert np.array_equal(Oracle.from_scene(s, b.nodes).pt_frame(p, x, y, w, h), g["pt"])


@pytest.mark.gpu
def test_gpu_hits_equal_golden(gold):
    print("case: golden config-1 rays through stk_trace_closest (wide, binary, reference order)", flush=True)
    g, _ = gold
    rays, hits = fixture_rays(g)
    s = sa.SceneData.generate(1)
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    for kw in ({}, {"binary": True}, {"reference_order": True}):
        assert ctx.trace_closest(rays, **kw).tobytes() == hits.tobytes(), kw
    ctx.close()


@pytest.mark.gpu
def test_gpu_path_samples_match_golden(gold):
    print("case: golden 16x16 crop of the 1080p frame, 4 spp, per-sample radiance", flush=True)
    g, meta = gold
    m = meta["pt"]
    s = sa.SceneData.generate(m["config"])
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    p = sa.render_params(s.camera, spp=m["spp"], max_depth=m["max_depth"], frame=m["frame"], sampling=m["sampling"],
                         sky=m["sky"], keep_samples=True)
    W, H = s.camera.dimension[0], s.camera.dimension[1]
    ctx.render(p)
    smp = ctx.samples(W * H * m["spp"] * 3).reshape(H, W, m["spp"], 3)
    ctx.close()
    x, y, w, h = m["crop"]
    got = smp[y:y + h, x:x + w].reshape(-1, 3)
    ref = g["pt"].reshape(-1, 3)
    # every sample within 1e-4 per channel (relative above 1): kernels and
    # oracle share their transcendentals, so no path takes another branch
    err = np.abs(got - ref) / np.maximum(1.0, np.abs(ref))
    print(f"max error {float(err.max()):.3e}")
    assert not np.any(err > 1e-4)
