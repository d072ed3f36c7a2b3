"""The flat tree built on the device (stk_build_bvh_gpu_method: a linear BVH
or a tree from locally-ordered clustering): its skip-link array is a valid
reference layout (the host packer accepts it and the reference walk over it
equals brute force), and every GPU walk over it - the wide and binary
closest-hit walks, the any-hit walk, the path tracer - returns exactly what
the oracle's reference walks return over the same nodes."""
import time

import numpy as np
import pytest

import strelka_amd as sa
from oracle import Oracle
from oracle.pyoracle import MODE_ANY, MODE_BRUTE, MODE_REF

from test_gpu_pathtrace import MAX_DIVERGENT, compare_samples

pytestmark = pytest.mark.gpu


def same(a, b):
    return a.view(np.uint8).tobytes() == b.view(np.uint8).tobytes()


@pytest.fixture(scope="module")
def ctx():
    c = sa.Context(0)
    yield c
    c.close()


def sah_cost(nodes):
    """Sum of inner-node box areas over the root's (the walk cost the SAH
    estimates, up to the leaf term every one-triangle-per-leaf tree shares)."""
    inner = nodes["instId"] == -1
    lo, hi = nodes["minBounds"][inner].astype(np.float64), nodes["maxBounds"][inner].astype(np.float64)
    d = hi - lo
    area = d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0]
    return float(area.sum() / area[0])


@pytest.mark.parametrize("method", ["lbvh", "ploc"])
@pytest.mark.parametrize("config", [1, 6, 7])
def test_gpu_tree_is_a_valid_reference_layout(ctx, config, method):
    print(f"case: config {config}, device build ({method}), closest and any hits", flush=True)
    s = sa.SceneData.generate(config)
    ctx.set_scene(s)
    ctx.build_bvh_gpu(method)
    nodes = ctx.get_bvh()
    assert nodes.size == 2 * int(s.inst["indexCount"].sum() // 3) - 1
    sa.HostBvh.from_nodes(s, nodes)  # the host packer validates preorder, skip links and leaves
    o = Oracle.from_scene(s, nodes)
    rays = s.jittered_rays(0x5354524B00000001)
    ref, st = o.trace(rays, MODE_REF)
    assert st.hits > 0
    if config != 6:
        assert same(ref, o.trace(rays, MODE_BRUTE)[0])  # the boxes hold their triangles
    assert same(ctx.trace_closest(rays), ref)
    assert same(ctx.trace_closest(rays, reference_order=True), ref)
    shadow = rays.copy()
    shadow["o"][:, 3] = 1.5
    a_ref, _ = o.trace(shadow, MODE_ANY)
    a_gpu = ctx.trace_any(shadow)
    assert np.array_equal(a_gpu[["instId", "primId", "t"]], a_ref[["instId", "primId", "t"]])


@pytest.mark.parametrize("method", ["lbvh", "ploc"])
def test_gpu_tree_config2_and_render(ctx, method):
    print(f"case: 1.16M-triangle scene, device build ({method}), 1080p rays subsample, 16x16x4 render crop",
          flush=True)
    s = sa.SceneData.generate(2)
    ctx.set_scene(s)
    t0 = time.perf_counter()
    ctx.build_bvh_gpu(method)
    print(f"device build {time.perf_counter() - t0:.3f} s (first call includes kernel loading)")
    t0 = time.perf_counter()
    ctx.build_bvh_gpu(method)
    print(f"device build {time.perf_counter() - t0:.3f} s")
    nodes = ctx.get_bvh()
    info = ctx.bvh_info()
    print(f"inner-area cost {sah_cost(nodes):.1f}, wide levels {info.wideDepth}, walk stack {info.walkStack}")
    assert info.gpuBuilt == 1 and info.walkStack == (11 if info.wideDepth <= 12 else 16)
    if method == "lbvh":
        # the linear tree is deeper: it is the case that runs the 16-entry kernel
        assert info.walkStack == 16
    o = Oracle.from_scene(s, nodes)
    rays = s.jittered_rays(0x5354524B00000002)[::37]
    assert same(ctx.trace_closest(rays), o.trace(rays, MODE_REF)[0])
    p = sa.render_params(s.camera, spp=4, max_depth=8, frame=0, sampling=1, sky=s.sky, keep_samples=True)
    w, h = p.pt.dimension[0], p.pt.dimension[1]
    ctx.reset_accumulation()
    ctx.render(p)
    gpu = ctx.samples(w * h * 4 * 3).reshape(h, w, 4, 3)
    x0, y0 = 952, 532
    cpu = o.pt_frame(p, x0, y0, 16, 16)
    frac, _ = compare_samples(gpu[y0:y0 + 16, x0:x0 + 16].reshape(-1), cpu)
    print(f"divergent samples {frac:.5f}")
    assert frac <= MAX_DIVERGENT


def test_clustered_tree_is_cheaper_than_the_linear_one(ctx):
    print("case: config 2, inner-area cost of the clustered and linear device trees", flush=True)
    s = sa.SceneData.generate(2)
    ctx.set_scene(s)
    costs = {}
    for method in ("lbvh", "ploc"):
        ctx.build_bvh_gpu(method)
        costs[method] = sah_cost(ctx.get_bvh())
    ctx.build_bvh()
    costs["host sah"] = sah_cost(ctx.get_bvh())
    print(costs)
    assert costs["ploc"] < costs["lbvh"]
