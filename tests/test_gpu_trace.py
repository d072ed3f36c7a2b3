"""GPU traversal parity through the C ABI: hit IDs, t and barycentrics must be
bit-identical to the CPU oracle (reference order, shaders/raytracing.h:198-286)
on the same BVH and rays."""
import ctypes

import numpy as np
import pytest

import strelka_amd as sa
from oracle import Oracle
from oracle.pyoracle import MODE_ANY, MODE_PRUNED, MODE_REF

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF


def same(a, b):
    return a.view(np.uint8).tobytes() == b.view(np.uint8).tobytes()


def first_diff(a, b):
    bad = np.flatnonzero(a.view(np.uint8).reshape(a.size, -1).any(axis=1) ^ False)
    return bad


@pytest.fixture(scope="module")
def ctx():
    c = sa.Context(0)
    yield c
    c.close()


def setup(ctx, config):
    s = sa.SceneData.generate(config)
    ctx.set_scene(s)
    ctx.build_bvh()
    nodes = ctx.get_bvh()
    return s, Oracle.from_scene(s, nodes)


def test_config1_all_rays_bit_exact(ctx):
    print("case: config 1, 65,536 jittered rays", flush=True)
    s, o = setup(ctx, 1)
    rays = s.jittered_rays(0x5354524B00000001)
    ref, _ = o.trace(rays, MODE_REF)
    gpu = ctx.trace_closest(rays)
    gpu_ref = ctx.trace_closest(rays, reference_order=True)
    mism = np.flatnonzero(gpu.view(np.uint8).reshape(-1, 20) != ref.view(np.uint8).reshape(-1, 20))
    assert same(gpu, ref), f"{np.unique(mism // 20).size} rays differ"
    assert same(gpu_ref, ref)
    anyh_o, _ = o.trace(rays, MODE_ANY)
    for order in (False, True):
        anyh = ctx.trace_any(rays, reference_order=order)
        assert np.array_equal(anyh["instId"], anyh_o["instId"]) and np.array_equal(anyh["primId"], anyh_o["primId"])
        assert np.array_equal(anyh["t"], anyh_o["t"])
    # shadow-ray-like segments: shorten to 70% of the closest hit (all miss)
    short = rays.copy()
    short["o"][:, 3] = ref["t"] * np.float32(0.7)
    anys, _ = o.trace(short, MODE_ANY)
    assert same(ctx.trace_any(short), anys)


def test_config2_subsample_bit_exact_and_full_frame_orders_agree(ctx):
    print("case: config 2, 1080p jittered rays", flush=True)
    s, o = setup(ctx, 2)
    info = ctx.bvh_info()
    print(f"SAH tree: depth {info.depth}, wide levels {info.wideDepth}, walk stack {info.walkStack}")
    # this tree is walked by the 11-entry kernel (at most 12 wide levels)
    assert info.wideDepth <= 12 and info.walkStack == 11
    rays = s.jittered_rays(0x5354524B00000002)
    sub = rays[::37].copy()
    ref, _ = o.trace(sub, MODE_PRUNED)
    gpu = ctx.trace_closest(sub)
    assert same(gpu, ref)
    gpu_ref = ctx.trace_closest(sub[:8192], reference_order=True)
    assert same(gpu_ref, ref[:8192])
    # full 2,073,600 rays: the pruned near-first kernel and the skip-link
    # kernel must agree bit for bit (order independence of the result rule)
    full = ctx.trace_closest(rays)
    full_ref = ctx.trace_closest(rays, reference_order=True)
    assert same(full, full_ref)
    # any-hit: packed left-first walk == skip-link walk on every ray, and both
    # equal the oracle on the subsample
    seg = rays.copy()
    seg["o"][:, 3] = full["t"] * np.float32(1.001)
    a_pk = ctx.trace_any(seg)
    a_rf = ctx.trace_any(seg, reference_order=True)
    assert same(a_pk, a_rf)
    a_or, _ = o.trace(seg[::37].copy(), MODE_ANY)
    assert same(a_pk[::37].copy(), a_or)
    # Moller-Trumbore is not watertight: a handful of rays slip through the
    # shared edges of heightfield triangles (the CPU oracle agrees on them)
    assert np.mean(full["instId"] != INVALID) > 0.9999


def trace_image(ctx, rays, width, height, binary=False):
    n = width * height
    d_rays = ctx.alloc(max(1, n) * 32)
    d_hits = ctx.alloc(max(1, n) * 20)
    try:
        if n:
            ctx.upload(d_rays, rays[:n].copy())
        ctx.trace_closest_image_device(d_rays, width, height, d_hits, binary=binary)
        ctx.sync()
        return ctx.download(d_hits, np.zeros(n, dtype=sa.HIT))
    finally:
        ctx.free(d_rays)
        ctx.free(d_hits)


def test_image_rays_traced_in_tiles_give_the_same_hits(ctx):
    """stk_trace_closest_image_device walks an image's rays one 8x8 tile per
    wave; every hit must equal the oracle's and the linear entry point's, at
    its own index, also where the image is not a whole number of tiles."""
    print("case: image entry point, config 1 256x256", flush=True)
    s, o = setup(ctx, 1)
    rays = s.jittered_rays(0x5354524B00000001)
    ref, _ = o.trace(rays, MODE_REF)
    assert same(trace_image(ctx, rays, 256, 256), ref)
    assert same(trace_image(ctx, rays, 256, 256, binary=True), ref)  # the packed binary walk
    assert same(trace_image(ctx, rays, 250, 262, binary=True), ref[:250 * 262].copy())
    print("case: image entry point, config 1 ragged sizes", flush=True)
    for w, h in ((13, 7), (1, 1), (1000, 65), (7, 9000), (0, 5), (5, 0)):
        n = w * h
        assert same(trace_image(ctx, rays, w, h), ref[:n].copy()), f"{w}x{h}"
    print("case: image entry point, config 2 1920x1080", flush=True)
    s, o = setup(ctx, 2)
    rays = s.jittered_rays(0x5354524B00000002)
    tiled = trace_image(ctx, rays, 1920, 1080)
    assert same(tiled, ctx.trace_closest(rays))
    sub = np.arange(0, rays.size, 41)
    ref, _ = o.trace(rays[sub].copy(), MODE_PRUNED)
    assert same(tiled[sub].copy(), ref)
    d = ctx.alloc(32)
    try:
        with pytest.raises(sa.StkError):
            ctx.trace_closest_image_device(d, 65536, 65536, d)
        with pytest.raises(sa.StkError):
            ctx.trace_closest_image_device(0, 4, 4, d)
    finally:
        ctx.free(d)


@pytest.mark.parametrize("sampling", [0, 1, 2, 3])
def test_device_entry_points_and_camera_rays(ctx, sampling):
    """Every stratifiedSamplingType (0 centre, the reference's default in
    src/hdRunner/main.cpp:419; 1 random; 2 stratified; 3 correlated
    multi-jittered) gives the oracle's generateCameraRay bit for bit."""
    print(f"case: device buffers + camera ray kernel, config 1, sampling type {sampling}", flush=True)
    s, o = setup(ctx, 1)
    cam = s.camera
    cam.stratifiedSamplingType = sampling
    cam.frameNumber = 3
    cam.spp = 2
    n = cam.dimension[0] * cam.dimension[1]
    d_rays = ctx.alloc(n * 32)
    d_hits = ctx.alloc(n * 20)
    try:
        ctx.camera_rays_device(cam, 1, d_rays)
        rays = ctx.download(d_rays, np.zeros(n, dtype=sa.RAY))
        # the oracle's generateCameraRay with the same RNG seed
        import ctypes as C
        from oracle import load
        lib = load()
        exp = np.zeros(1, dtype=sa.RAY)
        for i in list(range(0, n, 61)) + [n - 1]:
            px, py = i % cam.dimension[0], i // cam.dimension[0]
            frame = cam.frameNumber * cam.spp + 1
            st = C.c_uint32(lib.orc_init_rng(px, py, cam.dimension[0], frame))
            lib.orc_camera_ray(C.byref(cam), px, py, C.byref(st), frame, C.c_void_p(exp.ctypes.data))
            assert exp.tobytes() == rays[i:i + 1].tobytes(), f"pixel {px},{py}"
        ctx.trace_closest_device(d_rays, n, d_hits)
        ctx.sync()
        hits = ctx.download(d_hits, np.zeros(n, dtype=sa.HIT))
        ref, _ = o.trace(rays, MODE_REF)
        assert same(hits, ref)
    finally:
        ctx.free(d_rays)
        ctx.free(d_hits)


def test_edge_cases(ctx):
    print("case: edge cases (empty ray batch, tmax, axis-parallel, NaN rays, single triangle)", flush=True)
    s, o = setup(ctx, 1)
    assert ctx.trace_closest(np.zeros(0, dtype=sa.RAY)).size == 0
    rays = s.jittered_rays(5)[:1000].copy()
    rays["o"][:, 3] = 1e-3
    h = ctx.trace_closest(rays)
    assert np.all(h["instId"] == INVALID) and np.all(h["t"] == np.float32(1e9))
    ax = np.zeros(8, dtype=sa.RAY)
    for i, d in enumerate([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (0, 0, 0),
                           (np.nan, 0, 1)]):
        ax[i]["o"] = (0.0, 0.0, 1.5, 1e9)
        ax[i]["d"] = (*d, 0.0)
    ref, _ = o.trace(ax, MODE_REF)
    assert same(ctx.trace_closest(ax), ref)
    assert same(ctx.trace_closest(ax, reference_order=True), ref)
    # one-triangle scene: the root is a leaf
    one = sa.SceneData(vb=s.vb, ib=s.ib, inst=s.inst[:1].copy(), lights=s.lights[:0], mats=s.mats, camera=s.camera,
                       sky=s.sky)
    one.inst["indexCount"][0] = 3
    ctx.set_scene(one)
    ctx.build_bvh()
    o1 = Oracle.from_scene(one, ctx.get_bvh())
    r = s.jittered_rays(9)
    ref, _ = o1.trace(r, MODE_REF)
    assert same(ctx.trace_closest(r), ref)
    # empty scene: every ray misses
    empty = sa.SceneData(vb=s.vb, ib=s.ib, inst=s.inst[:0].copy(), lights=s.lights[:0], mats=s.mats,
                         camera=s.camera, sky=s.sky)
    ctx.set_scene(empty)
    ctx.build_bvh()
    h = ctx.trace_closest(r[:4096])
    assert np.all(h["instId"] == INVALID)
    h = ctx.trace_any(r[:4096])
    assert np.all(h["instId"] == INVALID)


def test_errors_are_reported_not_thrown(ctx):
    print("case: error paths", flush=True)
    c = sa.Context(0)
    with pytest.raises(sa.StkError) as e:
        c.trace_closest(np.zeros(4, dtype=sa.RAY))
    assert e.value.status == 4  # STK_E_STATE: no BVH yet
    with pytest.raises(sa.StkError) as e:
        c.build_bvh()
    assert e.value.status == 4
    s = sa.SceneData.generate(1)
    bad = s.inst.copy()
    bad["indexCount"][0] = 10 ** 8
    with pytest.raises(sa.StkError) as e:
        c.set_scene(sa.SceneData(vb=s.vb, ib=s.ib, inst=bad, lights=s.lights, mats=s.mats, camera=s.camera,
                                 sky=s.sky))
    assert e.value.status == 1
    c.close()


def test_render_image_pitch_and_rgba8():
    """stk_render_image: RGBA32F with a padded pitch equals stk_render's
    frame; RGBA8 stores round(clamp(v, 0, 1) * 255); mapped host memory works
    as a device-visible output."""
    print("case: stk_render_image pitch/RGBA8, config 6", flush=True)
    s = sa.SceneData.generate(6)
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    w, h = s.camera.dimension[0], s.camera.dimension[1]

    def params():
        return sa.render_params(s.camera, spp=2, max_depth=8, frame=4, sampling=1, sky=s.sky, accumulate=False)

    ref = ctx.render(params()).copy()
    f32 = ctx.render_image(params(), pitch=w * 16 + 48)
    assert np.array_equal(f32, ref)
    u8 = ctx.render_image(params(), rgba8=True, pitch=w * 4 + 12)
    expect = (np.clip(np.nan_to_num(ref, nan=0.0), 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
    assert np.array_equal(u8, expect)
    hp = ctx.host_alloc(w * h * 16)
    ctx.render_device(params(), hp)
    ctx.sync()
    got = np.ctypeslib.as_array((ctypes.c_float * (w * h * 4)).from_address(hp)).reshape(h, w, 4).copy()
    ctx.host_free(hp)
    ctx.close()
    assert np.array_equal(got, ref)


def test_failed_tree_updates_keep_the_previous_tree(ctx):
    """A rejected stk_set_bvh, stk_build_bvh or stk_build_two_level leaves the
    context tracing the tree it had (host description and device arrays of the
    same tree), so the next trace returns the previous hits."""
    print("case: config 1, rejected set_bvh / build_bvh / build_two_level between two traces", flush=True)
    s, o = setup(ctx, 1)
    rays = s.jittered_rays(0x5354524B00000001)[::8].copy()
    before = ctx.trace_closest(rays)
    info0 = ctx.bvh_info()
    nodes = ctx.get_bvh()
    broken = nodes.copy()
    broken[1]["nodeOffset"] = 10 ** 8  # a skip link past the array
    with pytest.raises(sa.StkError) as e:
        ctx.set_bvh(broken)
    assert e.value.status == 1
    assert same(ctx.trace_closest(rays), before)
    with pytest.raises(sa.StkError):
        ctx.build_bvh(max_depth=40)  # deeper than the walks' stacks allow
    assert same(ctx.trace_closest(rays), before)
    info1 = ctx.bvh_info()
    assert (info1.wideDepth, info1.numWide, info1.numNodes, info1.walkStack) == \
        (info0.wideDepth, info0.numWide, info0.numNodes, info0.walkStack)
    # a valid replacement still takes effect
    ctx.set_bvh(nodes)
    assert same(ctx.trace_closest(rays), before)


def test_spatial_split_tree_walks_match_the_oracle():
    """A tree with spatial splits (a triangle in several leaves): the GPU's
    wide, binary and reference-order walks return the oracle's restatements'
    hits bit for bit on a config-2 subsample, and a path-traced config-6
    frame over such a tree keeps every sample within 1e-4 of the oracle."""
    from oracle.pyoracle import MODE_PRUNED, MODE_REF, MODE_WIDE
    from test_gpu_pathtrace import compare_samples

    print("case: config 2 built with spatial splits (alpha 1e-5), rays[::53]", flush=True)
    s = sa.SceneData.generate(2)
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh(spatial_split_alpha=1e-5)
    nodes = ctx.get_bvh()
    b = sa.HostBvh.from_nodes(s, nodes)
    o = Oracle.from_scene(s, b.nodes, b.wide, b.wide_tris)
    rays = s.jittered_rays(0x5354524B00000001)[::53].copy()
    assert ctx.trace_closest(rays).tobytes() == o.trace(rays, MODE_WIDE)[0].tobytes()
    assert ctx.trace_closest(rays, binary=True).tobytes() == o.trace(rays, MODE_PRUNED)[0].tobytes()
    assert ctx.trace_closest(rays, reference_order=True).tobytes() == o.trace(rays, MODE_REF)[0].tobytes()
    ctx.close()

    print("case: config 6 path samples over a spatial-split tree vs oracle", flush=True)
    s6 = sa.SceneData.generate(6)
    ctx = sa.Context(0)
    ctx.set_scene(s6)
    ctx.build_bvh(spatial_split_alpha=1e-5)
    o6 = Oracle.from_scene(s6, ctx.get_bvh())
    p = sa.render_params(s6.camera, spp=1, max_depth=6, frame=0, sampling=1, sky=s6.sky, keep_samples=True)
    w, h = p.pt.dimension[0], p.pt.dimension[1]
    ctx.render(p)
    g = ctx.samples(w * h * 3)
    ctx.close()
    frac, err = compare_samples(g, o6.pt_frame(p, 0, 0, w, h))
    print(f"divergent samples {frac}, max error {err}")
    assert frac == 0


def test_tree_loaded_from_a_file_traces_like_the_built_tree(ctx, tmp_path):
    """A tree written to a BVH file (stk_bvh_save_nodes) and read back
    (stk_bvh_load_nodes) into stk_set_bvh: every config-1 ray's hit equals the
    oracle's on the built tree, and a config-6 path-traced frame over the
    loaded tree equals the frame over the built tree bit for bit."""
    print("case: config 1 rays and a config-6 frame over trees loaded from BVH files", flush=True)
    s, o = setup(ctx, 1)
    path = tmp_path / "config1.stkbvh"
    sa.save_bvh(path, ctx.get_bvh())
    rays = s.jittered_rays(0x5354524B00000001)
    ref, _ = o.trace(rays, MODE_REF)
    ctx.build_bvh(bins=4)  # a different tree in between
    ctx.set_bvh(sa.load_bvh(path))
    assert same(ctx.trace_closest(rays), ref)
    s6 = sa.SceneData.generate(6)
    ctx.set_scene(s6)
    ctx.build_bvh()
    p = sa.render_params(s6.camera, spp=4, max_depth=8, frame=2, sampling=1, sky=s6.sky, accumulate=False)
    built = ctx.render(p).copy()
    path6 = tmp_path / "config6.stkbvh"
    sa.save_bvh(path6, ctx.get_bvh())
    ctx.set_scene(s6)
    ctx.set_bvh(sa.load_bvh(path6))
    assert np.array_equal(ctx.render(p), built)
