"""Batched rendering, launch shapes and graph replay on the GPU.

A path's random numbers and arithmetic depend only on its view, pixel, frame
and sample, so every view of a batch must reproduce, bit for bit, what a
single-scene render of that view gives; and every launch shape (block size,
raygen tile, persistent extend grid, graph replay) must give the same image.
The single-scene renders are themselves checked against the CPU oracle
(test_gpu_pathtrace.py), which pins the batch to the oracle by transitivity;
one view is also checked against the oracle directly here."""
import ctypes as C

import numpy as np
import pytest

import strelka_amd as sa
from oracle import Oracle

pytestmark = pytest.mark.gpu

SCENES = [(6, 0), (7, 0), (3, 10)]


@pytest.fixture(scope="module")
def scenes():
    return [sa.SceneData.generate(c, v) for c, v in SCENES]


def params(camera, sky, spp=2, frame=3, keep=True, accumulate=True):
    return sa.render_params(camera, spp=spp, max_depth=8, frame=frame, sampling=1, sky=sky, keep_samples=keep,
                            accumulate=accumulate, enable_tonemap=False)


def single_render(s, spp=2, frame=3):
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    p = params(s.camera, s.sky, spp, frame)
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    img = ctx.render(p)
    smp = ctx.samples(w * h * spp * 3)
    bvh = ctx.get_bvh()
    ctx.close()
    return img, smp, bvh


def batch_context(scenes):
    ctx = sa.Context(0)
    ctx.batch_clear()
    for s in scenes:
        ctx.batch_add(s)
    ctx.batch_commit()
    return ctx


def test_batch_views_equal_single_scene_renders(scenes):
    print("case: batch of config 6, config 7 and config 3 variant 10, 2 spp, bit-exact vs single renders",
          flush=True)
    ctx = batch_context(scenes)
    views = ctx.batch_views()
    assert [v["width"] for v in views] == [s.camera.dimension[0] for s in scenes]
    p = params(scenes[0].camera, (0, 0, 0))
    imgs = [x.copy() for x in ctx.batch_render(p)]
    total = ctx.batch_pixels()
    smp = ctx.samples(total * 2 * 3)
    ctx.close()
    for i, s in enumerate(scenes):
        ref_img, ref_smp, _ = single_render(s)
        v = views[i]
        n = v["width"] * v["height"]
        got = smp[v["pixel_base"] * 6:(v["pixel_base"] + n) * 6]
        print(f"view {i}: {v}, mean {ref_img[..., :3].mean():.4f}")
        assert np.array_equal(got, ref_smp), i
        assert np.array_equal(imgs[i], ref_img), i
        assert ref_img[..., :3].mean() > 0.0


def test_batch_view_matches_oracle(scenes):
    print("case: batch view 0 (config 6) per-sample radiance vs the oracle", flush=True)
    ctx = batch_context(scenes)
    p = params(scenes[0].camera, (0, 0, 0), spp=2, frame=5)
    ctx.batch_render(p)
    v = ctx.batch_views()[0]
    w, h = v["width"], v["height"]
    smp = ctx.samples(ctx.batch_pixels() * 2 * 3)[:w * h * 6]
    ctx.close()
    s = scenes[0]
    _, _, bvh = single_render(s, spp=1, frame=0)
    o = Oracle.from_scene(s, bvh)
    cpu = o.pt_frame(params(s.camera, s.sky, spp=2, frame=5), 0, 0, w, h)
    err = np.abs(smp.reshape(-1, 3) - cpu.reshape(-1, 3)) / np.maximum(1.0, np.abs(cpu.reshape(-1, 3)))
    frac = np.any(err > 1e-4, axis=1).mean()
    print(f"divergent samples {frac:.5f}")
    assert frac == 0


SHAPES = [dict(block_size=64, tile=1), dict(block_size=128, tile=16), dict(block_size=256, tile=8),
          dict(block_size=256, tile=8, persistent_blocks=1024), dict(block_size=64, tile=4, persistent_blocks=512),
          dict(block_size=256, tile=8, graph=True), dict(block_size=128, tile=16, persistent_blocks=2048, graph=True),
          dict(block_size=256, tile=8, paths_in_flight=4096),
          dict(block_size=128, tile=8, paths_in_flight=100000, persistent_blocks=1024, graph=True),
          dict(block_size=128, tile=8, persistent_blocks=8192, graph=True),
          dict(block_size=64, tile=1, persistent_blocks=1, paths_in_flight=20000),
          dict(block_size=64, tile=8, stripe_threads=64), dict(block_size=256, tile=8, stripe_threads=2048, graph=True),
          dict(block_size=128, tile=4, stripe_threads=4096, paths_in_flight=20000, persistent_blocks=512),
          dict(block_size=64, tile=8, pass_streams=2), dict(block_size=256, tile=8, pass_streams=2, graph=True),
          dict(block_size=128, tile=8, pass_streams=2, paths_in_flight=20000, persistent_blocks=1024),
          dict(block_size=64, tile=4, pass_streams=3, paths_in_flight=8192),
          dict(block_size=256, tile=8, pass_streams=4, paths_in_flight=50000, stripe_threads=4096),
          dict(block_size=64, tile=8, pass_streams=1, graph=True),
          # one persistent launch per frame: resident workgroups pull chunks of slots
          dict(block_size=64, tile=8, frame_blocks=768), dict(block_size=256, tile=4, frame_blocks=3, frame_chunk=2048),
          dict(block_size=64, tile=8, frame_blocks=512, frame_chunk=1024),
          dict(block_size=128, tile=8, frame_blocks=256, binary_walk=True, pass_streams=4, graph=True)]


def test_launch_shapes_give_identical_images(scenes):
    print("case: launch shapes over the batch, 2 spp, bit-exact", flush=True)
    ctx = batch_context(scenes)
    p = params(scenes[0].camera, (0, 0, 0), accumulate=False)
    base = None
    npix = ctx.batch_pixels()
    for shape in SHAPES:
        print(f"shape {shape}", flush=True)
        ctx.set_launch_config(**shape)
        assert ctx.launch_config()["block_size"] == shape["block_size"]
        img = np.concatenate([x.reshape(-1, 4) for x in ctx.batch_render(p)])
        totals = ctx.ray_totals()
        if base is None:
            base, base_totals = img.copy(), totals
            # every path traces its camera ray; at most 8 extensions and 8 shadow rays per sample
            assert npix * 2 <= totals[0] <= npix * 2 * 8 and 0 < totals[1] <= npix * 2 * 8
        else:
            assert np.array_equal(img, base), shape
            assert totals == base_totals, shape
    ctx.close()


def test_pass_streams_across_render_calls(scenes):
    """Progressive frames issued back to back (device output, no host wait)
    with several pass streams, switching stream counts and single-stream graph
    replay on the way, give every frame, the last frame's samples and its ray
    totals bit for bit as one stream does."""
    print("case: config 3 variant 10, 6 accumulated frames of 2 spp, pass streams 3/3/1/2/2/4", flush=True)
    s = scenes[2]
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    shapes = [dict(block_size=64, tile=8, pass_streams=3, paths_in_flight=4096),
              dict(block_size=64, tile=8, pass_streams=3, paths_in_flight=4096),
              dict(block_size=64, tile=8, graph=True),
              dict(block_size=128, tile=4, pass_streams=2),
              dict(block_size=128, tile=4, pass_streams=2, paths_in_flight=8192),
              dict(block_size=64, tile=8, pass_streams=4, paths_in_flight=4096)]
    ref = []
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    for f in range(len(shapes)):
        ref.append(ctx.render(params(s.camera, s.sky, spp=2, frame=f)).copy())
    ref_smp = ctx.samples(w * h * 2 * 3)
    ref_totals = ctx.ray_totals()
    ctx.close()
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    outs = [ctx.alloc(w * h * 16) for _ in shapes]
    for f, shape in enumerate(shapes):
        ctx.set_launch_config(**shape)
        ctx.render_device(params(s.camera, s.sky, spp=2, frame=f), outs[f])
    smp = ctx.samples(w * h * 2 * 3)
    totals = ctx.ray_totals()
    for f, d in enumerate(outs):
        img = ctx.download(d, np.zeros((h, w, 4), np.float32))
        assert np.array_equal(img, ref[f]), f
        ctx.free(d)
    assert np.array_equal(smp, ref_smp)
    assert totals == ref_totals
    ctx.close()


def test_record_queries_write_only_their_record(scenes):
    """stk_get_launch_config and stk_batch_info fill the caller's record and
    nothing after it (a canary block follows each record)."""
    print("case: launch-config and view-info records followed by canaries", flush=True)
    ctx = batch_context(scenes)
    ctx.set_launch_config(block_size=128, tile=4, stripe_threads=64)
    for cls, call in ((sa._abi.LaunchConfig, lambda p: sa.lib().stk_get_launch_config(ctx._h, p)),
                      (sa._abi.BatchViewInfo, lambda p: sa.lib().stk_batch_info(ctx._h, 1, p))):
        n = C.sizeof(cls)
        buf = (C.c_uint8 * (n + 64))(*([0xA5] * (n + 64)))
        assert call(C.cast(buf, C.POINTER(cls))) == 0
        assert bytes(buf[n:]) == b"\xa5" * 64, cls.__name__
    ctx.close()


def test_graph_replays_match_eager_frames(scenes):
    print("case: graph-captured sample pass, 4 frames of 3 spp with eager work between", flush=True)
    s = scenes[2]
    frames = []
    for graph in (False, True):
        ctx = sa.Context(0)
        ctx.set_scene(s)
        ctx.build_bvh()
        ctx.set_launch_config(block_size=256, tile=8, graph=graph)
        rays = s.jittered_rays(9)[:4096]
        out = []
        for f in range(4):
            img = ctx.render(params(s.camera, s.sky, spp=3, frame=f, keep=False)).copy()
            hits = ctx.trace_closest(rays)  # eager work between replays
            ctx.sync()
            out.append((img, hits))
        frames.append(out)
        ctx.close()
    for f in range(4):
        assert np.array_equal(frames[0][f][0], frames[1][f][0]), f
        assert np.array_equal(frames[0][f][1], frames[1][f][1]), f


def test_batch_rejects_cache_and_bad_shapes(scenes):
    print("case: batch error paths", flush=True)
    ctx = sa.Context(0)
    with pytest.raises(sa.StkError):
        ctx.batch_render(params(scenes[0].camera, (0, 0, 0)))  # nothing committed
    with pytest.raises(sa.StkError):
        ctx.set_launch_config(block_size=96)
    with pytest.raises(sa.StkError):
        ctx.set_launch_config(tile=0)
    ctx.batch_add(scenes[0])
    ctx.batch_commit()
    ctx.set_scene(scenes[0])  # the cache takes its bounds from the single-scene BVH
    ctx.build_bvh()
    ctx.cache_create(log2_table=12)
    p = params(scenes[0].camera, (0, 0, 0))
    p.cacheMode = 1
    with pytest.raises(sa.StkError):
        ctx.batch_render(p)
    ctx.close()


def test_config3_all_32_scenes_against_oracle_and_single_renders():
    """Config 3 at its real workload: the 32 seeded scenes of 10k-2M triangles
    (generate(3, v), v = 0..31), every view 960x540, 8 bounces, one batch.
    A 16x16 crop of every view equals the oracle per sample, and four views
    (the smallest, the largest and two between) equal their single-scene
    renders bit for bit (samples and accumulated image)."""
    print("case: config 3, 32 scenes at 960x540, 2 spp, 8 bounces; oracle crop per view, 4 single renders",
          flush=True)
    scenes = [sa.SceneData.generate(3, v) for v in range(32)]
    tris = [int(s.inst["indexCount"].sum()) // 3 for s in scenes]
    print(f"instanced triangles: min {min(tris)}, max {max(tris)}, total {sum(tris)}", flush=True)
    assert min(tris) < 40_000 and max(tris) > 1_000_000
    ctx = batch_context(scenes)
    views = ctx.batch_views()
    assert len(views) == 32 and all((v["width"], v["height"]) == (960, 540) for v in views)
    p = params(scenes[0].camera, (0, 0, 0), spp=2, frame=7)
    imgs = [x.copy() for x in ctx.batch_render(p)]
    total = ctx.batch_pixels()
    smp = ctx.samples(total * 2 * 3)
    ctx.close()
    x0, y0 = 472, 262
    worst = 0.0
    for i, s in enumerate(scenes):
        v = views[i]
        n = v["width"] * v["height"]
        view = smp[v["pixel_base"] * 6:(v["pixel_base"] + n) * 6].reshape(v["height"], v["width"], 2, 3)
        crop = view[y0:y0 + 16, x0:x0 + 16].reshape(-1, 3)
        o = Oracle.from_scene(s, sa.HostBvh.build(s).nodes)
        cpu = o.pt_frame(params(s.camera, s.sky, spp=2, frame=7), x0, y0, 16, 16).reshape(-1, 3)
        err = np.abs(crop - cpu) / np.maximum(1.0, np.abs(cpu))
        worst = max(worst, float(err.max()))
        assert not np.any(err > 1e-4), (i, float(err.max()))
    print(f"32 views x 512 samples: max error {worst:.3e}")
    order = np.argsort(tris)
    for i in (int(order[0]), int(order[10]), int(order[21]), int(order[-1])):
        ref_img, ref_smp, _ = single_render(scenes[i], spp=2, frame=7)
        v = views[i]
        n = v["width"] * v["height"]
        assert np.array_equal(smp[v["pixel_base"] * 6:(v["pixel_base"] + n) * 6], ref_smp), i
        assert np.array_equal(imgs[i], ref_img), i
        print(f"view {i} ({tris[i]} triangles) equals its single render")
