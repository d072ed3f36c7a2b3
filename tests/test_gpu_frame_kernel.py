"""The persistent frame launch (StkLaunchConfig::frameBlocks): one launch per
frame whose resident workgroups pull chunks of slots and run all of a chunk's
samples with queues of their own.  Every image, kept sample and ray total must
equal the per-stage launches' bit for bit: on the flat tree in both estimator
modes, on the two-level structure, with material textures, over several
progressive frames; the radiance cache keeps the per-stage launches; the
watchdog abandons a persistent frame like any other.  (Batches: the frame
shapes in test_gpu_batch.py::SHAPES.)"""
import time

import numpy as np
import pytest

import strelka_amd as sa

from test_gpu_textures import checker_and_gradient, textured

pytestmark = pytest.mark.gpu

STAGES = dict(block_size=64, tile=8)
FRAMES = [dict(block_size=64, tile=8, frame_blocks=4096), dict(block_size=256, tile=8, frame_blocks=5, frame_chunk=1024),
          dict(block_size=64, tile=4, frame_blocks=256, frame_chunk=4096, binary_walk=True),
          dict(block_size=256, tile=8, frame_blocks=1024, frame_chunk=256)]


def run(s, launch, frames=3, two_level=False, **kw):
    ctx = sa.Context(0)
    ctx.set_scene(s)
    if two_level:
        ctx.build_two_level()
    else:
        ctx.build_bvh()
    ctx.set_launch_config(**launch)
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    out = []
    for f in range(frames):
        p = sa.render_params(s.camera, spp=3, max_depth=8, frame=f, sampling=1, sky=s.sky, keep_samples=True, **kw)
        img = ctx.render(p)
        out.append((img, ctx.samples(w * h * 3 * 3), ctx.ray_totals()))
    ctx.close()
    return out


def same_runs(a, b):
    for (ia, sa_, ta), (ib, sb, tb) in zip(a, b):
        assert ia.tobytes() == ib.tobytes()
        assert sa_.tobytes() == sb.tobytes()
        assert ta == tb


@pytest.mark.parametrize("ref_compat", [True, False])
def test_flat_tree_images_equal_per_stage_launches(ref_compat):
    print(f"case: config 6, 3 frames of 3 spp, refCompat {ref_compat}, frame launches vs per-stage", flush=True)
    s = sa.SceneData.generate(6)
    base = run(s, STAGES, ref_compat=ref_compat)
    assert base[-1][2][0] > 0 and base[-1][2][1] > 0
    for launch in FRAMES:
        print(f"  launch {launch}", flush=True)
        same_runs(run(s, launch, ref_compat=ref_compat), base)


def test_two_level_and_textures_equal_per_stage_launches():
    print("case: 60-instance two-level scene; textured config 6; frame launches vs per-stage", flush=True)
    s = sa.SceneData.generate(5, 60)
    same_runs(run(s, FRAMES[0], frames=2, two_level=True), run(s, STAGES, frames=2, two_level=True))
    textures, texels = checker_and_gradient()
    t = textured(textures, texels, lambda m: 1 + m % 2, rough=lambda m: 2 - m % 2)
    same_runs(run(t, FRAMES[1], frames=2), run(t, STAGES, frames=2))


def test_cache_renders_keep_the_per_stage_launches():
    print("case: query-only cache frame with frame_blocks set equals the per-stage frame", flush=True)
    s = sa.SceneData.generate(6)
    imgs = []
    for launch in (STAGES, FRAMES[0]):
        ctx = sa.Context(0)
        ctx.set_scene(s)
        ctx.build_bvh()
        ctx.set_launch_config(**launch)
        ctx.cache_create(log2_table=14, seed=9)
        p = sa.render_params(s.camera, spp=1, max_depth=6, frame=0, sampling=1, sky=s.sky, cache_mode=2,
                             cache_query_depth=1)
        imgs.append(ctx.render(p))
        ctx.close()
    assert imgs[0].tobytes() == imgs[1].tobytes()


def test_watchdog_abandons_a_persistent_frame():
    print("case: 2-ms watchdog on a 16-spp 1080p config-2 frame in one persistent launch", flush=True)
    s = sa.SceneData.generate(2)
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    ctx.set_launch_config(block_size=64, tile=8, frame_blocks=768)
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    p = lambda f: sa.render_params(s.camera, spp=16, max_depth=8, frame=f, sampling=1, sky=s.sky)
    ctx.render(p(0))
    t0 = time.perf_counter()
    ctx.render(p(1))
    full = time.perf_counter() - t0
    ctx.reset_accumulation()
    ctx.render(p(0))
    before = ctx.accumulation(w, h)
    ctx.set_watchdog(2)
    t0 = time.perf_counter()
    with pytest.raises(sa.StkError) as e:
        ctx.render(p(1))
    cut = time.perf_counter() - t0
    print(f"full frame {full * 1e3:.1f} ms, abandoned after {cut * 1e3:.1f} ms")
    assert e.value.status == 5 and cut < 0.5 * full
    assert ctx.accumulation(w, h).tobytes() == before.tobytes()
    ctx.set_watchdog(0)
    got = ctx.render(p(1))
    ctx.close()
    ref = sa.Context(0)
    ref.set_scene(s)
    ref.build_bvh()
    ref.render(p(0))
    assert ref.render(p(1)).tobytes() == got.tobytes()
    ref.close()
