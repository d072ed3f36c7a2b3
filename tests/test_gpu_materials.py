"""Specular, emissive and normal-mapped materials and the mirrored-repeat and
clip wraps on the GPU: per-sample radiance of the showcase scene against the
oracle in both estimator modes, the convex furnaces' exact values, the
persistent frame launch and the batch bit-identical to the per-stage single
renders, and the validation of material and texture records."""
import numpy as np
import pytest

import strelka_amd as sa
from material_scenes import furnace, material, showcase
from oracle import Oracle

from test_gpu_pathtrace import MAX_DIVERGENT, compare_samples

pytestmark = pytest.mark.gpu


def params(s, spp=2, frame=2, ref_compat=True, max_depth=8, rr_start=3):
    return sa.render_params(s.camera, spp=spp, max_depth=max_depth, frame=frame, sampling=1, sky=s.sky,
                            keep_samples=True, ref_compat=ref_compat, rr_start=rr_start)


def render_samples(s, p, launch=None):
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    if launch:
        ctx.set_launch_config(**launch)
    w, h = p.pt.dimension[0], p.pt.dimension[1]
    ctx.reset_accumulation()
    img = ctx.render(p)
    smp = ctx.samples(w * h * p.pt.spp * 3)
    nodes = ctx.get_bvh()
    ctx.close()
    return img, smp, nodes


@pytest.mark.parametrize("ref_compat", [True, False])
def test_showcase_samples_match_oracle(ref_compat):
    print(f"case: showcase scene (mirror, glass, emitters, normal maps, four wraps), 4 spp, refCompat {ref_compat}, "
          "every sample vs the oracle", flush=True)
    s = showcase()
    p = params(s, spp=4, ref_compat=ref_compat)
    _, smp, nodes = render_samples(s, p)
    o = Oracle.from_scene(s, nodes)
    w, h = p.pt.dimension[0], p.pt.dimension[1]
    cpu = o.pt_frame(p, 0, 0, w, h)
    frac, _ = compare_samples(smp, cpu)
    print(f"divergent samples {frac:.5f}, mean {cpu.mean():.4f}")
    assert frac <= MAX_DIVERGENT
    plain = sa.SceneData.generate(6)
    _, base, _ = render_samples(plain, params(plain, spp=4, ref_compat=ref_compat))
    assert not np.array_equal(base, smp)


@pytest.mark.parametrize("ref_compat", [True, False])
def test_furnaces_are_exact(ref_compat):
    print(f"case: mirror, clear glass, tinted glass and emitting cubes under a constant sky, refCompat {ref_compat}",
          flush=True)
    for mat, expect in ((material(sa.MAT_MIRROR), lambda S: [S]),
                        (material(sa.MAT_GLASS, ior=1.5), lambda S: [S]),
                        (material(sa.MAT_GLASS, albedo=(0.9, 0.6, 0.3), ior=1.5),
                         lambda S: [S, (np.float32([0.9, 0.6, 0.3]) * np.float32([0.9, 0.6, 0.3])) * S]),
                        (material(sa.MAT_LAMBERT, albedo=(0.5, 0.25, 0.75), emission=(2.0, 3.0, 0.5)),
                         lambda S: [S, np.float32([2.0, 3.0, 0.5]) + np.float32([0.5, 0.25, 0.75]) * S,
                                    np.float32([2.0, 3.0, 0.5])])):
        s = furnace(mat)
        specular = int(mat["type"][0]) in (sa.MAT_MIRROR, sa.MAT_GLASS)
        p = params(s, spp=4, ref_compat=ref_compat, max_depth=32 if specular else 8, rr_start=64 if specular else 3)
        _, smp, _ = render_samples(s, p)
        smp = smp.reshape(-1, 3)
        S = np.array(s.sky[:3], np.float32)
        kinds = [np.all(smp == v, axis=1) for v in expect(S)]
        zero = np.all(smp == 0, axis=1) if specular else np.zeros(len(smp), bool)
        fr = [round(float(k.mean()), 4) for k in kinds]
        print(f"type {int(mat['type'][0])}: fractions {fr}, zero {zero.mean():.5f}")
        assert np.all(np.any(kinds, axis=0) | zero)
        assert zero.mean() < 2e-3
        assert all(k.mean() > 0.02 for k in kinds[:2])  # sky, and the hit value (absorbed bounces are rare)


def test_frame_launch_and_shapes_equal_per_stage_launches():
    print("case: showcase scene, persistent frame launches and pass streams vs per-stage, 3 frames, bit-exact",
          flush=True)
    s = showcase()
    shapes = [dict(block_size=64, tile=8), dict(block_size=64, tile=8, frame_blocks=512),
              dict(block_size=256, tile=4, frame_blocks=5, frame_chunk=1024),
              dict(block_size=128, tile=8, pass_streams=2, persistent_blocks=256),
              dict(block_size=64, tile=8, binary_walk=True)]
    out = []
    for launch in shapes:
        ctx = sa.Context(0)
        ctx.set_scene(s)
        ctx.build_bvh()
        ctx.set_launch_config(**launch)
        w, h = s.camera.dimension[0], s.camera.dimension[1]
        frames = []
        for f in range(3):
            p = sa.render_params(s.camera, spp=3, max_depth=8, frame=f, sampling=1, sky=s.sky, keep_samples=True,
                                 ref_compat=False)
            frames.append((ctx.render(p).tobytes(), ctx.samples(w * h * 9).tobytes()))
        ctx.close()
        out.append(frames)
    for launch, frames in zip(shapes[1:], out[1:]):
        print(f"  launch {launch}", flush=True)
        assert frames == out[0]


def test_batch_view_equals_single_render():
    print("case: batch of the untextured showcase scene and config 7, 2 spp, bit-exact vs single renders",
          flush=True)
    scenes = [showcase(textured=False), sa.SceneData.generate(7)]
    sky = (0.3, 0.4, 0.5)
    p = sa.render_params(scenes[0].camera, spp=2, max_depth=8, frame=3, sampling=1, sky=sky, keep_samples=True,
                         enable_tonemap=False)
    ctx = sa.Context(0)
    ctx.batch_clear()
    for t in scenes:
        ctx.batch_add(t, sky=sky)
    ctx.batch_commit()
    views = ctx.batch_views()
    ctx.batch_render(p)
    smp = ctx.samples(ctx.batch_pixels() * 2 * 3)
    ctx.close()
    for i, t in enumerate(scenes):
        q = sa.render_params(t.camera, spp=2, max_depth=8, frame=3, sampling=1, sky=sky, keep_samples=True,
                             enable_tonemap=False)
        _, ref, _ = render_samples(t, q)
        v = views[i]
        n = v["width"] * v["height"]
        assert np.array_equal(smp[v["pixel_base"] * 6:(v["pixel_base"] + n) * 6], ref), i


def test_material_and_texture_validation():
    print("case: material and texture record validation", flush=True)
    ctx = sa.Context(0)
    for bad in (material(7), material(sa.MAT_GLASS, ior=-1.0), material(sa.MAT_GLASS, ior=float("nan"))):
        with pytest.raises(sa.StkError):
            ctx.set_materials(bad)
    ctx.set_materials(material(sa.MAT_GLASS, ior=1.7))
    t = np.zeros(1, sa.TEXTURE)
    for wrap in (sa.TEX_MIRRORED_REPEAT, sa.TEX_CLIP):
        t[0] = (2, 2, 0, wrap)
        ctx.set_textures(t, np.ones((4, 4), np.float32))
    t[0] = (2, 2, 0, 4)
    with pytest.raises(sa.StkError):
        ctx.set_textures(t, np.ones((4, 4), np.float32))
    ctx.batch_clear()
    with pytest.raises(sa.StkError):
        ctx.batch_add(showcase())  # textured materials serve single-scene renders
    ctx.close()
