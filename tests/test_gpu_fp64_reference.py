"""The converged GPU render against the binary64 CPU reference render
(oracle/oracle_fp64.c): the 1.16M-triangle scene's view at 96x54 (the 1080p
camera, each pixel covering 20x20 of the full frame), 4,096 spp on the GPU
and 512 spp of the binary64 twin per pixel, for the reference's estimator
and the corrected one.  The per-pixel means must agree within their
statistical error: the mean of z^2 over the 15,552 pixel channels within
[0.7, 1.4] (1 for unbiased estimates), at most 0.2% of them beyond 4.5
standard errors, and the RMSE of the means within 1.4 times the RMS
standard error of the difference - the stated RMSE of the FP32 GPU render
against the FP64 reference is the noise of the two estimates, and the test
prints it."""
import ctypes as C

import numpy as np
import pytest

import strelka_amd as sa
from material_scenes import showcase
from oracle import Oracle

pytestmark = pytest.mark.gpu

W, H = 96, 54
GPU_SPP, FP64_SPP = 4096, 512


@pytest.mark.parametrize("ref_compat", [True, False])
def test_converged_render_matches_fp64_reference(ref_compat):
    print(f"case: config 2 view at {W}x{H}, GPU {GPU_SPP} spp vs FP64 CPU {FP64_SPP} spp, ref_compat={ref_compat}",
          flush=True)
    s = sa.SceneData.generate(2)
    cam = s.camera
    cam.dimension[0], cam.dimension[1] = W, H
    cam.invDimension[0], cam.invDimension[1] = 1.0 / W, 1.0 / H
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    nodes = ctx.get_bvh()
    # the GPU estimate: 16 frames of 256 spp, per-pixel mean and its standard
    # error from the per-sample variance
    s1 = np.zeros((H, W, 3), np.float64)
    s2 = np.zeros((H, W, 3), np.float64)
    per = 256
    for f in range(GPU_SPP // per):
        p = sa.render_params(cam, spp=per, max_depth=8, frame=f, sampling=1, sky=s.sky, keep_samples=True,
                             accumulate=False, enable_tonemap=False, ref_compat=ref_compat)
        ctx.render(p)
        smp = ctx.samples(W * H * per * 3).reshape(H, W, per, 3).astype(np.float64)
        s1 += smp.sum(axis=2)
        s2 += (smp * smp).sum(axis=2)
    ctx.close()
    mg = s1 / GPU_SPP
    vg = np.maximum(s2 / GPU_SPP - mg * mg, 0.0) * GPU_SPP / (GPU_SPP - 1) / GPU_SPP
    o = Oracle.from_scene(s, nodes)
    p = sa.render_params(cam, spp=FP64_SPP, max_depth=8, frame=100000, sampling=1, sky=s.sky, ref_compat=ref_compat)
    m64, v64 = o.pt_frame_fp64(p, 0, 0, W, H)
    var = vg + v64
    ok = var > 0
    z = (mg - m64)[ok] / np.sqrt(var[ok])
    chi = float(np.mean(z * z))
    tail = float(np.mean(np.abs(z) > 4.5))
    rmse = float(np.sqrt(np.mean((mg - m64) ** 2)))
    noise = float(np.sqrt(np.mean(var)))
    print(f"mean pixel {m64.mean():.4f}; RMSE of GPU vs FP64 means {rmse:.5f} (RMS standard error {noise:.5f}); "
          f"mean z^2 {chi:.3f}; |z| > 4.5 on {tail:.5f}")
    assert 0.7 < chi < 1.4 and tail <= 0.002
    assert rmse < 1.4 * noise


@pytest.mark.parametrize("ref_compat", [True, False])
def test_specular_and_emissive_materials_converge_to_fp64_reference(ref_compat):
    """The untextured showcase scene (mirror, clear and tinted glass, an
    emitting sphere, GGX and Lambert) at 64x48: 8,192 spp on the GPU against
    8,192 spp of the binary64 twin, with the same statistical criteria."""
    spp = 8192
    print(f"case: untextured showcase scene, GPU {spp} spp vs FP64 CPU {spp} spp, ref_compat={ref_compat}", flush=True)
    s = showcase(textured=False)
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    nodes = ctx.get_bvh()
    s1 = np.zeros((h, w, 3), np.float64)
    s2 = np.zeros((h, w, 3), np.float64)
    per = 1024
    for f in range(spp // per):
        p = sa.render_params(s.camera, spp=per, max_depth=8, frame=f, sampling=1, sky=s.sky, keep_samples=True,
                             accumulate=False, enable_tonemap=False, ref_compat=ref_compat)
        ctx.render(p)
        smp = ctx.samples(w * h * per * 3).reshape(h, w, per, 3).astype(np.float64)
        s1 += smp.sum(axis=2)
        s2 += (smp * smp).sum(axis=2)
    ctx.close()
    mg = s1 / spp
    vg = np.maximum(s2 / spp - mg * mg, 0.0) * spp / (spp - 1) / spp
    o = Oracle.from_scene(s, nodes)
    p = sa.render_params(s.camera, spp=spp, max_depth=8, frame=100000, sampling=1, sky=s.sky, ref_compat=ref_compat)
    m64, v64 = o.pt_frame_fp64(p, 0, 0, w, h)
    var = vg + v64
    ok = var > 0
    z = (mg - m64)[ok] / np.sqrt(var[ok])
    chi = float(np.mean(z * z))
    tail = float(np.mean(np.abs(z) > 4.5))
    rmse = float(np.sqrt(np.mean((mg - m64) ** 2)))
    noise = float(np.sqrt(np.mean(var)))
    print(f"mean pixel {m64.mean():.4f}; RMSE of GPU vs FP64 means {rmse:.5f} (RMS standard error {noise:.5f}); "
          f"mean z^2 {chi:.3f}; |z| > 4.5 on {tail:.5f}")
    assert 0.7 < chi < 1.4 and tail <= 0.002
    assert rmse < 1.4 * noise


@pytest.mark.parametrize("ref_compat", [True, False])
def test_instanced_scene_converges_to_fp64_reference(ref_compat):
    """The large instanced scene (10,000 instances of 20 meshes, 18.8M
    triangles) through the two-level structure at 96x54 (the 4K camera, each
    pixel covering 40x40 of the full frame), 6 bounces: 4,096 spp on the GPU
    against 512 spp of the binary64 twin walking the same scene's instances
    expanded into one tree, with the same statistical criteria - the
    FP64-verified converged plain render of that workload."""
    print(f"case: config 5 view at {W}x{H} over the two-level structure, GPU {GPU_SPP} spp vs FP64 CPU "
          f"{FP64_SPP} spp, ref_compat={ref_compat}", flush=True)
    s = sa.SceneData.generate(5)
    cam = s.camera
    cam.dimension[0], cam.dimension[1] = W, H
    cam.invDimension[0], cam.invDimension[1] = 1.0 / W, 1.0 / H
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_two_level()
    s1 = np.zeros((H, W, 3), np.float64)
    s2 = np.zeros((H, W, 3), np.float64)
    per = 256
    for f in range(GPU_SPP // per):
        p = sa.render_params(cam, spp=per, max_depth=6, frame=f, sampling=1, sky=s.sky, keep_samples=True,
                             accumulate=False, enable_tonemap=False, ref_compat=ref_compat)
        ctx.render(p)
        smp = ctx.samples(W * H * per * 3).reshape(H, W, per, 3).astype(np.float64)
        s1 += smp.sum(axis=2)
        s2 += (smp * smp).sum(axis=2)
    ctx.close()
    mg = s1 / GPU_SPP
    vg = np.maximum(s2 / GPU_SPP - mg * mg, 0.0) * GPU_SPP / (GPU_SPP - 1) / GPU_SPP
    o = Oracle.from_two_level(s, sa.HostTwoLevel.build(s))
    p = sa.render_params(cam, spp=FP64_SPP, max_depth=6, frame=100000, sampling=1, sky=s.sky, ref_compat=ref_compat)
    m64, v64 = o.pt_frame_fp64(p, 0, 0, W, H)
    var = vg + v64
    ok = var > 0
    z = (mg - m64)[ok] / np.sqrt(var[ok])
    chi = float(np.mean(z * z))
    tail = float(np.mean(np.abs(z) > 4.5))
    rmse = float(np.sqrt(np.mean((mg - m64) ** 2)))
    noise = float(np.sqrt(np.mean(var)))
    print(f"mean pixel {m64.mean():.4f}; RMSE of GPU vs FP64 means {rmse:.5f} (RMS standard error {noise:.5f}); "
          f"mean z^2 {chi:.3f}; |z| > 4.5 on {tail:.5f}")
    assert m64.mean() > 1e-2
    assert 0.7 < chi < 1.4 and tail <= 0.002
    assert rmse < 1.4 * noise
