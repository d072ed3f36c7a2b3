"""The GPU path tracer against closed forms (FP64, tests/closed_forms.py):
the convex furnace, whose every sample is exactly albedo * sky or sky, and
the converged lit floor, whose pixel means are FP64 quadratures of the
estimator's expectation (the FP64 CPU reference of the converged image)."""
import numpy as np
import pytest

import strelka_amd as sa
from closed_forms import floor_expectations

pytestmark = pytest.mark.gpu


def test_convex_furnace_samples_are_exact_on_gpu():
    print("case: convex furnace (config 8), 96x96, 8 spp", flush=True)
    s = sa.SceneData.generate(8)
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    p = sa.render_params(s.camera, spp=8, max_depth=8, frame=3, sampling=1, sky=s.sky, keep_samples=True)
    ctx.render(p)
    smp = ctx.samples(w * h * 8 * 3).reshape(-1, 3)
    ctx.close()
    sky = np.array(s.sky[:3], np.float32)
    hit_val = np.array(s.mats["albedo"][0, :3], np.float32) * sky
    is_sky = np.all(smp == sky, axis=1)
    is_hit = np.all(smp == hit_val, axis=1)
    absorbed = np.all(smp == 0, axis=1)
    print(f"sky {is_sky.mean():.4f}, albedo*sky {is_hit.mean():.4f}, absorbed {absorbed.mean():.5f}")
    assert is_sky.mean() > 0.2 and is_hit.mean() > 0.2
    assert np.all(is_sky | is_hit | absorbed)
    assert absorbed.mean() < 1e-3


@pytest.mark.parametrize("ref_compat", [True, False])
def test_lit_floor_converges_to_fp64_quadrature_on_gpu(ref_compat):
    """64 frames of 64 spp of every pixel of the lit floor: each pixel's
    t statistic (mean minus closed form over the standard error of the 64
    frame means, 63 degrees of freedom) must stay below the two-sided
    quantile for a family-wise error of 1e-3 over the 4,096 pixels, and the
    mean relative deviation over the image (a bias) below 0.1%.  The closed
    form is the reference estimator's expectation (ref_compat) or, for the
    corrected estimator, the radiance the floor reflects (Lambert's polygon
    formula for the light)."""
    from scipy.stats import t as student_t

    print(f"case: lit floor (config 9), 64x64, 64 x 64 spp vs FP64 quadrature, ref_compat={ref_compat}", flush=True)
    s = sa.SceneData.generate(9)
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    frames = []
    for f in range(64):
        p = sa.render_params(s.camera, spp=64, max_depth=8, frame=f, sampling=1, sky=(0, 0, 0),
                             accumulate=False, enable_tonemap=False, ref_compat=ref_compat)
        frames.append(ctx.render(p)[..., 0].astype(np.float64).copy())
    ctx.close()
    frames = np.array(frames)
    mean = frames.mean(axis=0)
    se = frames.std(axis=0, ddof=1) / np.sqrt(len(frames))
    pix = [(x, y) for y in range(h) for x in range(w)]
    expect = floor_expectations(s, pix, corrected=not ref_compat).reshape(h, w)
    z = (mean - expect) / np.maximum(se, 1e-12)
    rel = (mean - expect) / expect
    print(f"max |z| {np.abs(z).max():.2f}, mean relative deviation {rel.mean():.2e}, "
          f"median SE/mean {np.median(se / mean):.2e}")
    assert np.abs(z).max() < student_t.isf(1e-3 / (2 * z.size), len(frames) - 1)
    assert abs(rel.mean()) < 1e-3
