"""Material textures on the GPU (stk_set_textures + StkMaterial.albedoTexture
and roughnessTexture):
per-sample radiance against the oracle path tracer with the same textures, a
white 1x1 texture leaving every sample bit-identical to the untextured
render, and the error paths."""
import numpy as np
import pytest

import strelka_amd as sa
from oracle import Oracle

from test_gpu_pathtrace import MAX_DIVERGENT, compare_samples

pytestmark = pytest.mark.gpu


def textured(textures, texels, pick, rough=lambda m: 0):
    s = sa.SceneData.generate(6)
    mats = s.mats.copy()
    for m in range(mats.size):
        mats["albedoTexture"][m] = pick(m)
        mats["roughnessTexture"][m] = rough(m)
    s.mats, s.textures, s.texels = mats, textures, texels
    return s


def checker_and_gradient():
    cw = 16
    ii, jj = np.meshgrid(np.arange(cw), np.arange(cw))
    checker = np.where((ii + jj) % 2 == 0, 1.0, 0.2).astype(np.float32)
    c = np.stack([checker, checker * 0.9, checker * 0.7, np.ones_like(checker)], -1).reshape(-1, 4)
    gw, gh = 8, 4
    gx, gy = np.meshgrid(np.linspace(0.1, 1.0, gw), np.linspace(0.3, 1.0, gh))
    g = np.stack([gx, gy, 1.0 - 0.5 * gx, np.ones_like(gx)], -1).reshape(-1, 4).astype(np.float32)
    textures = np.zeros(2, sa.TEXTURE)
    textures[0] = (cw, cw, 0, 0)           # repeat
    textures[1] = (gw, gh, cw * cw, 1)     # clamp
    return textures, np.concatenate([c, g])


def params(s, frame=2):
    return sa.render_params(s.camera, spp=2, max_depth=6, frame=frame, sampling=1, sky=s.sky, keep_samples=True)


def render_samples(s):
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    p = params(s)
    w, h = p.pt.dimension[0], p.pt.dimension[1]
    ctx.reset_accumulation()
    ctx.render(p)
    smp = ctx.samples(w * h * 2 * 3).reshape(h, w, 2, 3)
    nodes = ctx.get_bvh()
    ctx.close()
    return smp, nodes


def test_textured_samples_match_oracle():
    print("case: config 6 with a checker (repeat) and a gradient (clamp) texture, 2 spp, crop vs oracle", flush=True)
    textures, texels = checker_and_gradient()
    s = textured(textures, texels, lambda m: 1 + m % 2, rough=lambda m: 2 - m % 2)
    assert np.any(s.mats["type"] == sa.MAT_GGX)
    smp, nodes = render_samples(s)
    o = Oracle.from_scene(s, nodes)
    h, w = smp.shape[:2]
    x0, y0, cw, ch = w // 2 - 16, h // 2 - 8, 32, 24
    cpu = o.pt_frame(params(s), x0, y0, cw, ch)
    frac, _ = compare_samples(smp[y0:y0 + ch, x0:x0 + cw].reshape(-1), cpu)
    print(f"divergent samples {frac:.5f}, mean {cpu.mean():.4f}")
    assert frac <= MAX_DIVERGENT
    assert cpu.mean() > 1e-3
    plain, _ = render_samples(sa.SceneData.generate(6))
    assert not np.array_equal(plain, smp)  # the textures change the image


def test_white_texture_changes_nothing():
    print("case: config 6, every material on a 1x1 white texture, bit-exact vs untextured", flush=True)
    textures = np.zeros(1, sa.TEXTURE)
    textures[0] = (1, 1, 0, 0)
    white = np.ones((1, 4), np.float32)
    s = textured(textures, white, lambda m: 1, rough=lambda m: 1)
    smp, _ = render_samples(s)
    plain, _ = render_samples(sa.SceneData.generate(6))
    assert np.array_equal(smp, plain)


def test_texture_error_paths():
    print("case: texture validation and the batch's refusal of textured materials", flush=True)
    ctx = sa.Context(0)
    bad = np.zeros(1, sa.TEXTURE)
    bad[0] = (4, 4, 10, 0)  # 16 texels from 10: outside a 20-texel array
    with pytest.raises(sa.StkError):
        ctx.set_textures(bad, np.zeros((20, 4), np.float32))
    bad[0] = (4, 4, 0, 7)   # unknown wrap
    with pytest.raises(sa.StkError):
        ctx.set_textures(bad, np.zeros((20, 4), np.float32))
    textures, texels = checker_and_gradient()
    s = textured(textures, texels, lambda m: 1)
    with pytest.raises(sa.StkError):
        ctx.batch_add(s)
    ctx.close()
