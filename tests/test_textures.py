"""Albedo-texture lookup of the oracle (orc_texture_lookup), pinned by what
the filter's definition fixes (shaders/mdl_runtime.hlsl:93-207: wrap, the
quintic smootherstep of the fractional texel position, then a bilinear
blend): texel centres return their texel, the midpoint between two texels
their mean, repeat is periodic, clamp holds the edge texels, a constant
texture is constant, and in between the weight is the quintic 6f^5 - 15f^4 +
10f^3 of the offset.  The GPU lookup is compared with this one through
rendered samples (test_gpu_textures.py)."""
import numpy as np
import pytest

from oracle.pyoracle import texture_lookup

TEX = np.dtype([("width", "<u4"), ("height", "<u4"), ("offset", "<u4"), ("wrap", "<u4")])


def tex(w, h, offset=0, wrap=0):
    r = np.zeros((), TEX)
    r["width"], r["height"], r["offset"], r["wrap"] = w, h, offset, wrap
    return r


def texels(w, h, seed=0, pad=0):
    rng = np.random.default_rng(seed)
    return np.concatenate([np.full((pad, 4), 7.0, np.float32),
                           rng.uniform(0, 1, (w * h, 4)).astype(np.float32)])


@pytest.mark.parametrize("wrap", [0, 1])
def test_texel_centres_return_their_texel(wrap):
    w, h, pad = 8, 4, 5
    t = texels(w, h, 1, pad)
    for j in range(h):
        for i in range(w):
            got = texture_lookup(tex(w, h, pad, wrap), t, (i + 0.5) / w, (j + 0.5) / h)
            assert np.array_equal(got, t[pad + j * w + i, :3]), (i, j)


def test_midpoints_are_means_and_quintic_weights_in_between():
    w, h = 4, 1
    t = np.zeros((w * h, 4), np.float32)
    t[1, :3] = 1.0
    tx = tex(w, h)
    # halfway between the centres of texels 0 and 1: weight 1/2
    assert np.array_equal(texture_lookup(tx, t, 1.0 / w, 0.5), np.full(3, 0.5, np.float32))
    for f in (0.125, 0.25, 0.3, 0.75, 0.9):
        s = (0.5 + f) / w  # offset f from texel 0's centre towards texel 1
        q = f ** 3 * (f * (f * 6 - 15) + 10)
        got = texture_lookup(tx, t, s, 0.5)
        assert np.allclose(got, q, atol=2e-6), (f, got, q)


def test_repeat_is_periodic_and_clamp_holds_the_edges():
    w, h = 8, 8
    t = texels(w, h, 2)
    rep, clamp = tex(w, h, 0, 0), tex(w, h, 0, 1)
    rng = np.random.default_rng(3)
    for s, tt in rng.uniform(0, 1, (20, 2)):
        base = texture_lookup(rep, t, s, tt)
        for k in (-3, -1, 1, 2):
            assert np.allclose(texture_lookup(rep, t, s + k, tt + k), base, atol=1e-5)
    for s, tt in ((-2.0, 0.3), (1.7, 0.3), (0.3, -5.0), (0.3, 9.0)):
        cs, ct = min(max(s, 0.5 / w), 1 - 0.5 / w), min(max(tt, 0.5 / h), 1 - 0.5 / h)
        assert np.array_equal(texture_lookup(clamp, t, s, tt), texture_lookup(clamp, t, cs, ct))
    # outside, a clamped lookup equals the edge texel centre's value
    assert np.array_equal(texture_lookup(clamp, t, -1.0, 0.5 / h), t[0, :3])


def test_constant_texture_is_constant():
    w, h = 5, 3
    t = np.tile(np.array([[0.25, 0.5, 0.75, 1.0]], np.float32), (w * h, 1))
    rng = np.random.default_rng(4)
    for wrap in (0, 1):
        for s, tt in rng.uniform(-10, 10, (50, 2)):
            assert np.array_equal(texture_lookup(tex(w, h, 0, wrap), t, s, tt),
                                  np.array([0.25, 0.5, 0.75], np.float32))
