"""Pins the oracle's primitives: ray/box known answers from the reference's own
(disabled) test, FP64 Moller-Trumbore, the RNG against published values, and
the closed-form camera."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle.pyoracle import RAY, fptr, load


def box_test(o, d, pmin, pmax):
    lib = load()
    r = np.zeros(1, dtype=RAY)
    r["o"][0, :3] = o
    r["d"][0, :3] = d
    inv = np.array([np.float32(1.0 / np.float64(np.float32(x))) for x in d], dtype=np.float32)
    t = C.c_float()
    res = lib.orc_intersect_ray_box(C.c_void_p(r.ctypes.data), fptr(inv), fptr(np.array(pmin, np.float32)),
                                    fptr(np.array(pmax, np.float32)), C.byref(t))
    return bool(res), t.value


# The eight checks of tests/intersection/aabb.cpp:10-106 (commented out in the
# reference): ray origin (1,1,-4) + 1e-5, direction (-1,-1,4), unless stated.
OFF = 1e-5
KAT = [
    ("zero-volume box", (1, 1, -4), (-1, -1, 4), (0, 0, 0), (0, 0, 0), False),          # :30-36
    ("flat segment box", (1, 1, -4), (-1, -1, 4), (0, 0, 0), (1, 0, 0), False),         # :38-42
    ("plane box", (1, 1, -4), (-1, -1, 4), (-1, 0, -1), (1, 0, 1), True),               # :44-51
    ("average box", (1, 1, -4), (-1, -1, 4), (-1, 0, -1), (1, 2, 1), True),             # :53-59
    ("tiny box", (1, 1, -4), (-1, -1, 4), (-0.003, 0, -0.003), (0.003, 0.003, 0.003), True),  # :61-68
    ("angle", (1, 1, -4), (-1, -1, 4), (0, 0, 0), (1, 0, 1), True),                     # :70-77
    ("parallel miss", (1, -1, -4), (-1, -1, 4), (-1, 0, -1), (1, 0, 1), False),         # :79-91
    ("grazing", (2.03, -0.03, -2.03), (-2.03, -0.03, 2.03), (-1, 0, -1), (1, 0, 1), False),  # :93-106
]


@pytest.mark.parametrize("name,o,d,pmin,pmax,expected", KAT, ids=[k[0] for k in KAT])
def test_reference_box_known_answers(name, o, d, pmin, pmax, expected):
    o32 = [np.float32(np.float32(x) + np.float32(OFF)) for x in o]
    res, _ = box_test(o32, d, pmin, pmax)
    assert res == expected


def test_box_infinite_inverse_direction():
    # axis-parallel ray: 1/d = +-inf must be handled by the slab test
    res, t = box_test((0.5, 0.5, -3.0), (0.0, 0.0, 1.0), (0, 0, -1), (1, 1, 1))
    assert res and t == pytest.approx(2.0)
    res, _ = box_test((1.5, 0.5, -3.0), (0.0, 0.0, 1.0), (0, 0, -1), (1, 1, 1))
    assert not res
    # origin inside the box: t0 clamps to 0
    res, t = box_test((0.5, 0.5, 0.5), (0.3, -0.2, 0.9), (0, 0, 0), (1, 1, 1))
    assert res and t == 0.0


def tri(o, d, v0, v1, v2):
    lib = load()
    f = lambda x: np.array(x, dtype=np.float32)
    v0, v1, v2 = f(v0), f(v1), f(v2)
    t = np.zeros(1, np.float32)
    b = np.zeros(2, np.float32)
    hit = lib.orc_ray_triangle(fptr(f(o)), fptr(f(d)), fptr(v0), fptr(v1 - v0), fptr(v2 - v0), fptr(t), fptr(b))
    return bool(hit), float(t[0]), b.astype(np.float64)


def mt64(o, d, v0, v1, v2):
    o, d, v0, v1, v2 = (np.asarray(x, np.float64) for x in (o, d, v0, v1, v2))
    e0, e1 = v1 - v0, v2 - v0
    p = np.cross(d, e1)
    det = e0 @ p
    tv = o - v0
    u = (tv @ p) / det
    q = np.cross(tv, e0)
    v = (d @ q) / det
    t = (e1 @ q) / det
    return t, u, v


def test_triangle_against_fp64():
    rng = np.random.default_rng(1)
    n_hit = 0
    for _ in range(4000):
        v0, v1, v2 = (rng.uniform(-1, 1, 3).astype(np.float32) for _ in range(3))
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        target = (v0 + rng.uniform(0, 1) * (v1 - v0) + rng.uniform(0, 1) * (v2 - v0)).astype(np.float32)
        d = (target - o).astype(np.float64)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        hit, t, b = tri(o, d, v0, v1, v2)
        t64, u64, w64 = mt64(o, d, v0, v1, v2)
        det = np.float64(np.dot(v1 - v0, np.cross(d, v2 - v0)))
        inside = u64 >= 0 and w64 >= 0 and u64 + w64 <= 1 and t64 >= 1e-6 and abs(det) >= 1e-6
        margin = min(abs(u64), abs(w64), abs(1 - u64 - w64), abs(abs(det) - 1e-6))
        if margin < 1e-4:
            continue  # within rounding of a rejection threshold
        assert hit == inside
        # grazing rays are ill-conditioned in FP32: compare values only when the
        # normalised determinant is not small
        cond = abs(det) / (np.linalg.norm(v1 - v0) * np.linalg.norm(v2 - v0))
        if hit and cond > 0.05:
            n_hit += 1
            assert t == pytest.approx(t64, rel=1e-4, abs=1e-5)
            assert b[0] == pytest.approx(u64, abs=1e-4)
            assert b[1] == pytest.approx(w64, abs=1e-4)
    assert n_hit > 1000


def test_triangle_rejections():
    v = ((0, 0, 0), (1, 0, 0), (0, 1, 0))
    assert tri((0.25, 0.25, 1), (0, 0, -1), *v)[0]                     # front
    assert tri((0.25, 0.25, -1), (0, 0, 1), *v)[0]                     # back face: no culling
    assert not tri((0.25, 0.25, 1), (1, 0, 0), *v)[0]                  # parallel: |det| < 1e-6
    assert not tri((0.25, 0.25, 1), (0, 0, 1), *v)[0]                  # behind origin: t < 1e-6
    assert not tri((0.75, 0.75, 1), (0, 0, -1), *v)[0]                 # u + v > 1
    assert not tri((-0.1, 0.2, 1), (0, 0, -1), *v)[0]                  # u < 0
    assert not tri((0.2, 0.2, 5e-7), (0, 0, -1), *v)[0]                # t = 5e-7 < 1e-6
    hit, t, b = tri((0.25, 0.5, 2), (0, 0, -1), *v)
    assert hit and t == 2.0 and b[0] == 0.25 and b[1] == 0.5


def test_rng_known_answers():
    lib = load()
    # Bob Jenkins' one-at-a-time hash of the one-byte key "a" is 0xca2e9442;
    # jenkinsHash (shaders/random.h:14-22) is that hash's per-byte mixing plus
    # its final avalanche, so jenkinsHash(0x61) must reproduce it.
    assert lib.orc_jenkins(0x61) == 0xCA2E9442
    # Marsaglia (2003), "Xorshift RNGs": the 13/17/5 generator from
    # y = 2463534242 first returns 723471715.
    s = C.c_uint32(2463534242)
    assert lib.orc_xorshift(C.byref(s)) == 723471715
    # uintToFloat: asfloat(0x3f800000 | x >> 9) - 1 == (x >> 9) * 2^-23 exactly
    s = C.c_uint32(2463534242)
    f = lib.orc_rand(C.byref(s))
    assert f == (723471715 >> 9) * 2.0 ** -23
    # initRNG = jenkins((x + y*W) ^ jenkins(frame))
    assert lib.orc_init_rng(3, 2, 10, 5) == lib.orc_jenkins(23 ^ lib.orc_jenkins(5))
    # rand() stays in [0, 1)
    s = C.c_uint32(12345)
    vals = [lib.orc_rand(C.byref(s)) for _ in range(20000)]
    assert min(vals) >= 0.0 and max(vals) < 1.0
    assert abs(np.mean(vals) - 0.5) < 0.01


def test_permute_is_a_permutation():
    lib = load()
    for p in (0, 7, 0x51633E2D * 16 & 0xFFFFFFFF):
        out = sorted(lib.orc_permute(i, 16, p) for i in range(16))
        assert out == list(range(16))


def test_camera_closed_form():
    import strelka_amd as sa
    s = sa.SceneData.generate(1)
    cam = s.camera
    lib = load()
    # centre-of-pixel sampling (stratifiedSamplingType 0): the ray through the
    # image centre follows the view axis (-z from (0,0,1.9) towards the origin)
    cam.stratifiedSamplingType = 0
    W = cam.dimension[0]
    ray = np.zeros(1, dtype=RAY)
    rng = C.c_uint32(1)
    lib.orc_camera_ray(C.byref(cam), W // 2, W // 2, C.byref(rng), 0, C.c_void_p(ray.ctypes.data))
    o, d = ray["o"][0], ray["d"][0]
    assert np.allclose(o[:3], [0, 0, 1.9], atol=1e-6) and o[3] == np.float32(1e9)
    # pixel (128.5,128.5) of 256: ndc = 1/256; direction ~ (ndc*tan30, -ndc*tan30, -1)
    ndc = (128.5 / 256) * 2 - 1
    expect = np.array([ndc * math.tan(math.radians(30)), -ndc * math.tan(math.radians(30)), -1.0])
    expect /= np.linalg.norm(expect)
    assert np.allclose(d[:3], expect, atol=1e-6)
    # top-left pixel: 30 degrees half-angle on both axes (square image)
    lib.orc_camera_ray(C.byref(cam), 0, 0, C.byref(rng), 0, C.c_void_p(ray.ctypes.data))
    d = ray["d"][0][:3]
    ndc = (0.5 / 256) * 2 - 1
    expect = np.array([ndc * math.tan(math.radians(30)), -ndc * math.tan(math.radians(30)), -1.0])
    assert np.allclose(d, expect / np.linalg.norm(expect), atol=1e-6)
