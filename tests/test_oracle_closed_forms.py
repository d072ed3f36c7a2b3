"""Closed-form pins of the oracle's rect-light sampling, BSDFs and path loop
(oracle/oracle.c restating shaders/lights.h, shaders/pathtracer.hlsl,
shaders/helper.h and shaders/pathtracerMdl.hlsl).  Each expected value comes
from mathematics (tests/closed_forms.py), so a dropped cosine, a wrong pdf
measure, a transposed frame or a mis-mapped sample fails here."""
import ctypes as C

import numpy as np
import pytest

import strelka_amd as sa
from closed_forms import floor_expectations, gauss_rect, projected_solid_angle, solid_angle_quad
from oracle import Oracle, load
from oracle.pyoracle import fptr


class BsdfSample(C.Structure):
    _fields_ = [("k2", C.c_float * 3), ("pdf", C.c_float), ("bsdf_over_pdf", C.c_float * 3), ("event", C.c_int)]


def f32(*v):
    return np.array(v, np.float32)


def rect_light(corners, color=10.0):
    l = np.zeros(1, sa.LIGHT)
    l["points"][0, :, :3] = corners
    l["points"][0, :, 3] = 1.0
    l["color"][0] = color
    l["type"] = 0
    return l


LIGHT_CORNERS = np.array([[0.5, 2.0, -0.5], [-0.5, 2.0, -0.5], [-0.5, 2.0, 0.5], [0.5, 2.0, 0.5]], np.float32)
RECEIVERS = [(0.0, 0.0, 0.0), (0.7, 0.0, -0.3), (-2.5, 0.4, 1.5), (3.0, -1.0, 3.0)]


def sample_light(lib, light, P, u):
    pt, L, nrm = (np.zeros(3, np.float32) for _ in range(3))
    dist, area, pdf, S = (np.zeros(1, np.float32) for _ in range(4))
    lib.orc_sample_rect_light(C.c_void_p(light.ctypes.data), fptr(np.asarray(u, np.float32)), fptr(P), fptr(pt),
                              fptr(L), fptr(dist), fptr(nrm), fptr(area), fptr(pdf), fptr(S))
    return pt, L, float(dist[0]), nrm, float(area[0]), float(pdf[0]), float(S[0])


@pytest.mark.parametrize("P", RECEIVERS)
def test_rect_light_solid_angle_and_sample_distribution(P):
    lib = load()
    light = rect_light(LIGHT_CORNERS)
    P = f32(*P)
    S_ref = solid_angle_quad(P, LIGHT_CORNERS)
    n = 128
    grid = (np.arange(n) + 0.5) / n
    up = np.array([0.0, 1.0, 0.0])
    cos_sum, inv_pdf_cos_sum = 0.0, 0.0
    for u0 in grid:
        for u1 in grid:
            pt, L, dist, nl, area, pdf, S = sample_light(lib, light, P, (u0, u1))
            # the sample lies on the light
            assert abs(pt[1] - 2.0) < 1e-5 and abs(pt[0]) <= 0.5 + 1e-5 and abs(pt[2]) <= 0.5 + 1e-5
            assert np.allclose(L, (pt - P) / np.linalg.norm(pt - P), atol=1e-6)
            cos_sum += float(np.dot(L, up))
            inv_pdf_cos_sum += float(np.dot(L, up)) / pdf
    assert S == pytest.approx(S_ref, rel=2e-5)
    assert area == pytest.approx(1.0, rel=1e-6)
    assert np.allclose(nl, (0.0, -1.0, 0.0), atol=1e-6)
    # uniform in solid angle: E[cos] = projected solid angle / S
    assert cos_sum / n**2 == pytest.approx(projected_solid_angle(P, up, LIGHT_CORNERS) / S_ref, rel=2e-4)
    # the pdf is the area-measure density r^2 / (cos_l A) (quirk: not 1/S):
    # E[cos_P / pdf] = (A / S) * integral of cos_P cos_l^2 / r^4 dA
    pts, ws, A, _ = gauss_rect(LIGHT_CORNERS, 24)
    d = pts - P.astype(np.float64)
    r2 = np.sum(d * d, axis=1)
    cos_p = d[:, 1] / np.sqrt(r2)
    cos_l = d[:, 1] / np.sqrt(r2)
    expect = A / S_ref * np.sum(cos_p * cos_l**2 / r2**2 * ws)
    assert inv_pdf_cos_sum / n**2 == pytest.approx(expect, rel=5e-4)


def material(kind, albedo=(1.0, 1.0, 1.0), roughness=1.0):
    m = np.zeros(1, sa.MATERIAL)
    m["type"] = kind
    m["roughness"] = roughness
    m["albedo"][0, :3] = albedo
    m["albedo"][0, 3] = 1.0
    return m


def frame(n):
    n = np.asarray(n, np.float64)
    n = n / np.linalg.norm(n)
    t = np.cross(n, [1.0, 0.0, 0.0] if abs(n[0]) < 0.9 else [0.0, 0.0, 1.0])
    t /= np.linalg.norm(t)
    return n, t, np.cross(n, t)


def bsdf_sample(lib, m, N, k1, xi):
    out = BsdfSample()
    Nf = N.astype(np.float32)
    lib.orc_bsdf_sample(C.c_void_p(m.ctypes.data), fptr(Nf), fptr(Nf), fptr(k1.astype(np.float32)),
                        fptr(np.asarray(xi, np.float32)), C.byref(out))
    return out


def bsdf_eval(lib, m, N, k1, k2):
    d, g, pdf = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(1, np.float32)
    Nf = N.astype(np.float32)
    lib.orc_bsdf_eval(C.c_void_p(m.ctypes.data), fptr(Nf), fptr(Nf), fptr(k1.astype(np.float32)),
                      fptr(k2.astype(np.float32)), fptr(d), fptr(g), fptr(pdf))
    return d, g, float(pdf[0])


def hemisphere_quadrature(N, nt, nphi):
    """Gauss-Legendre in theta over [0, pi/2], midpoint in phi: directions and
    solid-angle weights."""
    n, t, b = frame(N)
    x, w = np.polynomial.legendre.leggauss(nt)
    th = 0.25 * np.pi * (x + 1.0)
    wt = 0.25 * np.pi * w * np.sin(th)
    ph = (np.arange(nphi) + 0.5) * 2.0 * np.pi / nphi
    dirs, ws = [], []
    for a, wa in zip(th, wt):
        for p in ph:
            dirs.append(np.sin(a) * np.cos(p) * t + np.sin(a) * np.sin(p) * b + np.cos(a) * n)
            ws.append(wa * 2.0 * np.pi / nphi)
    return np.array(dirs), np.array(ws)


def test_lambert_cosine_sampling_moments_and_weight():
    lib = load()
    m = material(sa.MAT_LAMBERT, (0.7, 0.5, 0.3))
    N, T, _ = frame([0.3, 0.8, -0.5])
    n = 160
    c1 = c2 = ct = 0.0
    for a in (np.arange(n) + 0.5) / n:
        for b in (np.arange(n) + 0.5) / n:
            s = bsdf_sample(lib, m, N, N, (a, b, 0.0, 0.0))
            assert s.event == 1
            assert np.array_equal(np.array(s.bsdf_over_pdf, np.float32), f32(0.7, 0.5, 0.3))  # f cos / pdf = albedo
            k2 = np.array(s.k2, np.float64)
            c = float(np.dot(k2, N))
            assert abs(np.linalg.norm(k2) - 1.0) < 1e-5 and c > 0
            assert s.pdf == pytest.approx((c + 1e-6) / np.pi, rel=1e-4)
            c1 += c
            c2 += c * c
            ct += float(np.dot(k2, T))
    # cosine-weighted hemisphere: E[cos] = 2/3, E[cos^2] = 1/2, tangential mean 0
    assert c1 / n**2 == pytest.approx(2.0 / 3.0, abs=2e-4)
    assert c2 / n**2 == pytest.approx(0.5, abs=2e-4)
    assert abs(ct / n**2) < 1e-3


def test_lambert_white_furnace():
    lib = load()
    m = material(sa.MAT_LAMBERT, (1.0, 0.5, 0.25))
    N, _, _ = frame([-0.2, 0.1, 0.97])
    dirs, ws = hemisphere_quadrature(N, 48, 32)
    tot, ptot = np.zeros(3), 0.0
    for k2, w in zip(dirs, ws):
        d, g, pdf = bsdf_eval(lib, m, N, N, k2)
        tot += d * w
        ptot += pdf * w
        assert np.all(g == 0)
    # integral of f cos over the hemisphere = albedo; the pdf integrates to 1
    assert tot == pytest.approx([1.0, 0.5, 0.25], rel=1e-4)
    assert ptot == pytest.approx(1.0, rel=1e-4)


@pytest.mark.parametrize("roughness", [0.3, 0.55, 0.8])
def test_ggx_pdf_normalisation_and_sampled_cdf(roughness):
    lib = load()
    alpha = max(roughness * roughness, 1e-4)
    m = material(sa.MAT_GGX, (1.0, 1.0, 1.0), roughness)
    N, _, _ = frame([0.0, 0.0, 1.0])
    # at normal incidence k2 = reflect(N, h) leaves the hemisphere iff theta_h >= 45 deg,
    # so the reflected pdf integrates to the GGX cdf at 45 deg: 1 / (1 + alpha^2)
    dirs, ws = hemisphere_quadrature(N, 800, 4)
    tot = sum(bsdf_eval(lib, m, N, N, k2)[2] * w for k2, w in zip(dirs, ws))
    assert tot == pytest.approx(1.0 / (1.0 + alpha * alpha), rel=2e-3)
    # inverse-cdf sampling of D cos(theta_h): P(theta_h < t) = tan^2 t / (alpha^2 + tan^2 t)
    n = 400
    th = []
    for a in (np.arange(n) + 0.5) / n:
        s = bsdf_sample(lib, m, N, N, (a, 0.37, 0.0, 0.0))
        th.append(0.5 * np.arccos(np.clip(s.k2[2], -1, 1)) if s.event == 2 else np.pi / 4 + 1e-3)
    th = np.array(th)
    for t in np.radians([10.0, 20.0, 30.0, 40.0]):
        cdf = np.tan(t) ** 2 / (alpha**2 + np.tan(t) ** 2)
        assert np.mean(th < t) == pytest.approx(cdf, abs=2.0 / n)


def test_ggx_sample_eval_consistency_and_reciprocity():
    lib = load()
    rng = np.random.default_rng(3)
    for roughness in (0.25, 0.6, 0.9):
        m = material(sa.MAT_GGX, (0.9, 0.6, 0.3), roughness)
        N, _, _ = frame([0.1, 0.2, 0.97])
        for _ in range(300):
            k1 = rng.normal(size=3)
            k1 /= np.linalg.norm(k1)
            if np.dot(k1, N) < 0.05:
                k1 = k1 - 2 * np.dot(k1, N) * N + 0.1 * N
                k1 /= np.linalg.norm(k1)
            s = bsdf_sample(lib, m, N, k1, rng.uniform(size=4))
            if s.event != 2:
                continue
            k2 = np.array(s.k2, np.float64)
            d, g, pdf = bsdf_eval(lib, m, N, k1, k2)
            assert pdf == pytest.approx(s.pdf, rel=2e-3)
            assert np.allclose(np.array(s.bsdf_over_pdf), g / pdf, rtol=5e-3)
            # f(k1,k2) cos(k2) / cos(k2) == f(k2,k1) cos(k1) / cos(k1)
            d2, g2, _ = bsdf_eval(lib, m, N, k2, k1)
            assert np.allclose(g * np.dot(N, k1), g2 * np.dot(N, k2), rtol=2e-3, atol=1e-7)


def scene_oracle(config):
    s = sa.SceneData.generate(config)
    return s, Oracle.from_scene(s, sa.HostBvh.build(s).nodes)


def test_convex_furnace_samples_are_exact():
    """Outward Lambert cube under a constant sky, no lights: a camera ray that
    hits the cube bounces once and escapes (convexity), so every sample is
    exactly albedo * sky; a miss is exactly sky."""
    s, o = scene_oracle(8)
    p = sa.render_params(s.camera, spp=4, max_depth=8, frame=1, sampling=1, sky=s.sky)
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    smp = o.pt_frame(p, 0, 0, w, h).reshape(-1, 3)
    sky = np.array(s.sky[:3], np.float32)
    hit_val = np.array(s.mats["albedo"][0, :3], np.float32) * sky
    is_sky = np.all(smp == sky, axis=1)
    is_hit = np.all(smp == hit_val, axis=1)
    assert is_sky.mean() > 0.2 and is_hit.mean() > 0.2
    # a few cosine samples may fall between the quantised shading normal's and
    # the face's horizons and are absorbed (value 0)
    absorbed = np.all(smp == 0, axis=1)
    assert np.all(is_sky | is_hit | absorbed)
    assert absorbed.mean() < 1e-3


def test_lit_floor_converges_to_quadrature():
    """Oracle pixel means (2048 spp) of the lit floor against the FP64
    quadrature of the estimator's expectation, within 5 standard errors."""
    s, o = scene_oracle(9)
    spp = 2048
    p = sa.render_params(s.camera, spp=spp, max_depth=8, frame=0, sampling=1, sky=(0, 0, 0))
    pixels = [(32, 32), (10, 50), (60, 3), (45, 20)]
    expect = floor_expectations(s, pixels)
    for (px, py), e in zip(pixels, expect):
        smp = o.pt_frame(p, px, py, 1, 1).reshape(-1, 3)[:, 0].astype(np.float64)
        mean, se = smp.mean(), smp.std() / np.sqrt(spp)
        print(f"pixel {px},{py}: oracle {mean:.6f} +- {se:.6f}, closed form {e:.6f}")
        assert abs(mean - e) < 5 * se + 1e-5 * e
        assert se < 0.01 * e  # the comparison is informative


def test_corrected_estimator_converges_to_the_reflected_radiance():
    """With refCompat = 0 the estimator is unbiased: on the lit floor each
    pixel's mean is the radiance the Lambert floor reflects from the rect
    light, albedo / pi * L * (Lambert's polygon formula for the light seen
    from the floor), for the binary32 oracle and for its binary64 twin, within
    5 standard errors.  The reference's estimator (refCompat = 1) does not
    converge there (its light pdf is in area measure and its next-event term
    carries cos^2): it lands at least 20 standard errors off."""
    s, o = scene_oracle(9)
    spp = 2048
    pixels = [(32, 32), (10, 50), (60, 3), (45, 20)]
    truth = floor_expectations(s, pixels, corrected=True)
    quirky = floor_expectations(s, pixels)
    fixed = sa.render_params(s.camera, spp=spp, max_depth=8, frame=0, sampling=1, sky=(0, 0, 0), ref_compat=False)
    ref = sa.render_params(s.camera, spp=spp, max_depth=8, frame=0, sampling=1, sky=(0, 0, 0))
    for (px, py), e, q in zip(pixels, truth, quirky):
        smp = o.pt_frame(fixed, px, py, 1, 1).reshape(-1, 3)[:, 0].astype(np.float64)
        mean, se = smp.mean(), smp.std() / np.sqrt(spp)
        m64, v64 = o.pt_frame_fp64(fixed, px, py, 1, 1)
        m64, se64 = float(m64[0, 0, 0]), float(np.sqrt(v64[0, 0, 0]))
        rsm = o.pt_frame(ref, px, py, 1, 1).reshape(-1, 3)[:, 0].astype(np.float64)
        print(f"pixel {px},{py}: corrected {mean:.6f} +- {se:.6f} (fp64 {m64:.6f} +- {se64:.6f}), "
              f"reflected radiance {e:.6f}; reference estimator {rsm.mean():.6f} (its expectation {q:.6f})")
        assert abs(mean - e) < 5 * se + 1e-5 * e
        assert abs(m64 - e) < 5 * se64 + 1e-5 * e
        assert se < 0.01 * e
        assert abs(rsm.mean() - e) > 20 * rsm.std() / np.sqrt(spp)
