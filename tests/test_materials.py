"""Closed-form pins of the oracle's specular, emissive and normal-mapped
materials and of the mirrored-repeat and clip texture wraps (oracle/oracle.c
standing in for the MDL scattering and emission calls of
shaders/pathtracerMdl.hlsl:299-372, CalcBumpedNormal of
shaders/helper.h:101-115 and apply_wrap_and_crop of
shaders/mdl_runtime.hlsl:93-120).  Expected values come from Fresnel's
equations, Snell's law, the law of reflection, Schlick's approximation and
the geometry of a convex object under a constant sky, in float64."""
import ctypes as C

import numpy as np
import pytest

import strelka_amd as sa
from material_scenes import furnace, material, showcase
from oracle import Oracle, load
from oracle.pyoracle import bumped_normal, fptr, texture_lookup


class BsdfSample(C.Structure):
    _fields_ = [("k2", C.c_float * 3), ("pdf", C.c_float), ("bsdf_over_pdf", C.c_float * 3), ("event", C.c_int)]


def sample(m, N, Ng, k1, xi):
    out = BsdfSample()
    N, Ng, k1 = (np.asarray(v, np.float32) for v in (N, Ng, k1))
    load().orc_bsdf_sample(C.c_void_p(m.ctypes.data), fptr(N), fptr(Ng), fptr(k1), fptr(np.asarray(xi, np.float32)),
                           C.byref(out))
    return np.array(out.k2[:], np.float64), np.array(out.bsdf_over_pdf[:], np.float64), out.event, out.pdf


def direction(theta, phi=0.3):
    """A unit vector at polar angle theta from +z."""
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])


def fresnel(c1, eta):
    """Unpolarised reflectance of a smooth dielectric for the cosine c1 on
    the incident side and eta = n1 / n2 (Fresnel's equations)."""
    s2 = eta * eta * (1.0 - c1 * c1)
    if s2 >= 1.0:
        return 1.0, None
    c2 = np.sqrt(1.0 - s2)
    rs = (eta * c1 - c2) / (eta * c1 + c2)
    rp = (c1 - eta * c2) / (c1 + eta * c2)
    return 0.5 * (rs * rs + rp * rp), c2


Z = np.array([0.0, 0.0, 1.0])


@pytest.mark.parametrize("ior", [1.5, 1.33, 2.4])
@pytest.mark.parametrize("deg", [0.0, 10.0, 35.0, 56.31, 70.0, 85.0])
@pytest.mark.parametrize("entering", [True, False])
def test_glass_chooses_reflection_with_the_fresnel_reflectance(ior, deg, entering):
    """The reflect/refract choice switches at xi = F, the reflectance of
    Fresnel's equations (the Brewster angle of n = 1.5, 56.31 deg, included;
    beyond the critical angle of an exiting ray every sample reflects); the
    weights are 1 for reflection and the tint for refraction, and the
    reflected direction is the mirror image of k1."""
    m = material(sa.MAT_GLASS, albedo=(0.9, 0.8, 0.7), ior=ior)
    th = np.radians(deg)
    k1 = direction(th) if entering else -direction(th)
    N = Z
    eta = 1.0 / ior if entering else ior
    F, c2 = fresnel(np.cos(th), eta)
    xs = np.linspace(0.0, 1.0, 4001)[:-1] + 1.0 / 8000
    refl = []
    for x in xs:
        k2, w, ev, pdf = sample(m, N, N, k1, (0.5, 0.5, x, 0.5))
        assert pdf == 0.0 and ev in (3, 4)
        n = N if entering else -N
        if ev == 3:
            refl.append(True)
            assert np.allclose(w, 1.0)
            assert np.allclose(k2, 2 * np.dot(n, k1) * n - k1, atol=2e-7)
        else:
            refl.append(False)
            assert np.allclose(w, (0.9, 0.8, 0.7), atol=1e-7)
    measured = np.mean(refl)
    assert abs(measured - F) <= 1.0 / 4000 + 1e-6, (measured, F)
    if c2 is None:
        assert measured == 1.0


@pytest.mark.parametrize("ior", [1.5, 1.33])
@pytest.mark.parametrize("deg", [0.0, 20.0, 45.0, 80.0])
@pytest.mark.parametrize("entering", [True, False])
def test_refraction_follows_snells_law(ior, deg, entering):
    """The refracted ray: unit length, on the far side, in the plane of
    incidence, with n1 sin(theta1) = n2 sin(theta2)."""
    m = material(sa.MAT_GLASS, ior=ior)
    th = np.radians(deg)
    k1 = direction(th, 1.1) if entering else -direction(th, 1.1)
    eta = 1.0 / ior if entering else ior
    if eta * np.sin(th) >= 1.0:
        pytest.skip("total internal reflection")
    k2, _, ev, _ = sample(m, Z, Z, k1, (0.5, 0.5, 0.999999, 0.5))
    assert ev == 4
    n = Z if entering else -Z
    assert abs(np.linalg.norm(k2) - 1.0) < 1e-6
    assert np.dot(k2, n) < 0.0
    assert abs(np.dot(k2, np.cross(n, k1))) < 1e-6
    sin2 = np.linalg.norm(np.cross(-n, k2))
    assert abs(sin2 - eta * np.sin(th)) < 1e-6
    # the exact transmitted direction
    c1 = np.dot(n, k1)
    c2 = np.sqrt(1.0 - eta * eta * (1.0 - c1 * c1))
    assert np.allclose(k2, -eta * k1 + (eta * c1 - c2) * n, atol=1e-6)


def test_glass_side_follows_the_geometric_normal():
    """With the shading normal tilted off the geometric one, the side is the
    geometric normal's: a ray that meets the back face is exiting, and a ray
    whose shading cosine is not positive on its side is absorbed."""
    m = material(sa.MAT_GLASS, ior=1.5)
    Ng = Z
    N = np.array([0.0, np.sin(0.2), np.cos(0.2)])
    k1 = direction(np.radians(30), 0.0)
    _, _, ev, _ = sample(m, N, Ng, k1, (0.5, 0.5, 0.99, 0.5))
    assert ev == 4  # entering, refracted
    _, _, ev, _ = sample(m, N, Ng, -k1, (0.5, 0.5, 0.99, 0.5))
    assert ev in (3, 4)  # exiting
    grazing = np.array([0.0, -np.cos(0.1), np.sin(0.1)])  # above Ng, below N's horizon
    assert np.dot(grazing, Ng) > 0 and np.dot(grazing, N) < 0
    _, _, ev, _ = sample(m, N, Ng, grazing, (0.5, 0.5, 0.99, 0.5))
    assert ev == 0


@pytest.mark.parametrize("deg", [0.0, 30.0, 60.0, 85.0])
def test_mirror_reflects_with_schlicks_weight(deg):
    """The mirror: the law of reflection about the shading normal, weight
    f0 + (1 - f0)(1 - cos)^5 with f0 = albedo, no pdf, and no evaluation
    off the delta; a ray from behind is absorbed."""
    f0 = np.array([0.9, 0.5, 0.1])
    m = material(sa.MAT_MIRROR, albedo=f0)
    k1 = direction(np.radians(deg), 2.0)
    k2, w, ev, pdf = sample(m, Z, Z, k1, (0.3, 0.7, 0.1, 0.9))
    assert ev == 3 and pdf == 0.0
    assert np.allclose(k2, 2 * np.dot(Z, k1) * Z - k1, atol=2e-7)
    c = np.cos(np.radians(deg))
    assert np.allclose(w, f0 + (1 - f0) * (1 - c) ** 5, atol=2e-6)
    _, _, ev, _ = sample(m, Z, Z, -k1, (0.3, 0.7, 0.1, 0.9))
    assert ev == 0
    for kind in (sa.MAT_MIRROR, sa.MAT_GLASS):
        d, g, pdf = (np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(1, np.float32))
        k1f, k2f, Zf = (np.asarray(v, np.float32) for v in (k1, k2, Z))
        load().orc_bsdf_eval(C.c_void_p(material(kind).ctypes.data), fptr(Zf), fptr(Zf), fptr(k1f), fptr(k2f),
                             fptr(d), fptr(g), fptr(pdf))
        assert not d.any() and not g.any() and pdf[0] == 0.0


def render(s, ref_compat, spp=4, max_depth=32, rr_start=64, frame=1):
    o = Oracle.from_scene(s, sa.HostBvh.build(s).nodes)
    p = sa.render_params(s.camera, spp=spp, max_depth=max_depth, frame=frame, sampling=1, sky=s.sky,
                         rr_start=rr_start, ref_compat=ref_compat)
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    return o.pt_frame(p, 0, 0, w, h).reshape(-1, 3)


@pytest.mark.parametrize("ref_compat", [True, False])
def test_white_mirror_and_clear_glass_furnace_return_the_sky_exactly(ref_compat):
    """A white mirror cube and a clear glass cube under a constant sky: a
    convex mirror sends every camera ray straight to the sky with weight
    Schlick(1, cos) = 1; inside the glass cube a ray reflects (weight 1) and
    refracts (weight 1) until it leaves, so every sample is exactly the sky
    (or 0 for the few absorbed at a shading normal's horizon or held past the
    depth limit)."""
    for mat in (material(sa.MAT_MIRROR), material(sa.MAT_GLASS, ior=1.5)):
        s = furnace(mat)
        smp = render(s, ref_compat)
        sky = np.array(s.sky[:3], np.float32)
        is_sky = np.all(smp == sky, axis=1)
        zero = np.all(smp == 0, axis=1)
        print(f"type {int(mat['type'][0])}: sky {is_sky.mean():.4f}, zero {zero.mean():.5f}")
        assert np.all(is_sky | zero)
        assert zero.mean() < 2e-3


def test_tinted_glass_tints_each_transmission():
    """Tinted glass cube: a path that refracts in and out carries tint^2 (two
    transmissions, reflections untinted); one that reflects off the first
    face carries 1, so every sample is sky, tint*tint*sky or 0."""
    t = np.array([0.9, 0.6, 0.3], np.float32)
    s = furnace(material(sa.MAT_GLASS, albedo=t, ior=1.5))
    smp = render(s, False)
    sky = np.array(s.sky[:3], np.float32)
    through = (t * t) * sky
    a = np.all(smp == sky, axis=1)
    b = np.all(smp == through, axis=1)
    z = np.all(smp == 0, axis=1)
    print(f"sky {a.mean():.4f} tint^2 {b.mean():.4f} zero {z.mean():.5f}")
    assert np.all(a | b | z) and b.mean() > 0.2 and z.mean() < 2e-3


@pytest.mark.parametrize("ref_compat", [True, False])
def test_emitting_cube_adds_its_emission_once(ref_compat):
    """An emitting Lambert cube (albedo a, emission E) under a constant sky S:
    a camera ray on a front face gathers E, then a * S after its one bounce
    (E alone if the bounce is absorbed); a miss gathers S."""
    a = np.array([0.5, 0.25, 0.75], np.float32)
    E = np.array([2.0, 3.0, 0.5], np.float32)
    s = furnace(material(sa.MAT_LAMBERT, albedo=a, emission=E))
    smp = render(s, ref_compat, max_depth=8, rr_start=3)
    S = np.array(s.sky[:3], np.float32)
    hit = E + a * S
    kinds = [np.all(smp == v, axis=1) for v in (S, hit, E)]
    print(f"sky {kinds[0].mean():.3f} hit {kinds[1].mean():.3f} absorbed {kinds[2].mean():.5f}")
    assert np.all(kinds[0] | kinds[1] | kinds[2]) and kinds[1].mean() > 0.2 and kinds[2].mean() < 1e-3


def test_emission_of_a_black_cube_under_a_black_sky():
    """Black cube, black sky, emission E: hits are exactly E, misses 0, and
    the back faces (never seen from outside) emit nothing."""
    E = np.array([1.5, 0.5, 4.0], np.float32)
    s = furnace(material(sa.MAT_LAMBERT, albedo=(0, 0, 0), emission=E), sky=(0, 0, 0, 0))
    smp = render(s, True, max_depth=4, rr_start=3)
    hit = np.all(smp == E, axis=1)
    miss = np.all(smp == 0, axis=1)
    assert np.all(hit | miss) and hit.mean() > 0.2


def frame_of(n, t):
    n = np.asarray(n, np.float64) / np.linalg.norm(n)
    t = -np.asarray(t, np.float64)
    t = t - n * np.dot(t, n)
    t /= np.linalg.norm(t)
    return n, t, np.cross(n, t)


@pytest.mark.parametrize("alpha", [0.0, 0.3, 1.0])
def test_bumped_normal_rotates_in_the_tangent_frame(alpha):
    """CalcBumpedNormal: a texel (sin a, 0, cos a) encoded as (n + 1) / 2
    tilts the normal by a towards -T (the negated tangent, made orthogonal to
    N), (0, sin a, cos a) towards B = N x (-T); the flat texel (0.5, 0.5, 1)
    returns N; a tangent with a component along N gives the same frame."""
    N = np.array([0.2, 0.9, -0.1])
    T = np.array([1.0, 0.1, 0.3])
    n, t, b = frame_of(N, T)
    for comp, axis in ((0, t), (1, b)):
        m = np.zeros(3)
        m[comp] = np.sin(alpha)
        m[2] = np.cos(alpha)
        got = bumped_normal(N / np.linalg.norm(N), T, (m + 1.0) / 2.0).astype(np.float64)
        assert abs(np.linalg.norm(got) - 1.0) < 1e-6
        assert abs(np.dot(got, n) - np.cos(alpha)) < 2e-6
        assert abs(np.dot(got, axis) - np.sin(alpha)) < 2e-6
        assert abs(np.dot(got, np.cross(n, axis))) < 2e-6
    flat = bumped_normal(N / np.linalg.norm(N), T, (0.5, 0.5, 1.0)).astype(np.float64)
    assert np.allclose(flat, n, atol=2e-7)
    skew = bumped_normal(N / np.linalg.norm(N), T + 0.7 * N, (0.8, 0.3, 0.9)).astype(np.float64)
    assert np.allclose(skew, bumped_normal(N / np.linalg.norm(N), T, (0.8, 0.3, 0.9)), atol=2e-6)


def test_bumped_normal_without_a_frame_is_nan():
    """A tangent along the normal leaves no frame (the renderers then keep
    the interpolated normal)."""
    got = bumped_normal((0.0, 1.0, 0.0), (0.0, 2.0, 0.0), (0.7, 0.5, 0.9))
    assert np.isnan(got).any()


def tex(wrap, texels, w, h):
    t = np.zeros(1, sa.TEXTURE)
    t[0] = (w, h, 0, wrap)
    return t[0], texels


def test_mirrored_repeat_and_clip_wraps():
    """Mirrored repeat equals clamp on [0, 1] exactly, is even (f(-s) = f(s))
    and has period 2; clip equals clamp inside [0, 1) and is zero outside it
    on either axis."""
    rng = np.random.default_rng(5)
    texels = rng.random((5 * 3, 4)).astype(np.float32)
    clamp = tex(sa.TEX_CLAMP, texels, 5, 3)
    mirror = tex(sa.TEX_MIRRORED_REPEAT, texels, 5, 3)
    clip = tex(sa.TEX_CLIP, texels, 5, 3)
    for s, t in rng.random((300, 2)):
        s, t = float(np.float32(s)), float(np.float32(t))
        c = texture_lookup(*clamp, s, t)
        assert np.array_equal(texture_lookup(*mirror, s, t), c)
        assert np.array_equal(texture_lookup(*clip, s, t), c)
        assert np.allclose(texture_lookup(*mirror, -s, t), c, atol=2e-6)
        assert np.allclose(texture_lookup(*mirror, s, -t), c, atol=2e-6)
        assert np.allclose(texture_lookup(*mirror, s + 2.0, t - 4.0), c, atol=4e-6)
        assert np.allclose(texture_lookup(*mirror, 2.0 - s, t), c, atol=4e-6)
        for so, to in ((s - 1.0, t), (s + 1.0, t), (s, t - 1.0), (s, t + 1.0)):
            assert not texture_lookup(*clip, so, to).any()
    assert np.array_equal(texture_lookup(*mirror, 1.0, 0.5), texture_lookup(*clamp, 1.0, 0.5))
    assert not texture_lookup(*clip, 1.0, 0.5).any()


@pytest.mark.parametrize("ref_compat", [True, False])
def test_showcase_scene_renders_every_material(ref_compat):
    """The showcase scene (every material kind and wrap) renders finite,
    non-negative samples, and each specular or emissive sphere is visible:
    the oracle's frame changes when any one sphere's material goes back to
    Lambert."""
    s = showcase()
    smp = render(s, ref_compat, spp=1, max_depth=8, rr_start=3)
    assert np.all(np.isfinite(smp)) and np.all(smp >= 0)
    for i in (1, 2, 3, 4, 7):
        t = showcase()
        mats = t.mats.copy()
        mats[i] = material(sa.MAT_LAMBERT, (0.5, 0.5, 0.5))[0]
        t.mats = mats
        assert not np.array_equal(render(t, ref_compat, spp=1, max_depth=8, rr_start=3), smp), i


@pytest.mark.parametrize("ref_compat", [True, False])
def test_oracle_converges_to_the_binary64_twin_on_the_specular_scene(ref_compat):
    """The binary32 oracle (1,024 spp) against the binary64 twin (512 spp) on
    the untextured showcase scene: per-pixel means within their standard
    errors (mean z^2 in [0.7, 1.4], no |z| > 4.5).  The twin restates the
    mirror, the dielectric and the emission in binary64 from the same
    formulas, so a dropped weight or a wrong side shows as a bias."""
    s = showcase(textured=False)
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    o = Oracle.from_scene(s, sa.HostBvh.build(s).nodes)
    spp, per = 1024, 256
    s1 = np.zeros((h, w, 3))
    s2 = np.zeros((h, w, 3))
    for f in range(spp // per):
        p = sa.render_params(s.camera, spp=per, max_depth=8, frame=f, sampling=1, sky=s.sky, ref_compat=ref_compat)
        smp = o.pt_frame(p, 0, 0, w, h).reshape(h, w, per, 3).astype(np.float64)
        s1 += smp.sum(2)
        s2 += (smp * smp).sum(2)
    m32 = s1 / spp
    v32 = np.maximum(s2 / spp - m32 * m32, 0) / (spp - 1)
    p = sa.render_params(s.camera, spp=512, max_depth=8, frame=100000, sampling=1, sky=s.sky, ref_compat=ref_compat)
    m64, v64 = o.pt_frame_fp64(p, 0, 0, w, h)
    var = v32 + v64
    ok = var > 0
    z = (m32 - m64)[ok] / np.sqrt(var[ok])
    print(f"mean z^2 {np.mean(z * z):.3f}, max |z| {np.abs(z).max():.2f}")
    assert 0.7 < np.mean(z * z) < 1.4 and np.abs(z).max() < 4.5


def test_binary64_twin_refuses_textured_materials():
    s = showcase()
    o = Oracle.from_scene(s, sa.HostBvh.build(s).nodes)
    with pytest.raises(ValueError):
        o.pt_frame_fp64(sa.render_params(s.camera, spp=1, sky=s.sky), 0, 0, 4, 4)
