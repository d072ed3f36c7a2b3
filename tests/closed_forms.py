"""FP64 closed forms used to pin the oracle and the GPU path tracer.

Everything here is mathematics written from the textbook formulas named in
each function, in float64 numpy; none of it calls the oracle or the product.
"""
from __future__ import annotations

import numpy as np


def solid_angle_triangle(a, b, c):
    """Solid angle of the triangle (a, b, c) seen from the origin
    (van Oosterom and Strackee, IEEE Trans. Biomed. Eng. 30(2), 1983):
    tan(omega/2) = |a.(b x c)| / (|a||b||c| + (a.b)|c| + (a.c)|b| + (b.c)|a|)."""
    a, b, c = (np.asarray(v, np.float64) for v in (a, b, c))
    la, lb, lc = (np.linalg.norm(v, axis=-1) for v in (a, b, c))
    num = np.abs(np.sum(a * np.cross(b, c), axis=-1))
    den = la * lb * lc + np.sum(a * b, axis=-1) * lc + np.sum(a * c, axis=-1) * lb + np.sum(b * c, axis=-1) * la
    return 2.0 * np.arctan2(num, den)


def solid_angle_quad(p, corners):
    """Solid angle of a planar convex quad (corners in order) seen from p."""
    v = np.asarray(corners, np.float64)[:, :3] - np.asarray(p, np.float64)
    return solid_angle_triangle(v[0], v[1], v[2]) + solid_angle_triangle(v[0], v[2], v[3])


def projected_solid_angle(p, normal, corners):
    """Integral of cos(theta) over the solid angle of a polygon seen from p
    (Lambert's polygon formula: 1/2 sum_i Theta_i n_i . N, with Theta_i the
    angle subtended by edge i and n_i the unit normal of the plane through p
    and that edge)."""
    v = np.asarray(corners, np.float64)[:, :3] - np.asarray(p, np.float64)
    u = v / np.linalg.norm(v, axis=1, keepdims=True)
    n = np.asarray(normal, np.float64)
    total = 0.0
    for i in range(len(u)):
        a, b = u[i], u[(i + 1) % len(u)]
        theta = np.arccos(np.clip(np.dot(a, b), -1.0, 1.0))
        cr = np.cross(a, b)
        total += theta * np.dot(cr / np.linalg.norm(cr), n)
    return abs(0.5 * total)


def gauss_rect(corners, n: int):
    """Gauss-Legendre nodes and weights (area measure) over the parallelogram
    p0 + s (p1 - p0) + t (p3 - p0), s, t in [0, 1]."""
    c = np.asarray(corners, np.float64)[:, :3]
    x, w = np.polynomial.legendre.leggauss(n)
    x = 0.5 * (x + 1.0)
    w = 0.5 * w
    e1, e2 = c[1] - c[0], c[3] - c[0]
    area = np.linalg.norm(np.cross(e1, e2))
    s, t = np.meshgrid(x, x, indexing="ij")
    ws = np.outer(w, w) * area
    pts = c[0] + s[..., None] * e1 + t[..., None] * e2
    return pts.reshape(-1, 3), ws.reshape(-1), area, np.cross(e1, e2) / area


def unpack_normal(v: int):
    """10:10:10 normal unpacking (shaders/pack.h): component = q / 511.99999 * 2 - 1."""
    x = (v & 0x3FF) / 511.99999 * 2.0 - 1.0
    y = ((v >> 10) & 0x3FF) / 511.99999 * 2.0 - 1.0
    z = ((v >> 20) & 0xFFF) / 511.99999 * 2.0 - 1.0
    n = np.array([x, y, z])
    return n / np.linalg.norm(n)


def nee_lambert_expectation(P, N, corners, radiance, albedo, light_normal):
    """Expected next-event contribution at the points P (shape (..., 3)) of
    the renderer's estimator for one rect light sampled uniformly in solid
    angle (spherical rectangle, Urena et al. 2013) with the area-measure pdf
    pl = r^2 / (cos_l A), the Lambert evaluation f cos = a/pi cos_P, the
    light-only MIS weight pl / (pl + pb) with pb = (cos_P + 1e-6) / pi, and
    the radiance term L cos_P (so cos_P enters twice):
        c(w) = L a cos_P^2 / (pi (pl + pb))
        E    = 1/S * integral over the light of c cos_l / r^2 dA."""
    pts, ws, area, _ = gauss_rect(corners, 16)
    P = np.asarray(P, np.float64)
    d = pts - P[..., None, :]
    r2 = np.sum(d * d, axis=-1)
    L = d / np.sqrt(r2)[..., None]
    cos_p = L @ np.asarray(N, np.float64)
    cos_l = -(L @ np.asarray(light_normal, np.float64))
    ok = (cos_p > 0) & (cos_l > 0)
    pl = r2 / (np.where(ok, cos_l, 1.0) * area)
    pb = (cos_p + 1e-6) / np.pi
    c = radiance * albedo * cos_p ** 2 / (np.pi * (pl + pb))
    c4 = np.asarray(corners, np.float64)[:, :3]
    v = c4[None, :, :] - P.reshape(-1, 1, 3)
    S = (solid_angle_triangle(v[:, 0], v[:, 1], v[:, 2]) + solid_angle_triangle(v[:, 0], v[:, 2], v[:, 3]))
    S = S.reshape(P.shape[:-1])
    return np.sum(np.where(ok, c * cos_l / r2, 0.0) * ws, axis=-1) / S


def lambert_reflected(P, N, corners, radiance, albedo):
    """Radiance a Lambert surface at each point P reflects from a rect emitter
    of constant radiance: albedo / pi * L * integral of cos over the rect's
    solid angle (Lambert's polygon formula), what an unbiased estimator
    converges to."""
    P = np.asarray(P, np.float64).reshape(-1, 3)
    proj = np.array([projected_solid_angle(p, N, corners) for p in P])
    return albedo / np.pi * radiance * proj


def floor_expectations(s, pixels, sub=2, corrected=False):
    """Closed-form pixel means of the lit-floor scene: the footprint of each
    pixel (sub x sub Gauss points) mapped to the floor by the camera's
    matrices, then at each point the expectation of the reference's estimator
    (next-event estimate with its quirks) or, corrected, the reflected
    radiance itself."""
    cam = s.camera
    W, H = cam.dimension[0], cam.dimension[1]
    Cm = np.array(cam.clipToView, np.float64).reshape(4, 4)  # column-major: Cm[col][row]
    Vm = np.array(cam.viewToWorld, np.float64).reshape(4, 4)
    eye = Vm[3, :3]
    N = unpack_normal(int(s.vb["normal"][0]))
    lightN = np.array([0.0, -1.0, 0.0])
    corners = s.lights["points"][0, :, :3].astype(np.float64)
    radiance = float(s.lights["color"][0, 0])
    albedo = float(s.mats["albedo"][0, 0])
    x, wq = np.polynomial.legendre.leggauss(sub)
    x, wq = 0.5 * (x + 1.0), 0.5 * wq
    pix = np.asarray(pixels, np.float64)
    acc = np.zeros(len(pix))
    for jx, wx in zip(x, wq):
        for jy, wy in zip(x, wq):
            ndc = np.stack([(pix[:, 0] + jx) / W * 2.0 - 1.0, (pix[:, 1] + jy) / H * 2.0 - 1.0,
                            np.ones(len(pix)), np.ones(len(pix))], axis=1)
            v = ndc @ Cm  # clipToView * ndc (Cm holds columns)
            d = v[:, :3] @ Vm[:3, :3]  # viewToWorld * (v, 0)
            P = eye + d * (-eye[1] / d[:, 1])[:, None]
            if corrected:
                acc += wx * wy * lambert_reflected(P, N, corners, radiance, albedo)
            else:
                acc += wx * wy * nee_lambert_expectation(P, N, corners, radiance, albedo, lightN)
    return acc
