"""Scenes for the specular, emissive and normal-mapped materials, shared by
the oracle tests and the GPU tests."""
from __future__ import annotations

import numpy as np

import strelka_amd as sa


def material(kind, albedo=(1.0, 1.0, 1.0), roughness=1.0, emission=(0.0, 0.0, 0.0), ior=0.0, **textures):
    m = np.zeros(1, sa.MATERIAL)
    m["type"] = kind
    m["roughness"] = roughness
    m["albedo"][0, :3] = albedo
    m["albedo"][0, 3] = 1.0
    m["emission"][0] = emission
    m["ior"] = ior
    for k, v in textures.items():
        m[k] = v
    return m


def furnace(mat, sky=None):
    """The outward cube under a constant sky (no lights) with its one
    material replaced."""
    s = sa.SceneData.generate(8)
    s.mats = mat.copy()
    if sky is not None:
        s.sky = tuple(sky)
    return s


def unpack_uv(p):
    p = np.asarray(p, np.uint32)
    t = ((p >> 16) & 0xFFFF).astype(np.float32) / np.float32(16383.99999) * np.float32(20.0) - np.float32(10.0)
    s = (p & 0xFFFF).astype(np.float32) / np.float32(16383.99999) * np.float32(20.0) - np.float32(10.0)
    return s, t


def bump_and_stripes():
    """A normal map of sinusoidal bumps (16 x 16, repeat), a stripe texture
    (8 x 8, mirrored repeat), a roughness ramp (4 x 4, clip) and an emission
    spot (4 x 4, clamp)."""
    n = 16
    ii, jj = np.meshgrid(np.arange(n), np.arange(n))
    a = 0.6 * np.sin(2 * np.pi * ii / n)
    b = 0.4 * np.cos(2 * np.pi * jj / n)
    nm = np.stack([a, b, np.ones_like(a)], -1)
    nm /= np.linalg.norm(nm, axis=-1, keepdims=True)
    normal = np.concatenate([(nm + 1.0) * 0.5, np.ones((n, n, 1))], -1).reshape(-1, 4)
    ii, jj = np.meshgrid(np.arange(8), np.arange(8))
    st = np.where((ii // 2) % 2 == 0, 1.0, 0.3)
    stripes = np.stack([st, 0.8 * st + 0.1, 0.5 + 0.0 * st, np.ones_like(st)], -1).reshape(-1, 4)
    ii, jj = np.meshgrid(np.arange(4), np.arange(4))
    ramp = np.stack([(ii + jj + 1) / 7.0] * 3 + [np.ones((4, 4))], -1).reshape(-1, 4)
    spot = np.zeros((4, 4, 4))
    spot[1:3, 1:3, :3] = (3.0, 2.0, 1.0)
    spot[..., 3] = 1.0
    texels = np.concatenate([normal, stripes, ramp, spot.reshape(-1, 4)]).astype(np.float32)
    textures = np.zeros(4, sa.TEXTURE)
    textures[0] = (n, n, 0, sa.TEX_REPEAT)
    textures[1] = (8, 8, n * n, sa.TEX_MIRRORED_REPEAT)
    textures[2] = (4, 4, n * n + 64, sa.TEX_CLIP)
    textures[3] = (4, 4, n * n + 64 + 16, sa.TEX_CLAMP)
    return textures, texels


def showcase(textured: bool = True):
    """The small heightfield-and-spheres scene (64 x 48) with every material
    kind: the ground normal-mapped with mirrored-repeat stripes (its texture
    coordinates stretched over [-1.5, 2.5] so every wrap is exercised), and
    the eight spheres mirror, clear glass (1.5), tinted glass (1.33), an
    emitter, a GGX with a clip-wrapped roughness, a normal-mapped GGX, an
    emission-textured Lambert, and a plain Lambert.  textured=False drops
    every texture (for the batch, which refuses them)."""
    s = sa.SceneData.generate(6)
    mats = s.mats.copy()
    assert mats.size == 9  # the ground, then one per sphere
    kinds = [
        material(sa.MAT_LAMBERT, (0.55, 0.5, 0.45), normalTexture=1, albedoTexture=2),
        material(sa.MAT_MIRROR, (0.9, 0.8, 0.6)),
        material(sa.MAT_GLASS, (1.0, 1.0, 1.0), ior=1.5),
        material(sa.MAT_GLASS, (0.7, 0.9, 0.8), ior=1.33),
        material(sa.MAT_LAMBERT, (0.2, 0.2, 0.2), emission=(4.0, 3.0, 1.5)),
        material(sa.MAT_GGX, (0.8, 0.6, 0.3), roughness=0.6, roughnessTexture=3),
        material(sa.MAT_GGX, (0.5, 0.7, 0.9), roughness=0.35, normalTexture=1),
        material(sa.MAT_LAMBERT, (0.6, 0.6, 0.6), emission=(1.0, 1.0, 1.0), emissionTexture=4),
        material(sa.MAT_LAMBERT, (0.3, 0.6, 0.3)),
    ]
    for i, m in enumerate(kinds):
        mats[i] = m[0]
        if not textured:
            for f in ("albedoTexture", "roughnessTexture", "normalTexture", "emissionTexture"):
                mats[f][i] = 0
    s.mats = mats
    if textured:
        vb = s.vb.copy()
        u, v = unpack_uv(vb["uv"])
        packed = [sa.pack_uv(float(a) * 4.0 - 1.5, float(b) * 4.0 - 1.5) for a, b in zip(u, v)]
        vb["uv"] = np.array(packed, np.uint32)
        s.vb = vb
        s.textures, s.texels = bump_and_stripes()
    return s
