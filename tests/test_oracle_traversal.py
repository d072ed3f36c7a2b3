"""Pins the oracle's traversals: the reference-order skip-link walk
(shaders/raytracing.h:198-247) against brute force over every leaf, the pruned
near-first order against both, and any-hit (:249-286) by set membership."""
import numpy as np
import pytest

import strelka_amd as sa
from oracle import Oracle
from oracle.pyoracle import MODE_ANY, MODE_BRUTE, MODE_PRUNED, MODE_REF

INVALID = 0xFFFFFFFF
SEED1 = 0x5354524B00000001


@pytest.fixture(scope="module")
def cfg1():
    s = sa.SceneData.generate(1)
    b = sa.HostBvh.build(s)
    return s, b, Oracle.from_scene(s, b.nodes), s.jittered_rays(SEED1)


def same(a, b):
    return a.view(np.uint8).tobytes() == b.view(np.uint8).tobytes()


def test_config1_every_ray_ref_equals_brute_and_pruned(cfg1):
    s, b, o, rays = cfg1
    ref, st = o.trace(rays, MODE_REF)
    brute, _ = o.trace(rays, MODE_BRUTE)
    pruned, st2 = o.trace(rays, MODE_PRUNED)
    assert same(ref, brute)
    assert same(ref, pruned)
    # closed box: every ray hits something
    assert st.hits == rays.size and np.all(ref["instId"] != INVALID)
    # the pruned order visits far fewer inner nodes
    assert st2.inner < st.inner


def test_config1_hits_lie_on_their_triangles(cfg1):
    s, b, o, rays = cfg1
    hits, _ = o.trace(rays[::64], MODE_REF)
    for r, h in zip(rays[::64], hits):
        v0, e0, e1 = o.get_triangle(int(h["instId"]), int(h["primId"]))
        p = r["o"][:3].astype(np.float64) + float(h["t"]) * r["d"][:3].astype(np.float64)
        q = v0 + h["bary"][0] * e0.astype(np.float64) + h["bary"][1] * e1.astype(np.float64)
        assert np.allclose(p, q, atol=1e-4)


def test_any_hit_set_membership(cfg1):
    s, b, o, rays = cfg1
    closest, _ = o.trace(rays, MODE_REF)
    # shorten every ray to 60% of its closest hit: nothing may be found
    short = rays.copy()
    short["o"][:, 3] = closest["t"] * np.float32(0.6)
    anyh, _ = o.trace(short, MODE_ANY)
    assert np.all(anyh["instId"] == INVALID)
    # full length: a hit whenever the closest hit exists, on a triangle that
    # really is hit within tmax
    anyh, _ = o.trace(rays, MODE_ANY)
    assert np.all(anyh["instId"] != INVALID)
    for r, h in zip(rays[::97], anyh[::97]):
        v0, e0, e1 = o.get_triangle(int(h["instId"]), int(h["primId"]))
        d = r["d"][:3].astype(np.float64)
        t = float(h["t"])
        assert 0 < t < r["o"][3]
        p = r["o"][:3] + t * d
        n = np.cross(e0, e1)
        assert abs(np.dot(p - v0, n)) / np.linalg.norm(n) < 1e-3


def test_config2_subsample_pruned_equals_reference_order():
    s = sa.SceneData.generate(2)
    b = sa.HostBvh.build(s)
    o = Oracle.from_scene(s, b.nodes)
    rays = s.jittered_rays(0x5354524B00000002)[::251].copy()
    ref, st = o.trace(rays, MODE_REF)
    pruned, st2 = o.trace(rays, MODE_PRUNED)
    assert same(ref, pruned)
    assert st.hits == rays.size
    assert st2.inner * 4 < st.inner


def test_tmax_and_parallel_rays(cfg1):
    s, b, o, rays = cfg1
    r = rays[:256].copy()
    r["o"][:, 3] = 1e-3  # tmax below every hit
    h, _ = o.trace(r, MODE_REF)
    assert np.all(h["instId"] == INVALID) and np.all(h["t"] == np.float32(1e9))
    # axis-parallel directions from inside the box (1/d = inf on two axes)
    ax = np.zeros(6, dtype=sa.RAY)
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for i, d in enumerate(dirs):
        ax[i]["o"] = (0.0, 0.0, 1.5, 1e9)
        ax[i]["d"] = (*d, 0.0)
    ref, _ = o.trace(ax, MODE_REF)
    brute, _ = o.trace(ax, MODE_BRUTE)
    pruned, _ = o.trace(ax, MODE_PRUNED)
    assert same(ref, brute) and same(ref, pruned)
    assert np.all(ref["instId"] != INVALID)
