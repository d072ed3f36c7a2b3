"""The GPU kernel's compressed 8-wide walk, restated in the oracle
(orc_closest_hit_wide), must return exactly the reference walk's hits
(orc_closest_hit, shaders/raytracing.h:198-247) on the same rays: this pins
the wide algorithm on the CPU (quantised boxes conservative, node-group stack
complete, acceptance rule order-free) before the GPU kernel is compared with
it.  Its counters give the wide kernel's algorithmic bytes."""
import numpy as np
import pytest

import strelka_amd as sa
from oracle import Oracle
from oracle.pyoracle import MODE_BRUTE, MODE_PRUNED, MODE_REF, MODE_WIDE


def same(a, b):
    return a.view(np.uint8).tobytes() == b.view(np.uint8).tobytes()


def scene_oracle(config, variant=0):
    s = sa.SceneData.generate(config, variant)
    b = sa.HostBvh.build(s)
    return s, b, Oracle.from_scene(s, b.nodes, b.wide, b.wide_tris)


@pytest.mark.parametrize("config,variant,seed", [(1, 0, 0x5354524B00000001), (6, 0, 11), (7, 0, 12), (3, 10, 13)])
def test_wide_walk_equals_reference_walk(config, variant, seed):
    s, b, o = scene_oracle(config, variant)
    rays = s.jittered_rays(seed)
    ref, st_ref = o.trace(rays, MODE_REF)
    wide, st_wide = o.trace(rays, MODE_WIDE)
    pruned, st_pr = o.trace(rays, MODE_PRUNED)
    assert same(ref, wide)
    assert st_wide.hits == st_ref.hits > 0
    # an 8-wide node visit stands for several binary node expansions
    assert st_wide.inner < st_pr.inner
    print(f"config {config}: per ray {st_pr.inner / rays.size:.1f} binary expansions, "
          f"{st_wide.inner / rays.size:.1f} wide visits, {st_wide.leaf / rays.size:.1f} triangle tests")


def test_wide_walk_secondary_rays_config2_sample():
    """Incoherent rays (random origins in the scene box, random directions)
    on the 1M-triangle scene: the case the GPU path tracer's bounces hit."""
    s, b, o = scene_oracle(2)
    rng = np.random.default_rng(5)
    n = 20000
    rays = np.zeros(n, sa.RAY)
    lo, hi = np.array(b.header.lo), np.array(b.header.hi)
    rays["o"][:, :3] = rng.uniform(lo, hi, (n, 3))
    rays["o"][:, 3] = 1e9
    d = rng.normal(size=(n, 3))
    rays["d"][:, :3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    # a few exactly axis-aligned directions (infinite inverse components)
    rays["d"][:50, :3] = [0.0, -1.0, 0.0]
    rays["d"][50:100, :3] = [1.0, 0.0, 0.0]
    ref, _ = o.trace(rays, MODE_REF)
    wide, st = o.trace(rays, MODE_WIDE)
    assert same(ref, wide)
    assert st.hits > n // 4


def test_config2_full_frame_fp_edge_count():
    """Every camera ray of the 1080p config-2 frame (2,073,600 rays): the
    number of rays on which the reference's slab test and the wide walk's
    conservative 8-bit boxes reach different hits (a box the reference's
    test rejects although it encloses the closest hit, or the reverse) is
    counted and must be 0; an exhaustive triangle test over all 1.16 M
    triangles agrees with the reference walk on 192 rays spread across the
    frame."""
    s, b, o = scene_oracle(2)
    rays = s.jittered_rays(0x5354524B00000002)
    assert rays.size == 1920 * 1080
    ref, st = o.trace(rays, MODE_REF)
    wide, _ = o.trace(rays, MODE_WIDE)
    rb, wb = ref.view(np.uint8).reshape(rays.size, -1), wide.view(np.uint8).reshape(rays.size, -1)
    edge = int(np.count_nonzero((rb != wb).any(axis=1)))
    print(f"config 2, {rays.size} camera rays: {st.hits} hits, {edge} rays where the walks differ")
    assert edge == 0
    pick = np.arange(0, rays.size, rays.size // 192)[:192]
    brute, _ = o.trace(rays[pick].copy(), MODE_BRUTE)
    assert same(ref[pick], brute)


def test_axis_parallel_rays_cost_what_oblique_rays_cost():
    """Rays parallel to two axes' planes (infinite inverse components): the
    slot test clamps the inverse, so such an axis still rejects the slabs
    that do not hold the origin.  Were the axis left unconstrained, these
    rays would visit a large part of the tree; here they cost about what
    oblique rays do, on the flat tree of the 1M-triangle scene and on the
    two-level structure of the instanced scene (mesh trees in object space)."""
    from oracle.pyoracle import MODE_TWO_LEVEL
    s, b, o = scene_oracle(2)
    rng = np.random.default_rng(9)
    n = 400
    lo, hi = np.array(b.header.lo), np.array(b.header.hi)

    def rays_with(dirs):
        r = np.zeros(n, sa.RAY)
        r["o"][:, :3] = rng.uniform(lo, hi, (n, 3))
        r["o"][:, 3] = 1e9
        r["d"][:, :3] = dirs
        return r

    d = rng.normal(size=(n, 3))
    oblique = rays_with(d / np.linalg.norm(d, axis=1, keepdims=True))
    axis = rays_with(np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, -1.0], [0.0, 1.0, 0.0]] * (n // 4)))
    _, so = o.trace(oblique, MODE_WIDE)
    ref, _ = o.trace(axis, MODE_REF)
    wide, sa_ = o.trace(axis, MODE_WIDE)
    assert same(ref, wide)
    print(f"flat tree: {so.inner / n:.1f} visits per oblique ray, {sa_.inner / n:.1f} per axis-parallel ray")
    assert sa_.inner < 3 * so.inner
    s5 = sa.SceneData.generate(5, 60)
    tl = sa.HostTwoLevel.build(s5)
    o5 = Oracle.from_two_level(s5, tl)
    lo, hi = np.min(tl.top_leaves["lo"], axis=0), np.max(tl.top_leaves["hi"], axis=0)
    oblique, axis = rays_with(d / np.linalg.norm(d, axis=1, keepdims=True)), rays_with(axis["d"][:, :3])
    _, so = o5.trace(oblique, MODE_TWO_LEVEL)
    ref, _ = o5.trace(axis, MODE_REF)
    two, st = o5.trace(axis, MODE_TWO_LEVEL)
    assert same(ref, two)
    print(f"two-level: {so.inner / n:.1f} visits per oblique ray, {st.inner / n:.1f} per axis-parallel ray")
    assert st.inner < 3 * so.inner
