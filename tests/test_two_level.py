"""The two-level structure (a tree per mesh in object space under a tree of
instances, stk_build_two_level): its walk, restated in the oracle
(orc_closest_two_level / orc_any_two_level), must return exactly what the
reference walks (closestHit, anyHit: shaders/raytracing.h:198-286) return over
the same structure expanded into one skip-link tree - the tree the two-level
structure stands for.  The expansion itself is pinned by brute force, and the
hits by the flat tree of the same scene (same closest t)."""
import numpy as np
import pytest

import strelka_amd as sa
from oracle import Oracle
from oracle.pyoracle import MODE_ANY, MODE_ANY_TWO_LEVEL, MODE_BRUTE, MODE_REF, MODE_TWO_LEVEL, fptr

INVALID = 0xFFFFFFFF


def same(a, b):
    return a.view(np.uint8).tobytes() == b.view(np.uint8).tobytes()


def scene_rays(s, box_lo, box_hi, n, seed, tmax=1e9):
    """Rays from the camera towards random instance centres (mostly hits),
    random rays inside the scene box, and a few axis-aligned ones."""
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, sa.RAY)
    cam = np.array(s.camera.camPos[:3], dtype=np.float64)
    centres = s.inst["objectToWorld"].reshape(-1, 16)[:, 12:15].astype(np.float64)
    k = n // 2
    tgt = centres[rng.integers(0, centres.shape[0], k)] + rng.normal(scale=1.0, size=(k, 3))
    d = tgt - cam
    rays["o"][:k, :3] = cam
    rays["d"][:k, :3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["o"][k:, :3] = rng.uniform(box_lo, box_hi, (n - k, 3))
    d = rng.normal(size=(n - k, 3))
    rays["d"][k:, :3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["d"][k:k + 20, :3] = [0.0, -1.0, 0.0]
    rays["d"][k + 20:k + 40, :3] = [1.0, 0.0, 0.0]
    rays["o"][:, 3] = tmax
    return rays


@pytest.fixture(scope="module")
def small_instanced():
    s = sa.SceneData.generate(5, 60)  # 60 instances of the 20 meshes + the light
    tl = sa.HostTwoLevel.build(s)
    o = Oracle.from_two_level(s, tl)
    return s, tl, o


def test_structure_invariants(small_instanced):
    s, tl, _ = small_instanced
    n = s.inst.size
    assert tl.top_leaves.size == int(np.sum(tl.instance_mesh != INVALID)) == n
    # keys: the preorder index of each instance's leaf in the top binary tree
    leaves = np.nonzero(tl.top_nodes["instId"] != -1)[0]
    assert sorted(tl.instance_keys.tolist()) == leaves.tolist()
    assert np.all(tl.top_nodes["instId"][tl.instance_keys] == np.arange(n))
    assert set(tl.top_leaves["instId"].tolist()) == set(range(n))
    assert tl.top_depth + tl.mesh_depth <= 16
    # every instance's world box holds its triangles as getTriangle forms them
    ob = Oracle.from_scene(s, sa.HostBvh.build(s).nodes)
    for leaf in tl.top_leaves:
        i = int(leaf["instId"])
        ic = s.inst[i]
        off, cnt = int(ic["indexOffset"]), int(ic["indexCount"])
        m = np.ascontiguousarray(ic["objectToWorld"], dtype=np.float32).reshape(16)
        out = np.zeros(3, np.float32)
        for vi in s.ib[off:off + cnt][:: max(1, cnt // 60)]:
            p = np.ascontiguousarray(s.vb["position"][vi], dtype=np.float32)
            ob.lib.orc_transform_point(fptr(m), fptr(p), fptr(out))
            assert np.all(out >= leaf["lo"]) and np.all(out <= leaf["hi"])
    # the inverse the walk uses moves mesh points by less than its pad term
    assert np.all(tl.instance_inverses[:, 12] < 1e-4)


def test_expansion_against_brute_force(small_instanced):
    s, tl, o = small_instanced
    lo = np.min(tl.top_leaves["lo"], axis=0)
    hi = np.max(tl.top_leaves["hi"], axis=0)
    rays = scene_rays(s, lo, hi, 1500, 3)
    ref, st = o.trace(rays, MODE_REF)
    brute, _ = o.trace(rays, MODE_BRUTE)
    assert st.hits > rays.size // 3
    assert same(ref, brute)


@pytest.mark.parametrize("seed", [1, 2])
def test_two_level_walk_equals_expanded_reference_walk(small_instanced, seed):
    s, tl, o = small_instanced
    lo = np.min(tl.top_leaves["lo"], axis=0)
    hi = np.max(tl.top_leaves["hi"], axis=0)
    rays = scene_rays(s, lo, hi, 40000, seed)
    ref, st_ref = o.trace(rays, MODE_REF)
    two, st_two = o.trace(rays, MODE_TWO_LEVEL)
    assert st_ref.hits > rays.size // 3
    assert same(ref, two)
    print(f"per ray: {st_ref.inner / rays.size:.1f} boxes (reference walk), "
          f"{st_two.inner / rays.size:.1f} wide visits, {st_two.leaf / rays.size:.1f} triangle tests (two-level)")


def test_two_level_any_hit_equals_expanded_any_hit(small_instanced):
    s, tl, o = small_instanced
    lo = np.min(tl.top_leaves["lo"], axis=0)
    hi = np.max(tl.top_leaves["hi"], axis=0)
    rays = scene_rays(s, lo, hi, 20000, 4, tmax=30.0)
    ref, st = o.trace(rays, MODE_ANY)
    two, _ = o.trace(rays, MODE_ANY_TWO_LEVEL)
    assert st.hits > 100
    got = two["instId"] != INVALID
    assert np.array_equal(ref["instId"] != INVALID, got)
    for f in ("instId", "primId", "t"):
        assert np.array_equal(ref[f][got], two[f][got])


def test_two_level_hits_match_flat_tree(small_instanced):
    """Against the flat tree of the same scene: the same closest t on every
    ray, and the same triangle wherever no other triangle ties that t."""
    s, tl, o = small_instanced
    flat = Oracle.from_scene(s, sa.HostBvh.build(s).nodes)
    lo = np.min(tl.top_leaves["lo"], axis=0)
    hi = np.max(tl.top_leaves["hi"], axis=0)
    rays = scene_rays(s, lo, hi, 20000, 5)
    a, _ = flat.trace(rays, MODE_REF)
    b, _ = o.trace(rays, MODE_TWO_LEVEL)
    assert np.array_equal(a["t"], b["t"])
    differ = (a["instId"] != b["instId"]) | (a["primId"] != b["primId"])
    assert differ.sum() <= max(2, rays.size // 2000)
