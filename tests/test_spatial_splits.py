"""Spatial splits in the host SAH builder (StkBvhOptions::spatialSplitAlpha;
the reference builds with Embree's spatial splits, src/render/bvh.cpp:80-105,
154).  A split tree keeps the skip-link preorder layout with one triangle
reference per leaf, holds every triangle at least once, and bounds every
triangle's parts: the reference walk over it equals an exhaustive test of all
triangles on every config-1 and config-7 camera ray, and the walks the kernels
run (pruned binary, compressed 8-wide) equal the reference walk except on rays
that hit two triangles at exactly the same t, where a triangle held in two
leaves can reach the reference walk's preorder earlier."""
import numpy as np
import pytest

import strelka_amd as sa
from oracle import Oracle
from oracle.pyoracle import MODE_BRUTE, MODE_PRUNED, MODE_REF, MODE_WIDE

INVALID = 0xFFFFFFFF


def walk_structure(nodes):
    """(leaf keys in preorder, max depth); asserts preorder and skip links."""
    inst = nodes["instId"].view(np.uint32)
    is_leaf = inst != INVALID
    stack = [(0, INVALID, 0)]
    order, keys, depth = [], [], 0
    while stack:
        i, skip, d = stack.pop()
        order.append(i)
        depth = max(depth, d)
        assert int(nodes["nodeOffset"][i]) & 0xFFFFFFFF == skip
        if is_leaf[i]:
            keys.append((int(inst[i]), int(nodes["minBounds"][i][:1].view(np.uint32)[0])))
            continue
        left, right = i + 1, int(nodes["nodeOffset"][i + 1])
        lo, hi = nodes["minBounds"][i], nodes["maxBounds"][i]
        for c in (left, right):
            if not is_leaf[c]:
                assert np.all(lo <= nodes["minBounds"][c]) and np.all(nodes["maxBounds"][c] <= hi)
        stack.append((right, skip, d + 1))
        stack.append((left, right, d + 1))
    assert order == list(range(nodes.size))
    return keys, depth


def same(a, b):
    return a.view(np.uint8).tobytes() == b.view(np.uint8).tobytes()


@pytest.mark.parametrize("config,seed", [(1, 0x5354524B00000001), (7, 3)])
def test_split_tree_layout_coverage_and_walks(config, seed):
    s = sa.SceneData.generate(config)
    ntri = int(s.inst["indexCount"].sum()) // 3
    plain = sa.HostBvh.build(s)
    split = sa.HostBvh.build(s, spatial_split_alpha=1e-5)
    keys, depth = walk_structure(split.nodes)
    every = {(i, p) for i in range(s.inst.size) for p in range(int(s.inst["indexCount"][i]) // 3)}
    assert set(keys) == every                            # every triangle at least once
    assert ntri < len(keys) <= int(ntri * 1.3) + 2       # some triangles split, within the budget
    assert split.nodes.size == 2 * len(keys) - 1 and depth <= 32
    rays = s.jittered_rays(seed)
    o = Oracle.from_scene(s, split.nodes, split.wide, split.wide_tris)
    ref, st = o.trace(rays, MODE_REF)
    assert st.hits == rays.size                          # closed rooms
    assert same(ref, o.trace(rays, MODE_BRUTE)[0])
    assert same(ref, o.trace(rays, MODE_PRUNED)[0])
    wide, sw = o.trace(rays, MODE_WIDE)
    assert same(ref, wide)
    # the walls' large triangles split: the binary walk expands far fewer nodes
    po = Oracle.from_scene(s, plain.nodes, plain.wide, plain.wide_tris)
    _, sp = po.trace(rays, MODE_PRUNED)
    _, ss = o.trace(rays, MODE_PRUNED)
    print(f"config {config}: {len(keys)} references for {ntri} triangles; binary node expansions per ray "
          f"{sp.inner / rays.size:.1f} -> {ss.inner / rays.size:.1f}")
    assert ss.inner < 0.85 * sp.inner


def test_config2_frame_differs_only_on_exact_ties_and_builds_deterministically():
    s = sa.SceneData.generate(2)
    split = sa.HostBvh.build(s, spatial_split_alpha=1e-5)
    # the threaded build is deterministic (the split budget goes down the tree)
    assert sa.HostBvh.build(s, spatial_split_alpha=1e-5).nodes.tobytes() == split.nodes.tobytes()
    o = Oracle.from_scene(s, split.nodes, split.wide, split.wide_tris)
    rays = s.jittered_rays(0x5354524B00000002)
    ref, _ = o.trace(rays, MODE_REF)
    wide, sw = o.trace(rays, MODE_WIDE)
    diff = np.nonzero((ref.view(np.uint8).reshape(rays.size, -1) != wide.view(np.uint8).reshape(rays.size, -1))
                      .any(axis=1))[0]
    print(f"config 2 split tree: {diff.size} of {rays.size} rays differ between the reference and wide walks; "
          f"{sw.inner / rays.size:.2f} wide visits, {sw.leaf / rays.size:.2f} tests per ray")
    assert diff.size <= 16
    for i in diff:
        # the same t on two different triangles: a tie at a shared edge
        assert ref["t"][i] == wide["t"][i]
        assert (ref["instId"][i], ref["primId"][i]) != (wide["instId"][i], wide["primId"][i])


def test_two_level_builds_ignore_spatial_splits():
    s = sa.SceneData.generate(5, 60)
    a = sa.HostTwoLevel.build(s)
    lib = sa.lib()
    import ctypes as C
    opts = sa.BvhOptions(32, 32, 1.0, 1.0, 1e-5)
    h = C.c_void_p()
    assert lib.stk_tlas_build_host(sa._ptr(s.vb), s.vb.size, sa._ptr(s.ib), s.ib.size, sa._ptr(s.inst),
                                   s.inst.size, C.byref(opts), C.byref(h)) == 0
    cnt = C.c_uint32()
    top = lib.stk_tlas_top_wide(h, C.byref(cnt))
    assert cnt.value == a.top_wide.size and top
    lib.stk_tlas_free(h)
