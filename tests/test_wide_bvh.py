"""Invariants of the compressed 8-wide BVH (StkWideNode) built from the binary
SAH tree: every triangle in exactly one leaf slot, consecutive child and
triangle ranges, and decoded child boxes (in the traversal's float32
arithmetic) that contain everything below them.  The GPU traversal's hits
are checked bit-exact against the oracle in the GPU tests; a box that failed
to contain its subtree would lose hits there, and fails here first."""
import numpy as np
import pytest

import strelka_amd as sa


def decode(node, slot):
    scale = np.array([np.float32(2.0) ** (int(e) - 127) for e in node["exps"]], np.float32)
    o = node["origin"].astype(np.float32)
    lo = o + node["qlo"][:, slot].astype(np.float32) * scale
    hi = o + node["qhi"][:, slot].astype(np.float32) * scale
    return lo.astype(np.float32), hi.astype(np.float32)


def tri_boxes(tris):
    v0 = tris["v0"].astype(np.float32)
    v1 = (v0 + tris["e0"].astype(np.float32)).astype(np.float32)
    v2 = (v0 + tris["e1"].astype(np.float32)).astype(np.float32)
    return np.minimum(np.minimum(v0, v1), v2), np.maximum(np.maximum(v0, v1), v2)


def binary_parents(nodes):
    """Parent of every node of the preorder skip-link tree."""
    parent = np.full(nodes.size, -1, np.int64)
    for i in range(nodes.size):
        if nodes["instId"][i] == -1:  # inner: left = i + 1, right = skip of the left child
            parent[i + 1] = i
            parent[nodes["nodeOffset"][i + 1]] = i
    return parent


def lca(parent, a, b):
    seen = set()
    while a != -1:
        seen.add(a)
        a = parent[a]
    while b not in seen:
        b = parent[b]
    return b


@pytest.mark.parametrize("config,variant", [(1, 0), (6, 0), (7, 0), (9, 0), (3, 10)])
def test_wide_bvh_invariants(config, variant):
    s = sa.SceneData.generate(config, variant)
    b = sa.HostBvh.build(s)
    W, T = b.wide, b.wide_tris
    n = W.size
    assert n >= 1 and T.size == b.tris.size
    # same triangles, each once
    assert np.array_equal(np.sort(T["leaf"]), np.sort(b.tris["leaf"]))
    assert np.unique(T["leaf"]).size == T.size
    tlo, thi = tri_boxes(T)
    parent = binary_parents(b.nodes)
    child_seen = np.zeros(n, np.int32)
    tri_seen = np.zeros(T.size, np.int32)
    # preorder leaf range of every wide subtree, bottom-up (children are created
    # after their parent, so walk the nodes in reverse)
    leaf_min = np.full(n, np.iinfo(np.int64).max, np.int64)
    leaf_max = np.full(n, -1, np.int64)
    for i in range(n - 1, -1, -1):
        nd = W[i]
        vm, im = int(nd["vmask"]) & 0xFF, int(nd["imask"])
        assert vm and (im & ~vm) == 0
        inner = leaf = 0
        for sl in range(8):
            if not (vm >> sl) & 1:
                continue
            lo, hi = decode(nd, sl)
            if (im >> sl) & 1:
                c = int(nd["childBase"]) + inner
                inner += 1
                assert i < c < n
                child_seen[c] += 1
                # the binary node the child stands for is the lowest common
                # ancestor of its subtree's first and last leaves: its box
                # must lie inside the decoded one
                r = lca(parent, leaf_min[c], leaf_max[c])
                assert b.nodes["instId"][r] == -1
                assert np.all(lo <= b.nodes["minBounds"][r]) and np.all(hi >= b.nodes["maxBounds"][r]), (i, sl)
                cmin, cmax = leaf_min[c], leaf_max[c]
            else:
                t = int(nd["triBase"]) + leaf
                leaf += 1
                tri_seen[t] += 1
                assert np.all(lo <= tlo[t]) and np.all(hi >= thi[t]), (i, sl)
                cmin = cmax = int(T["leaf"][t])
            leaf_min[i] = min(leaf_min[i], cmin)
            leaf_max[i] = max(leaf_max[i], cmax)
    assert child_seen[0] == 0 and np.all(child_seen[1:] == 1)
    assert np.all(tri_seen == 1)
    assert 1 <= b.wide_depth <= 16
    # an 8-wide tree is much shallower than the binary one
    assert b.wide_depth <= max(2, (b.depth + 2) // 2)
    lights = s.inst["lightId"] != -1
    assert b.num_emitters == int(np.sum(s.inst["indexCount"][lights]) // 3)


def test_wide_bvh_single_triangle():
    s = sa.SceneData.generate(6)
    keep = 3
    s.ib = s.ib[:keep]
    inst = s.inst[:1].copy()
    inst["indexOffset"] = 0
    inst["indexCount"] = 3
    inst["lightId"] = -1
    s.inst = inst
    b = sa.HostBvh.build(s)
    assert b.wide.size == 1 and b.wide_tris.size == 1
    assert int(b.wide[0]["vmask"]) == 1 << [sl for sl in range(8) if (int(b.wide[0]["vmask"]) >> sl) & 1][0]
    assert int(b.wide[0]["imask"]) == 0
