"""The 8-wide collapse keeps the set of binary nodes of least summed surface
area.  On small trees every legal collapse is enumerated (each kept node's
children, cut at the next kept nodes or leaves, fill at most 8 slots; the root
is always kept), and the collapse read back from the built StkWideNode array
must cost exactly the least of them.  A greedy opening rule, a wrong slot
budget or a dropped term in the plan's recursion costs more on some of these
trees."""
import numpy as np
import pytest

import strelka_amd as sa

from test_wide_bvh import binary_parents, lca


def small_scene(config, ntri, first):
    s = sa.SceneData.generate(config)
    s.ib = s.ib[3 * first:3 * (first + ntri)].copy()
    inst = s.inst[:1].copy()
    inst["indexOffset"] = 0
    inst["indexCount"] = 3 * ntri
    inst["lightId"] = -1
    s.inst = inst
    return s


def areas(nodes):
    lo, hi = nodes["minBounds"].astype(np.float64), nodes["maxBounds"].astype(np.float64)
    d = hi - lo
    return d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0]


def least_collapse_cost(nodes):
    """Minimum over every legal choice of kept inner nodes, by enumeration."""
    inner = [i for i in range(nodes.size) if nodes["instId"][i] == -1]
    kids = {i: (i + 1, int(nodes["nodeOffset"][i + 1])) for i in inner}
    a = areas(nodes)
    others = inner[1:]
    best = np.inf
    for mask in range(1 << len(others)):
        kept = {0} | {others[b] for b in range(len(others)) if (mask >> b) & 1}

        def slots(i):
            if i not in kids or i in kept:
                return 1
            return slots(kids[i][0]) + slots(kids[i][1])

        if all(slots(kids[k][0]) + slots(kids[k][1]) <= 8 for k in kept):
            best = min(best, sum(a[k] for k in kept))
    return best


def built_collapse(b):
    """The binary nodes the wide tree kept: for each wide node, the lowest
    common ancestor of the leaves below it."""
    W, T = b.wide, b.wide_tris
    parent = binary_parents(b.nodes)

    def leaves(w):
        node = W[w]
        out = list(T["leaf"][int(node["triBase"]):int(node["triBase"]) + bin(int(node["vmask"]) & ~int(node["imask"])).count("1")])
        for c in range(bin(int(node["imask"])).count("1")):
            out += leaves(int(node["childBase"]) + c)
        return out

    kept = []
    for w in range(W.size):
        ls = [int(x) for x in leaves(w)]
        k = ls[0]
        for x in ls[1:]:
            k = lca(parent, k, x)
        kept.append(k)
    return kept


@pytest.mark.parametrize("config,ntri,first", [(6, 12, 0), (6, 14, 40), (1, 13, 100), (1, 15, 400), (7, 14, 7)])
def test_collapse_is_least_summed_area(config, ntri, first):
    s = small_scene(config, ntri, first)
    b = sa.HostBvh.build(s)
    nodes = b.nodes
    assert nodes.size == 2 * ntri - 1
    kept = built_collapse(b)
    assert len(set(kept)) == len(kept) and kept[0] == 0
    a = areas(nodes)
    got = float(sum(a[k] for k in kept))
    want = float(least_collapse_cost(nodes))
    assert got == pytest.approx(want, rel=1e-12, abs=0.0)
