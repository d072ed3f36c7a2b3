"""Index arithmetic of the device's locally-ordered clustering
(strelka_amd/csrc/gpubuild.hip: k_ploc_nn, k_ploc_flags, the two exclusive
scans, k_ploc_write), mirrored in numpy: every inner index is written exactly
once, the last merge is the root (inner node 0), every leaf is reached once
with consistent counts, and the pair order (area, distance, parity of the
lower position, lower position) keeps the number of rounds logarithmic even
for equal boxes.  The GPU tree itself is checked by test_gpu_build.py."""
import sys

import numpy as np

RADIUS = 32
LEAF = 1 << 31


def half_area(lo, hi):
    d = hi - lo
    return (d[..., 0] * d[..., 1] + d[..., 1] * d[..., 2]) + d[..., 2] * d[..., 0]


def cluster(lo0, hi0):
    n = len(lo0)
    ninner = n - 1
    C = np.array([k | LEAF for k in range(n)], np.int64)
    lo, hi, cnt = lo0.copy(), hi0.copy(), np.ones(n, np.int64)
    child = np.full((ninner, 2), -1, np.int64)
    count = np.zeros(ninner, np.int64)
    written = np.zeros(ninner, np.int64)
    m, base, rounds = n, 0, 0
    while m > 1:
        rounds += 1
        nn = np.zeros(m, np.int64)
        for i in range(m):
            j0, j1 = max(0, i - RADIUS), min(m - 1, i + RADIUS)
            js = np.array([j for j in range(j0, j1 + 1) if j != i])
            a = half_area(np.minimum(lo[i], lo[js]), np.maximum(hi[i], hi[js]))
            low = np.minimum(i, js)
            rank = (np.maximum(i, js) - low) * 2 + (low & 1)
            nn[i] = js[np.lexsort((js, rank, a))[0]]  # least (area, rank, position)
        idx = np.arange(m)
        mutual = nn[nn] == idx
        merge = (mutual & (idx < nn)).astype(np.int64)
        keep = (~(mutual & (idx > nn))).astype(np.int64)
        mrank = np.cumsum(merge) - merge
        pos = np.cumsum(keep) - keep
        m2 = int(pos[-1] + keep[-1])
        assert 0 < m2 < m
        C2, lo2, hi2, cnt2 = np.zeros(m2, np.int64), np.zeros((m2, 3), lo.dtype), np.zeros((m2, 3), lo.dtype), \
            np.zeros(m2, np.int64)
        for i in np.flatnonzero(keep):
            p = pos[i]
            if merge[i]:
                j = nn[i]
                k = ninner - 1 - (base + mrank[i])
                assert 0 <= k < ninner
                written[k] += 1
                child[k] = (C[i], C[j])
                count[k] = cnt[i] + cnt[j]
                C2[p], lo2[p], hi2[p], cnt2[p] = k, np.minimum(lo[i], lo[j]), np.maximum(hi[i], hi[j]), count[k]
            else:
                C2[p], lo2[p], hi2[p], cnt2[p] = C[i], lo[i], hi[i], cnt[i]
        base += m - m2
        m, C, lo, hi, cnt = m2, C2, lo2, hi2, cnt2
    assert (written == 1).all() and C[0] == 0
    seen = np.zeros(n, np.int64)
    sys.setrecursionlimit(10000)

    def walk(r):
        if r & LEAF:
            seen[r & ~LEAF] += 1
            return 1
        c = walk(child[r][0]) + walk(child[r][1])
        assert c == count[r]
        return c

    assert walk(0) == n and (seen == 1).all()
    return rounds


def test_random_boxes():
    rng = np.random.default_rng(1)
    for n in (2, 3, 17, 400):
        c = np.sort(rng.uniform(0, 10, (n, 3)).astype(np.float32), axis=0)
        e = rng.uniform(0.01, 0.2, (n, 3)).astype(np.float32)
        assert cluster(c - e, c + e) <= 4 * int(np.log2(n)) + 4


def test_equal_boxes_finish_in_logarithmic_rounds():
    n = 1000
    rounds = cluster(np.zeros((n, 3), np.float32), np.ones((n, 3), np.float32))
    assert rounds <= 2 * int(np.log2(n)) + 2, rounds
    rng = np.random.default_rng(2)
    c = np.repeat(rng.uniform(0, 10, (20, 3)).astype(np.float32), 30, axis=0)
    assert cluster(c - 0.1, c + 0.1) <= 40
