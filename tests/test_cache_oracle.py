"""Pins the FP64 cache oracle (oracle/cache_ref.py) on the CPU: finite
differences of the relative-L2 loss against its analytic gradients, the hash
and dense indexing rules, and the first Adam step's closed form."""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle.cache_ref import CacheOracle, adam

REC = np.dtype([("position", "<f4", 3), ("roughness", "<f4"), ("normal", "<f4", 3), ("payload", "<u4"),
                ("direction", "<f4", 3), ("pad0", "<f4"), ("target", "<f4", 3), ("pad1", "<f4")])


def make_info(log2t=8, base=2, maxres=64):
    b = np.exp((np.log(maxres) - np.log(base)) / 15)
    res = [int(np.floor(base * b ** l + 1e-9)) for l in range(16)]
    T = 1 << log2t
    dense = [1 if (r + 1) ** 3 <= T else 0 for r in res]
    return SimpleNamespace(lo=[-1.0, -1.0, -1.0], scale=0.5, resolution=res, dense=dense, log2TableSize=log2t)


def make_problem(seed=0, n=24, log2t=8):
    rng = np.random.default_rng(seed)
    info = make_info(log2t)
    T = 1 << log2t
    grid = rng.uniform(-0.5, 0.5, 16 * T * 2).astype(np.float32)
    mlp = rng.uniform(-0.3, 0.3, 4 * 4096 + 64 * 16).astype(np.float32)
    recs = np.zeros(n, REC)
    recs["position"] = rng.uniform(-1, 1, (n, 3))
    nrm = rng.normal(size=(n, 3))
    recs["normal"] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    recs["direction"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    recs["roughness"] = rng.uniform(0, 1, n)
    recs["target"] = rng.uniform(0, 2, (n, 3))
    return info, grid, mlp, recs


def loss_of(info, grid, mlp, recs):
    o = CacheOracle(info, grid, mlp, round16=False)
    p = o.predict(recs)
    t = recs["target"].astype(np.float64)
    # the denominator is held constant in the gradient: evaluate it at the
    # unperturbed prediction
    return p, t


def test_mlp_gradient_matches_finite_differences():
    info, grid, mlp, recs = make_problem()
    o = CacheOracle(info, grid, mlp, round16=False)
    g_grid, dW, _ = o.gradients(recs)
    gm = o.flat_mlp_grad(dW)
    p0, t = loss_of(info, grid, mlp, recs)
    den = p0 * p0 + 0.01
    rng = np.random.default_rng(3)
    mlp64 = mlp.astype(np.float64)
    for k in list(rng.choice(16384, 12, replace=False)) + list(16384 + 16 * rng.choice(64, 4) + rng.choice(3, 4)):
        eps = 1e-6
        up, dn = mlp64.copy(), mlp64.copy()
        up[k] += eps
        dn[k] -= eps
        o_up = CacheOracle(info, grid, up, round16=False)
        o_dn = CacheOracle(info, grid, dn, round16=False)
        lu = np.sum((o_up.predict(recs) - t) ** 2 / den)
        ld = np.sum((o_dn.predict(recs) - t) ** 2 / den)
        fd = (lu - ld) / (2 * eps)
        assert gm[k] == pytest.approx(fd, rel=1e-4, abs=1e-6), k


def test_grid_gradient_matches_finite_differences():
    info, grid, mlp, recs = make_problem(seed=1)
    o = CacheOracle(info, grid, mlp, round16=False)
    g_grid, _, _ = o.gradients(recs)
    p0, t = loss_of(info, grid, mlp, recs)
    den = p0 * p0 + 0.01
    flat = g_grid.ravel()
    touched = np.flatnonzero(flat)
    assert touched.size > 100
    rng = np.random.default_rng(4)
    g64 = grid.astype(np.float64)
    for k in rng.choice(touched, 16, replace=False):
        eps = 1e-6
        up, dn = g64.copy(), g64.copy()
        up[k] += eps
        dn[k] -= eps
        lu = np.sum((CacheOracle(info, up, mlp, round16=False).predict(recs) - t) ** 2 / den)
        ld = np.sum((CacheOracle(info, dn, mlp, round16=False).predict(recs) - t) ** 2 / den)
        assert flat[k] == pytest.approx((lu - ld) / (2 * eps), rel=1e-4, abs=1e-7), k


def test_encoding_indexing_rules():
    info = make_info(log2t=12, base=4, maxres=256)
    o = CacheOracle(info, np.zeros(16 * 4096 * 2, np.float32), np.zeros(17408, np.float32), round16=False)
    p = np.array([[0.5, 0.25, 0.75]])
    # dense level: corner (0,0,0) of the cell containing p
    lvl = 0
    r = info.resolution[lvl]
    assert info.dense[lvl] == 1
    idx, w = o.corners(lvl, p)
    i = np.floor(p[0] * r).astype(int)
    assert idx[0, 0] == i[0] + (r + 1) * (i[1] + (r + 1) * i[2])
    assert w.sum() == pytest.approx(1.0)
    # hashed level: (x * 1) ^ (y * 2654435761) ^ (z * 805459861) mod T, 32-bit products
    lvl = 15
    assert info.dense[lvl] == 0
    r = info.resolution[lvl]
    idx, w = o.corners(lvl, p)
    i = np.floor(p[0] * r).astype(np.uint64)
    h = (int(i[0]) ^ ((int(i[1]) * 2654435761) & 0xFFFFFFFF) ^ ((int(i[2]) * 805459861) & 0xFFFFFFFF)) % 4096
    assert int(idx[0, 0]) == h
    # a point on a lattice vertex puts all weight on one corner
    q = np.array([[2.0 / r, 3.0 / r, 5.0 / r]])
    _, w = o.corners(lvl, q)
    assert w[0, 0] == pytest.approx(1.0) and np.sum(w[0, 1:]) == pytest.approx(0.0, abs=1e-12)


def test_adam_first_step_closed_form():
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    p = np.array([1.0, 1.0, 1.0, 1.0])
    p1, m1, v1 = adam(p, np.zeros(4), np.zeros(4), g, step=1, lr=1e-2, sparse=True)
    # bias-corrected first step: m^ = g, v^ = g^2 -> p -= lr * sign(g); zero gradient untouched
    assert np.allclose(p1[:3], 1.0 - 1e-2 * np.sign(g[:3]), atol=1e-9)
    assert p1[3] == 1.0 and m1[3] == 0.0
