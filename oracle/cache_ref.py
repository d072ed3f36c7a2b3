"""FP64 CPU oracle of the radiance cache (TEST INFRASTRUCTURE ONLY).

The cache has no counterpart in the reference renderer; this module restates
the published methods the HIP kernels implement so their outputs and gradients
can be checked:

* multi-resolution hash encoding (Muller et al. 2022, "Instant Neural Graphics
  Primitives"): per level a grid of resolution floor(N_min * b^l), dense
  indexing while (res+1)^3 fits the table, otherwise the spatial hash
  (x * 1) ^ (y * 2654435761) ^ (z * 805459861) mod T, trilinear weights;
* a bias-free MLP 64 -> 4 x 64 (ReLU) -> 16 (3 used), inputs = 32 grid
  features, normal, outgoing direction, roughness, constant 1, zero padding;
* relative L2 loss sum_c (p - t)^2 / (p^2 + 0.01) with the denominator held
  constant (Muller et al. 2021, "Real-time Neural Radiance Caching");
* Adam (Kingma & Ba 2015) with bias correction; grid entries with an exactly
  zero gradient are skipped.

`round16=True` rounds to FP16 exactly where the kernels do (grid and weight
copies, MLP inputs, hidden activations, output gradients); with False the
whole computation is FP64, which finite differences pin.
"""
from __future__ import annotations

import numpy as np

LEVELS, FEATS, WIDTH, OUT = 16, 2, 64, 16
P1, P2 = np.uint64(2654435761), np.uint64(805459861)


def h16(x):
    return np.asarray(x, dtype=np.float64).astype(np.float16).astype(np.float64)


class CacheOracle:
    def __init__(self, info, grid, mlp, round16=True):
        """info: StkCacheInfo-like (lo, scale, resolution, dense, log2TableSize);
        grid: flat FP32 [level][entry][feature]; mlp: flat FP32 W[l][in][out]."""
        self.lo = np.array(list(info.lo), np.float64)
        self.scale = float(info.scale)
        self.res = [int(r) for r in info.resolution]
        self.dense = [int(d) for d in info.dense]
        self.T = 1 << int(info.log2TableSize)
        self.round16 = round16
        r = h16 if round16 else (lambda x: np.asarray(x, np.float64))
        self.grid = r(np.asarray(grid)).reshape(LEVELS, self.T, FEATS)
        m = np.asarray(mlp)
        self.W = [r(m[l * 4096:(l + 1) * 4096]).reshape(WIDTH, WIDTH) for l in range(4)]
        self.W.append(r(m[16384:16384 + WIDTH * OUT]).reshape(WIDTH, OUT))

    # -- encoding -------------------------------------------------------------
    def corners(self, lvl, p):
        """p: (N,3) normalised positions -> indices (N,8) uint64, weights (N,8).

        The lattice coordinate is formed in binary32 (x = p * res, floor,
        fraction), the arithmetic of the kernels, so both pick the same cell."""
        res = self.res[lvl]
        x = np.asarray(p, np.float32) * np.float32(res)
        f = np.floor(x)
        t = (x - f).astype(np.float64)
        i = f.astype(np.int64)
        over = i >= res
        i = np.where(over, res - 1, i)
        t = np.where(over, 1.0, t)
        idx = np.zeros((p.shape[0], 8), np.uint64)
        w = np.ones((p.shape[0], 8), np.float64)
        for c in range(8):
            o = np.array([c & 1, (c >> 1) & 1, c >> 2])
            cc = (i + o).astype(np.uint64)
            if self.dense[lvl]:
                s = np.uint64(res + 1)
                idx[:, c] = cc[:, 0] + s * (cc[:, 1] + s * cc[:, 2])
            else:
                with np.errstate(over="ignore"):
                    h = (cc[:, 0] * np.uint64(1)) ^ ((cc[:, 1] * P1) & np.uint64(0xFFFFFFFF)) ^ \
                        ((cc[:, 2] * P2) & np.uint64(0xFFFFFFFF))
                idx[:, c] = (h & np.uint64(0xFFFFFFFF)) & np.uint64(self.T - 1)
            for a in range(3):
                w[:, c] *= np.where(o[a] == 1, t[:, a], 1.0 - t[:, a])
        return idx, w

    def normalised(self, pos):
        """(p - lo) * scale clamped to [0, 1], in binary32 like the kernels."""
        p = (np.asarray(pos, np.float32) - self.lo.astype(np.float32)) * np.float32(self.scale)
        return np.clip(p, np.float32(0.0), np.float32(1.0))

    def inputs(self, recs):
        p = self.normalised(recs["position"])
        n = p.shape[0]
        x = np.zeros((n, WIDTH), np.float64)
        self._corner_cache = []
        for lvl in range(LEVELS):
            idx, w = self.corners(lvl, p)
            self._corner_cache.append((idx, w))
            vals = self.grid[lvl][idx.astype(np.int64)]  # (N,8,2)
            x[:, 2 * lvl:2 * lvl + 2] = np.einsum("nc,ncf->nf", w, vals)
        x[:, 32:35] = recs["normal"]
        x[:, 35:38] = recs["direction"]
        x[:, 38] = recs["roughness"]
        x[:, 39] = 1.0
        return h16(x) if self.round16 else x

    # -- MLP ------------------------------------------------------------------
    def forward(self, recs):
        x = self.inputs(recs)
        acts = [x]
        h = x
        for l in range(4):
            h = np.maximum(h @ self.W[l], 0.0)
            if self.round16:
                h = h16(h)
            acts.append(h)
        out = h @ self.W[4]
        return acts, out

    def predict(self, recs):
        return self.forward(recs)[1][:, :3]

    def gradients(self, recs):
        """Gradients of the summed relative-L2 loss: (grid (L,T,2), [dW0..dW4], loss)."""
        acts, out = self.forward(recs)
        p = out[:, :3]
        t = np.asarray(recs["target"], np.float64)
        den = p * p + 0.01
        loss = float(np.sum((p - t) ** 2 / den))
        dy = np.zeros_like(out)
        dy[:, :3] = 2.0 * (p - t) / den
        if self.round16:
            dy = h16(dy)
        dW = [None] * 5
        dW[4] = acts[4].T @ dy
        d = dy @ self.W[4].T
        for l in range(3, -1, -1):
            d = d * (acts[l + 1] > 0)
            if self.round16:
                d = h16(d)
            dW[l] = acts[l].T @ d
            d = d @ self.W[l].T
        dfeat = d[:, :32]
        g = np.zeros((LEVELS, self.T, FEATS), np.float64)
        for lvl in range(LEVELS):
            idx, w = self._corner_cache[lvl]
            for f in range(FEATS):
                np.add.at(g[lvl, :, f], idx.astype(np.int64).ravel(), (w * dfeat[:, 2 * lvl + f][:, None]).ravel())
        return g, dW, loss

    def flat_mlp_grad(self, dW):
        return np.concatenate([dW[l].ravel() for l in range(4)] + [dW[4].ravel()])


def adam(p, m, v, g, step, lr=1e-2, b1=0.9, b2=0.99, eps=1e-15, sparse=False):
    """One Adam step in FP64; returns new (p, m, v)."""
    p, m, v, g = (np.asarray(a, np.float64).copy() for a in (p, m, v, g))
    mask = (g != 0) if sparse else np.ones_like(g, bool)
    m[mask] = b1 * m[mask] + (1 - b1) * g[mask]
    v[mask] = b2 * v[mask] + (1 - b2) * g[mask] ** 2
    bc1, bc2 = 1 - b1 ** step, 1 - b2 ** step
    p[mask] -= lr * (m[mask] / bc1) / (np.sqrt(v[mask] / bc2) + eps)
    return p, m, v
