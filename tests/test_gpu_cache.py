"""Radiance cache on the GPU through the C ABI: the hand-written encode /
forward / backward / Adam kernels against the FP64 CPU oracle
(oracle/cache_ref.py, FP16 rounding emulated where the kernels round), and
what the trained cache gains over the plain path tracer."""
import numpy as np
import pytest

import strelka_amd as sa
from oracle.cache_ref import CacheOracle, adam

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sa.Context(0)
    s = sa.SceneData.generate(6)
    c.set_scene(s)
    c.build_bvh()
    c.scene = s
    yield c
    c.close()


def records(info, n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, sa.CACHE_RECORD)
    lo = np.array(list(info.lo))
    ext = 1.0 / info.scale
    r["position"] = lo + rng.uniform(0.02, 0.98, (n, 3)) * ext
    nrm = rng.normal(size=(n, 3))
    r["normal"] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    r["direction"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["roughness"] = rng.uniform(0, 1, n)
    r["target"] = rng.uniform(0, 2, (n, 3))
    return r


def random_params(info, seed):
    rng = np.random.default_rng(seed)
    grid = rng.uniform(-0.5, 0.5, info.gridParams).astype(np.float32)
    mlp = rng.uniform(-0.25, 0.25, info.mlpParams).astype(np.float32)
    mlp[16384:].reshape(64, 16)[:, 3:] = 0.0
    return grid, mlp


def test_inference_matches_oracle(ctx):
    print("case: cache inference, 4096 records, log2 table 14", flush=True)
    ctx.cache_create(log2_table=14, seed=7)
    info = ctx.cache_info()
    grid, mlp = random_params(info, 1)
    ctx.cache_set_params(grid, mlp)
    recs = records(info, 4096, 2)
    gpu = ctx.cache_infer(recs)[:, :3]
    ora = CacheOracle(info, grid, mlp, round16=True).predict(recs)
    scale = np.abs(ora).max()
    assert scale > 0.05
    err = np.abs(gpu - ora).max() / scale
    print(f"max relative error {err:.2e}")
    assert err < 5e-3
    exact = CacheOracle(info, grid, mlp, round16=False).predict(recs)
    assert np.abs(gpu - exact).max() / scale < 3e-2


def test_gradients_match_oracle(ctx):
    print("case: cache gradients, 2048 records", flush=True)
    ctx.cache_create(log2_table=14, seed=7)
    info = ctx.cache_info()
    grid, mlp = random_params(info, 3)
    ctx.cache_set_params(grid, mlp)
    recs = records(info, 2048, 4)
    g_grid, g_mlp, loss = ctx.cache_gradients(recs)
    o = CacheOracle(info, grid, mlp, round16=True)
    og, dW, oloss = o.gradients(recs)
    om = o.flat_mlp_grad(dW)
    assert loss == pytest.approx(oloss, rel=1e-2)
    rel_mlp = np.linalg.norm(g_mlp - om) / np.linalg.norm(om)
    rel_grid = np.linalg.norm(g_grid - og.ravel()) / np.linalg.norm(og)
    print(f"relative gradient error: mlp {rel_mlp:.2e}, grid {rel_grid:.2e}")
    assert rel_mlp < 2e-2 and rel_grid < 2e-2
    # per layer too (a transposed or misplaced fragment would fail here)
    for l in range(5):
        a = g_mlp[l * 4096:(l + 1) * 4096] if l < 4 else g_mlp[16384:]
        b = om[l * 4096:(l + 1) * 4096] if l < 4 else om[16384:]
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 3e-2, l
    # unused output columns get no gradient
    assert np.all(g_mlp[16384:].reshape(64, 16)[:, 3:] == 0)


def test_adam_step_matches_oracle(ctx):
    print("case: one Adam step", flush=True)
    ctx.cache_create(log2_table=14, seed=7, lr=1e-2)
    info = ctx.cache_info()
    grid, mlp = random_params(info, 5)
    ctx.cache_set_params(grid, mlp)
    recs = records(info, 1024, 6)
    g_grid, g_mlp, _ = ctx.cache_gradients(recs)
    ctx.cache_train_step(recs)
    p_grid, p_mlp = ctx.cache_get_params()
    n = recs.size
    e_mlp, _, _ = adam(mlp, np.zeros_like(g_mlp), np.zeros_like(g_mlp), g_mlp / n, 1, lr=1e-2)
    e_grid, _, _ = adam(grid, np.zeros_like(g_grid), np.zeros_like(g_grid), g_grid / n, 1, lr=1e-2, sparse=True)
    assert np.mean(np.abs(p_mlp - e_mlp) < 1e-5) > 0.999
    assert np.mean(np.abs(p_grid - e_grid) < 1e-5) > 0.999
    assert ctx.cache_info().step == 1


def test_training_reduces_loss(ctx):
    print("case: 150 training steps on a smooth radiance field", flush=True)
    ctx.cache_create(log2_table=16, seed=3)
    info = ctx.cache_info()
    rng = np.random.default_rng(8)
    losses = []
    for it in range(150):
        r = records(info, 4096, 100 + it)
        p = (r["position"] - np.array(list(info.lo))) * info.scale
        r["target"] = 0.5 + 0.4 * np.sin(6.0 * p) * np.array([1.0, 0.7, 0.4])
        losses.append(ctx.cache_train_step(r))
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.5f}")
    assert np.isfinite(losses[-1])
    assert losses[-1] < 0.1 * losses[0]


def rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


def test_config4_cache_on_the_1m_triangle_scene():
    """Config 4 at its real workload: the 1.16M-triangle scene at 1920x1080,
    a 16-level grid of 2^19 entries per level, 2^16 training paths per frame,
    queries after 2 bounces, 32 frames of 1 spp.  On the records the renderer
    itself produced, GPU inference and gradients equal the FP64 oracle's (FP16
    rounding emulated).  Training is judged against the best constant
    predictor of each frame's fresh records (their mean target): one-sample
    targets are so noisy that this constant's squared error is 0.2-1.4 from
    frame to frame, so only a prediction that has learned how radiance varies
    with position and direction gets below it.  The untrained cache (output
    layer zero) starts above it; after 24 frames it is below it on every
    frame.  The error on the first frame's records also falls."""
    print("case: config 4, 1080p, 2^19 grid, 65,536 training paths, query depth 2, 32 frames", flush=True)
    s = sa.SceneData.generate(2)
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    ctx.cache_create(log2_table=19)
    info = ctx.cache_info()
    assert info.log2TableSize == 19 and info.gridParams == 16 * (1 << 19) * 2
    w, h = s.camera.dimension[0], s.camera.dimension[1]

    def mse(r):
        return float(np.mean((ctx.cache_infer(r)[:, :3] - r["target"]) ** 2))

    evals, first, ratio = None, [], []
    for f in range(32):
        p = sa.render_params(s.camera, spp=1, max_depth=8, frame=f, sampling=1, sky=s.sky, cache_mode=1,
                             cache_query_depth=2, cache_train_paths=65536)
        img = ctx.render(p)
        recs = ctx.cache_last_records("train")
        assert recs.size > 65536  # several vertices per training path
        sub = recs[:: max(1, recs.size // 16384)][:16384].copy()
        if f == 0:
            evals = sub
        t = sub["target"]
        const = float(np.mean((t - t.mean(axis=0)) ** 2))   # best constant predictor
        ratio.append(mse(sub) / const)
        if f < 4 or f >= 28:
            first.append(mse(evals))
    ratio = np.array(ratio)
    print("squared error over the best constant's, per frame: " + " ".join(f"{r:.3f}" for r in ratio))
    print(f"frames 1-8 mean {ratio[:8].mean():.4f}, frames 25-32 mean {ratio[24:].mean():.4f}; on frame 0's "
          f"records, frames 1-4 / 29-32: {np.mean(first[:4]):.4f} / {np.mean(first[4:]):.4f}")
    assert np.all(np.isfinite(img))
    assert ratio[:8].mean() > 0.99
    assert np.all(ratio[24:] < 1.0) and ratio[24:].mean() < 0.98 and ratio[24:].mean() < ratio[:8].mean() - 0.03
    assert np.mean(first[4:]) < np.mean(first[:4])
    cnt = ctx.counters()
    queries = ctx.cache_last_records("query")
    train = ctx.cache_last_records("train")
    print(f"queries {queries.size} (counter {cnt[3]}), training records {train.size} (counter {cnt[4]})")
    assert 0 < queries.size == cnt[3] <= w * h and train.size > 0
    grid, mlp = ctx.cache_get_params()
    ctx.cache_set_params(grid, mlp)  # serve the trained weights themselves (the moving average restarts)
    o = CacheOracle(info, grid, mlp, round16=True)
    q = queries[:: max(1, queries.size // 4096)][:4096].copy()
    gpu = ctx.cache_infer(q)[:, :3]
    ora = o.predict(q)
    scale = np.abs(ora).max()
    err = np.abs(gpu - ora).max() / scale
    print(f"inference on renderer queries: max relative error {err:.2e} (scale {scale:.3f})")
    assert scale > 0.01 and err < 5e-3
    t = train[:: max(1, train.size // 2048)][:2048].copy()
    g_grid, g_mlp, loss = ctx.cache_gradients(t)
    og, dW, oloss = o.gradients(t)
    om = o.flat_mlp_grad(dW)
    rel_mlp = np.linalg.norm(g_mlp - om) / np.linalg.norm(om)
    rel_grid = np.linalg.norm(g_grid - og.ravel()) / np.linalg.norm(og)
    print(f"gradients on renderer training records: mlp {rel_mlp:.2e}, grid {rel_grid:.2e}, "
          f"loss {loss:.4f} vs {oloss:.4f}")
    assert loss == pytest.approx(oloss, rel=1e-2)
    assert rel_mlp < 2e-2 and rel_grid < 2e-2
    ctx.close()


def test_trained_cache_beats_plain_path_tracing_at_equal_samples():
    print("case: config 7 closed room, cached vs plain path tracing, 8 frames of 1 spp", flush=True)
    s = sa.SceneData.generate(7)
    ctx = sa.Context(0)
    ctx.set_scene(s)
    ctx.build_bvh()
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    sky = (0.0, 0.0, 0.0)
    ctx.reset_accumulation()
    ref = ctx.render(sa.render_params(s.camera, spp=1024, max_depth=8, frame=1000, sampling=1, sky=sky,
                                      enable_tonemap=False))[..., :3].copy()
    frames = 8

    def run(mode, base):
        ctx.reset_accumulation()
        out = None
        for f in range(frames):
            p = sa.render_params(s.camera, spp=1, max_depth=8, frame=base + f, sampling=1, sky=sky,
                                 enable_tonemap=False, cache_mode=mode, cache_query_depth=2, cache_train_paths=4096)
            out = ctx.render(p)[..., :3].copy()
        return out

    plain = run(0, 0)
    ctx.cache_create(log2_table=16, seed=11)
    # warm-up: training-only frames (extended paths, no queries)
    for f in range(300):
        p = sa.render_params(s.camera, spp=1, max_depth=8, frame=5000 + f, sampling=1, sky=sky,
                             enable_tonemap=False, cache_mode=3, cache_train_paths=4096)
        ctx.render(p)
    cnt = ctx.counters()
    assert cnt[4] > 4096 and cnt[3] == 0  # training records, no queries in train-only mode
    g, m = ctx.cache_get_params()
    assert np.all(np.isfinite(g)) and np.all(np.isfinite(m))
    cached = run(1, 0)
    cnt = ctx.counters()
    assert 0 < cnt[3] <= w * h  # queries answered this frame
    assert np.all(np.isfinite(cached))
    e_plain, e_cached = rmse(plain, ref), rmse(cached, ref)
    print(f"RMSE vs 1024-spp reference: plain {e_plain:.4f}, cached {e_cached:.4f}, queries {cnt[3]}")
    ctx.close()
    assert e_cached < e_plain


def test_cached_frames_on_pass_streams_equal_one_stream():
    """With several pass streams a pass's bounces overlap the previous pass's
    inference and training, which its own still follow.  Query-only frames
    over a fixed network and train-only frames (their images do not read the
    network) are bit-identical to one stream; query-and-train frames are close
    to one stream's (training itself is not bit-reproducible: records are
    appended and grid gradients scattered with atomics)."""
    print("case: config 7, cached frames of 2 spp, pass streams 1 / 2 / 3", flush=True)
    s = sa.SceneData.generate(7)
    w, h = s.camera.dimension[0], s.camera.dimension[1]

    def run(streams, mode, frames=3):
        ctx = sa.Context(0)
        ctx.set_scene(s)
        ctx.build_bvh()
        ctx.set_launch_config(block_size=64, tile=8, pass_streams=streams)
        ctx.cache_create(log2_table=14, seed=5)
        if mode == 2:
            info = ctx.cache_info()
            grid, mlp = random_params(info, 3)
            ctx.cache_set_params(grid, mlp)
        imgs = []
        for f in range(frames):
            p = sa.render_params(s.camera, spp=2, max_depth=8, frame=f, sampling=1, sky=(0.0, 0.0, 0.0),
                                 keep_samples=True, cache_mode=mode, cache_query_depth=1, cache_train_paths=2048,
                                 cache_train_steps=2)
            imgs.append(ctx.render(p).copy())
        smp = ctx.samples(w * h * 2 * 3)
        ctx.close()
        return imgs, smp

    for mode in (2, 3):
        base = run(1, mode)
        for streams in (2, 3):
            got = run(streams, mode)
            for f in range(3):
                assert np.array_equal(got[0][f], base[0][f]), (mode, streams, f)
            assert np.array_equal(got[1], base[1]), (mode, streams)
    one = run(1, 1)[0][-1][..., :3]
    two = run(2, 1)[0][-1][..., :3]
    rel = np.abs(two - one).mean() / one.mean()
    print(f"query + train: mean relative difference between one and two streams {rel:.2e}")
    assert rel < 0.02


def test_self_training_suffix_ends_training_paths_with_a_query():
    """cacheTrainSuffix n: a training path ends with a query of its own at the
    first vertex of depth >= query depth + n where the other paths query
    (self-training).  A fresh cache answers 0 everywhere (its output layer
    starts at zero), so the frame's samples are plain path-traced samples cut
    after a vertex's direct light: a path that does not train equals the plain
    render with max_depth = query depth + 1, a training path the plain render
    with max_depth = query depth + n + 1, bit for bit.  Training records: at
    most query depth + n per training path, fewer than full-depth training
    paths leave."""
    print("case: config 7 (Lambert room), query depth 1, suffix 2, fresh cache", flush=True)
    s = sa.SceneData.generate(7)
    assert np.all(s.mats["type"] == 0)  # Lambert only: every vertex may query
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    n = w * h
    ctx = sa.Context(0)
    try:
        ctx.set_scene(s)
        ctx.build_bvh()
        qd, suffix, paths = 1, 2, 2048

        def plain(depth):
            p = sa.render_params(s.camera, spp=1, max_depth=depth, frame=5, sampling=1, sky=(0.0, 0.0, 0.0),
                                 keep_samples=True, accumulate=False)
            ctx.render(p)
            return ctx.samples(n * 3).reshape(n, 3).copy()

        short, long = plain(qd + 1), plain(qd + suffix + 1)
        counts = {}
        for sfx in (suffix, 0):
            ctx.cache_create(log2_table=14, seed=11)
            p = sa.render_params(s.camera, spp=1, max_depth=8, frame=5, sampling=1, sky=(0.0, 0.0, 0.0),
                                 keep_samples=True, accumulate=False, cache_mode=1, cache_query_depth=qd,
                                 cache_train_paths=paths, cache_train_suffix=sfx)
            ctx.render(p)
            got = ctx.samples(n * 3).reshape(n, 3).copy()
            counts[sfx] = ctx.cache_last_records("train").size
            if sfx == suffix:
                eq_short = np.all(got.view(np.uint32) == short.view(np.uint32), axis=1)
                eq_long = np.all(got.view(np.uint32) == long.view(np.uint32), axis=1)
                assert np.all(eq_short | eq_long), f"{np.sum(~(eq_short | eq_long))} samples are neither"
                trained = int(np.sum(eq_long & ~eq_short))
                print(f"training samples told apart: {trained} of about {paths}")
                assert 0.5 * paths <= trained <= 1.01 * paths
        print(f"training records: suffix {suffix}: {counts[suffix]}, full depth: {counts[0]}")
        assert 0 < counts[suffix] <= paths * (qd + suffix) * 1.01
        assert counts[suffix] < counts[0]
    finally:
        ctx.close()


def test_queries_covering_direct_light_send_no_shadow_ray():
    """cacheQueryDirect: a query answers for all the radiance its vertex
    reflects, direct light included, so the vertex sends no shadow ray.  With
    a fresh cache (answers 0) in the closed Lambert room (black sky; a light
    hit after a diffuse bounce adds nothing in the reference estimator), a
    frame's paths that do not train equal the plain render with max_depth =
    query depth, the training paths (suffix 2) the plain render with max_depth
    = query depth + 2, bit for bit; the training targets then include each
    vertex's light sample, so they exceed those of the default mode."""
    print("case: config 7, query depth 1, suffix 2, fresh cache, queries covering direct light", flush=True)
    s = sa.SceneData.generate(7)
    assert np.all(s.mats["type"] == 0)
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    n = w * h
    ctx = sa.Context(0)
    try:
        ctx.set_scene(s)
        ctx.build_bvh()
        qd, suffix, paths = 1, 2, 2048

        def plain(depth):
            p = sa.render_params(s.camera, spp=1, max_depth=depth, frame=9, sampling=1, sky=(0.0, 0.0, 0.0),
                                 keep_samples=True, accumulate=False)
            ctx.render(p)
            return ctx.samples(n * 3).reshape(n, 3).copy()

        short, long = plain(qd), plain(qd + suffix)
        targets = {}
        for direct in (True, False):
            ctx.cache_create(log2_table=14, seed=11)
            p = sa.render_params(s.camera, spp=1, max_depth=8, frame=9, sampling=1, sky=(0.0, 0.0, 0.0),
                                 keep_samples=True, accumulate=False, cache_mode=1, cache_query_depth=qd,
                                 cache_train_paths=paths, cache_train_suffix=suffix, cache_query_direct=direct)
            ctx.render(p)
            got = ctx.samples(n * 3).reshape(n, 3).copy()
            targets[direct] = float(ctx.cache_last_records("train")["target"].mean())
            if direct:
                eq_short = np.all(got.view(np.uint32) == short.view(np.uint32), axis=1)
                eq_long = np.all(got.view(np.uint32) == long.view(np.uint32), axis=1)
                assert np.all(eq_short | eq_long), f"{np.sum(~(eq_short | eq_long))} samples are neither"
                trained = int(np.sum(eq_long & ~eq_short))
                print(f"training samples told apart: {trained} of about {paths}")
                assert 0.5 * paths <= trained <= 1.01 * paths
        print(f"mean training target: direct light included {targets[True]:.4f}, excluded {targets[False]:.4f}")
        assert targets[True] > targets[False]
    finally:
        ctx.close()
