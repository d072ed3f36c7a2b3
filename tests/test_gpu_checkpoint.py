"""Checkpoint / resume (stk_checkpoint_save / stk_checkpoint_load): a
progressive render stopped, saved, restored into a fresh context and
continued gives the same bytes as the uninterrupted run; a trained radiance
cache comes back with its weights, Adam moments, moving averages and step
count; damaged files and mismatched caches are refused before any state
changes."""
import numpy as np
import pytest

import strelka_amd as sa
from test_gpu_cache import records

pytestmark = pytest.mark.gpu


def fresh(scene):
    c = sa.Context(0)
    c.set_scene(scene)
    c.build_bvh()
    return c


def params(scene, frame, **kw):
    return sa.render_params(scene.camera, spp=2, max_depth=6, frame=frame, sampling=1, sky=scene.sky, **kw)


def test_accumulation_resumes_bit_identical(tmp_path):
    print("case: checkpoint after 4 of 8 accumulated frames, config 7", flush=True)
    s = sa.SceneData.generate(7)
    path = tmp_path / "run.stkckpt"
    a = fresh(s)
    for f in range(4):
        a.render(params(s, f))
    a.checkpoint_save(path, user_frame=4)
    for f in range(4, 8):
        want = a.render(params(s, f))
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    want_acc = a.accumulation(w, h)
    a.close()

    b = fresh(s)
    info = b.checkpoint_load(path)
    assert info["accumulation"] and not info["cache"]
    assert info["iteration"] == 4 and info["user_frame"] == 4 and info["pixels"] == w * h
    for f in range(info["user_frame"], 8):
        got = b.render(params(s, f))
    assert got.tobytes() == want.tobytes()
    assert b.accumulation(w, h).tobytes() == want_acc.tobytes()
    b.close()


def test_refused_checkpoints_change_nothing(tmp_path):
    print("case: damaged / mismatched checkpoints are refused", flush=True)
    s = sa.SceneData.generate(6)
    path = tmp_path / "c.stkckpt"
    a = fresh(s)
    a.cache_create(log2_table=14, seed=3)
    a.render(params(s, 0, cache_mode=1, cache_query_depth=1, cache_train_paths=4096))
    a.checkpoint_save(path, user_frame=1)
    a.close()

    b = fresh(s)
    ref = b.render(params(s, 0))
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    acc0 = b.accumulation(w, h)
    with pytest.raises(sa.StkError):          # holds a cache, b has none
        b.checkpoint_load(path)
    b.cache_create(log2_table=15, seed=3)
    with pytest.raises(sa.StkError):          # another table size
        b.checkpoint_load(path)
    raw = bytearray(path.read_bytes())
    raw[len(raw) // 2] ^= 0x40
    bad = tmp_path / "bad.stkckpt"
    bad.write_bytes(bytes(raw))
    with pytest.raises(sa.StkError):          # checksum
        b.checkpoint_load(bad)
    short = tmp_path / "short.stkckpt"
    short.write_bytes(path.read_bytes()[:-100])
    with pytest.raises(sa.StkError):          # truncated
        b.checkpoint_load(short)
    with pytest.raises(sa.StkError):
        b.checkpoint_load(tmp_path / "missing.stkckpt")
    # nothing changed: the accumulation continues from b's own first frame
    assert b.accumulation(w, h).tobytes() == acc0.tobytes()
    nxt = b.render(params(s, 1))
    b.close()
    c = fresh(s)
    c.render(params(s, 0))
    assert c.render(params(s, 1)).tobytes() == nxt.tobytes()
    c.close()
    assert np.isfinite(ref).all()


def test_trained_cache_resumes(tmp_path):
    print("case: cache checkpoint after 6 trained frames, config 6, log2 table 16", flush=True)
    s = sa.SceneData.generate(6)
    path = tmp_path / "cache.stkckpt"
    cfg = dict(log2_table=16, seed=5, ema=0.99)
    train = dict(cache_mode=1, cache_query_depth=1, cache_train_paths=8192)
    a = fresh(s)
    a.cache_create(**cfg)
    for f in range(6):
        a.render(params(s, f, **train))
    a.checkpoint_save(path, user_frame=6)
    step = a.cache_info().step
    grid_a, mlp_a = a.cache_get_params()

    b = fresh(s)
    b.cache_create(**cfg)
    info = b.checkpoint_load(path)
    assert info["cache"] and info["accumulation"] and info["cache_step"] == step > 0
    assert b.cache_info().step == step
    grid_b, mlp_b = b.cache_get_params()
    assert grid_b.tobytes() == grid_a.tobytes() and mlp_b.tobytes() == mlp_a.tobytes()

    # queries read the moving averages: a frozen-cache frame is the same bytes
    frozen = params(s, 6, cache_mode=2, cache_query_depth=1)
    assert b.render(frozen).tobytes() == a.render(frozen).tobytes()

    # one optimiser step on the same records: restored moments and step give
    # the continuing run's update (up to the order of the gradient atomics);
    # the weights alone (fresh moments, step 0) give a different one
    recs = records(a.cache_info(), 16384, 11)
    a.cache_train_step(recs)
    b.cache_train_step(recs)
    c = fresh(s)
    c.cache_create(**cfg)
    c.cache_set_params(grid_a, mlp_a)
    c.cache_train_step(recs)
    ga, ma = a.cache_get_params()
    gb, mb = b.cache_get_params()
    gc, mc = c.cache_get_params()
    d_resumed = max(np.abs(ga - gb).max(), np.abs(ma - mb).max())
    d_weights_only = max(np.abs(ga - gc).max(), np.abs(ma - mc).max())
    print(f"after one step: resumed differs by {d_resumed:.3e}, weights-only by {d_weights_only:.3e}")
    assert d_resumed < 1e-6
    assert d_weights_only > 1e-4
    for x in (a, b, c):
        x.close()
