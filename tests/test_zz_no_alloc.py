"""stk_trace_closest_noalloc (the entry point that allocates no device memory:
rays and hits staged through the library's static device buffer) against
stk_trace_closest, through the C ABI: bit-identical hits on two calls in a
row, the second with fewer rays, and a clear refusal above the buffer's
1,048,576 rays."""
import ctypes as C

import numpy as np
import pytest

import strelka_amd as sa

pytestmark = pytest.mark.gpu

MAX_RAYS = 1 << 20


def call_noalloc(ctx, rays, flags=0):
    hits = np.zeros(rays.size, dtype=sa.HIT)
    st = sa.lib().stk_trace_closest_noalloc(ctx.handle, C.c_void_p(rays.ctypes.data), rays.size,
                                            C.c_void_p(hits.ctypes.data), flags)
    return st, hits


def test_noalloc_trace_equals_allocating_trace():
    print("case: config 2 (1.16M triangles), noalloc closest hit: 1,048,576 rays then 65,536 rays", flush=True)
    with sa.Context(0) as ctx:
        s = sa.SceneData.generate(2)
        ctx.set_scene(s)
        ctx.build_bvh()
        rays = s.jittered_rays(0x5354524B00000002)
        first = np.ascontiguousarray(rays[:MAX_RAYS])
        second = np.ascontiguousarray(rays[MAX_RAYS + 12345:MAX_RAYS + 12345 + 65536])
        for batch in (first, second):
            st, got = call_noalloc(ctx, batch)
            assert st == 0, sa.lib().stk_last_error(ctx.handle)
            exp = ctx.trace_closest(batch)
            assert got.tobytes() == exp.tobytes()
            assert np.mean(got["instId"] != 0xFFFFFFFF) > 0.99
        # the reference-order walk through the same buffer
        st, got = call_noalloc(ctx, second[:4096].copy(), flags=1)
        assert st == 0 and got.tobytes() == ctx.trace_closest(second[:4096], reference_order=True).tobytes()


def test_noalloc_trace_refuses_more_rays_than_the_buffer_holds():
    print("case: noalloc closest hit, 1,048,577 rays refused, then a small call still works", flush=True)
    with sa.Context(0) as ctx:
        s = sa.SceneData.generate(1)
        ctx.set_scene(s)
        ctx.build_bvh()
        rays = np.zeros(MAX_RAYS + 1, dtype=sa.RAY)
        st, _ = call_noalloc(ctx, rays)
        assert st == 1
        assert b"1048576" in sa.lib().stk_last_error(ctx.handle)
        small = s.jittered_rays(0x5354524B00000001)[:1000].copy()
        st, got = call_noalloc(ctx, small)
        assert st == 0 and got.tobytes() == ctx.trace_closest(small).tobytes()
        st, _ = call_noalloc(ctx, np.zeros(0, dtype=sa.RAY))
        assert st == 0
