"""Writes the committed golden vectors (tests/golden/golden.npz + golden.json).

The reference ships no vectors for this path beyond the disabled box test
(tests/intersection/aabb.cpp, checked in test_oracle_primitives.py), and its
shaders cannot run here, so these vectors are produced by the CPU oracle
(oracle/oracle.c) on the seeded synthetic workloads, after the oracle itself
is pinned by the closed-form and brute-force tests.  They freeze:

  * rays + hits: 4,096 config-1 jittered primary rays (every 16th of the
    65,536, PCG seed 0x5354524B00000001) and their reference-order closest
    hits; the BVH is this library's deterministic SAH build;
  * rng: xorshift32 streams (16 draws) seeded by initRNG for 8 (pixel, frame)
    pairs, shaders/random.h;
  * pt: per-sample radiance of a 16x16 crop at (952, 532) of the 1080p frame of
    the 1M-triangle scene, 4 spp, frame 0, 8 bounces, random jitter, its sky;
  * tonemap: 64 linear colours through the three tonemappers
    (shaders/tonemap.hlsl).

Run from the repository root:  python tests/golden/make_golden.py
"""
import ctypes as C
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import strelka_amd as sa  # noqa: E402
from oracle import Oracle, load  # noqa: E402
from oracle.pyoracle import MODE_REF, fptr  # noqa: E402

OUT = Path(__file__).resolve().parent
SEED1 = 0x5354524B00000001


def rng_streams(lib):
    cases = [(0, 0, 256, 0), (17, 3, 256, 0), (255, 255, 256, 7), (952, 532, 1920, 0), (1919, 1079, 1920, 63),
             (5, 9, 64, 1000), (100, 0, 960, 123456), (0, 1, 1, 4294967295)]
    out = np.zeros((len(cases), 17), np.uint32)
    for i, (px, py, w, frame) in enumerate(cases):
        st = C.c_uint32(lib.orc_init_rng(px, py, w, frame))
        out[i, 0] = st.value
        for k in range(16):
            out[i, k + 1] = lib.orc_xorshift(C.byref(st))
    return np.array(cases, np.uint64), out


def main():
    lib = load()
    s1 = sa.SceneData.generate(1)
    b1 = sa.HostBvh.build(s1)
    rays = s1.jittered_rays(SEED1)[::16].copy()
    hits, _ = Oracle.from_scene(s1, b1.nodes).trace(rays, MODE_REF)

    rng_cases, rng = rng_streams(lib)

    s2 = sa.SceneData.generate(2)
    b2 = sa.HostBvh.build(s2)
    p = sa.render_params(s2.camera, spp=4, max_depth=8, frame=0, sampling=1, sky=s2.sky)
    pt = Oracle.from_scene(s2, b2.nodes).pt_frame(p, 952, 532, 16, 16)

    rng_np = np.random.default_rng(2026)
    colours = np.concatenate([rng_np.uniform(0, 4, (60, 4)), [[0, 0, 0, 1], [1, 1, 1, 1], [100, 0.5, 0, 1],
                                                            [0.18, 0.18, 0.18, 1]]]).astype(np.float32)
    tm = np.zeros((3, 64, 4), np.float32)
    for t in range(3):
        src = np.ascontiguousarray(colours)
        lib.orc_tonemap(fptr(src), 64, t, fptr(tm[t]))

    np.savez_compressed(OUT / "golden.npz", rays=rays.view(np.uint8), hits=hits.view(np.uint8),
                        rng_cases=rng_cases, rng=rng, pt=pt, tonemap_in=colours, tonemap=tm)
    meta = {"rays": {"config": 1, "seed": hex(SEED1), "stride": 16, "count": int(rays.size),
                     "bvh_nodes": int(b1.nodes.size)},
            "pt": {"config": 2, "crop": [952, 532, 16, 16], "spp": 4, "frame": 0, "max_depth": 8, "sampling": 1,
                   "sky": list(s2.sky[:3])},
            "tonemap": {"types": ["reinhard", "aces_fitted", "aces_film"]}}
    (OUT / "golden.json").write_text(json.dumps(meta, indent=1) + "\n")
    print("wrote", OUT / "golden.npz")


if __name__ == "__main__":
    main()
