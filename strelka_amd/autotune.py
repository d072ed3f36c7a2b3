"""Launch-shape autotuner of the wavefront path tracer.

The reference dispatches its compute passes with one fixed 16x16 workgroup
(src/render/pathtracer.cpp); here the shape of the wavefront launches is a
per-device, per-workload choice: threads per block of the wavefront kernels,
the raygen tile edge (spatial coherence of a wave's rays), a persistent grid
for the extend kernel, the traversal (compressed 8-wide or packed binary
BVH), a cap on the threads per queue stripe of the consumer grids (fewer
blocks that loop, for short late-bounce queues), the number of paths traced
per pass (so a pass's path state fits the last-level cache), the number of
pass streams (overlapping a pass's last bounces with the next passes), graph
replay of the pass, and the persistent frame launch (one launch per frame
whose resident workgroups pull 256-slot chunks and run all of a chunk's
samples; its size in workgroups per CU).  None of them
changes a single output bit (the launch-shape tests check that), so the tuner
only measures time: it runs the caller's workload under each candidate with
device events on the context stream and keeps the fastest.

Search: coordinate descent from the default shape, one knob at a time, two
sweeps; every measured shape is kept in the returned table.  Choices can be
stored in a JSON file keyed by device and workload name and reused.
"""
from __future__ import annotations

import json
from pathlib import Path
from typing import Callable

BLOCK_SIZES = (64, 128, 256)
TILES = (1, 4, 8, 16)
PERSISTENT_PER_CU = (0, 4, 8, 16)
PATHS_IN_FLIGHT = (0, 1 << 20, 1 << 22)
STRIPE_THREADS_PER_CU = (0, 8, 16, 32)  # x CUs threads per queue stripe (0: one per entry)
FRAME_BLOCKS_PER_CU = (0, 4, 8, 16)  # x CUs workgroups of the persistent frame launch (0: per-stage launches)
FRAME_CHUNKS = (256, 1024, 4096)     # slots per chunk of that launch

DEFAULT = {"block_size": 256, "tile": 8, "persistent_blocks": 0, "graph": True, "paths_in_flight": 0,
           "binary_walk": False, "stripe_threads": 0, "pass_streams": 1, "frame_blocks": 0,
           "frame_chunk": 256}


def _key(shape: dict) -> tuple:
    return tuple(sorted(shape.items()))


def measure(ctx, run: Callable[[], None], shape: dict, repeats: int = 2) -> float:
    """Milliseconds per call of run() under a launch shape (one untimed call
    first: it builds buffers and captures the graph)."""
    ctx.set_launch_config(**shape)
    run()
    ctx.sync()
    ctx.record(14)
    for _ in range(repeats):
        run()
    ctx.record(15)
    ctx.sync()
    return ctx.elapsed_ms(14, 15) / repeats


def tune(ctx, run: Callable[[], None], total_paths: int, repeats: int = 2, sweeps: int = 2,
         log: Callable[[str], None] | None = None) -> dict:
    """Returns {"best": shape, "ms": time, "table": [{shape, ms}, ...]} and
    leaves the context on the best shape."""
    cus = ctx.device_info()["compute_units"]
    axes = {
        "block_size": list(BLOCK_SIZES),
        "tile": list(TILES),
        "persistent_blocks": [k * cus for k in PERSISTENT_PER_CU],
        "paths_in_flight": [q for q in PATHS_IN_FLIGHT if q == 0 or q < total_paths],
        "graph": [True, False],
        "binary_walk": [False, True],
        "stripe_threads": [k * cus for k in STRIPE_THREADS_PER_CU],
        "pass_streams": [1, 2, 3, 4],
        "frame_blocks": [k * cus for k in FRAME_BLOCKS_PER_CU],
        "frame_chunk": list(FRAME_CHUNKS),
    }
    seen: dict = {}

    def cost(shape: dict) -> float:
        k = _key(shape)
        if k not in seen:
            seen[k] = measure(ctx, run, shape, repeats)
            if log:
                log(f"autotune {shape}: {seen[k]:.3f} ms")
        return seen[k]

    best = dict(DEFAULT)
    best_ms = default_ms = cost(best)
    for _ in range(sweeps):
        improved = False
        for axis, values in axes.items():
            for v in values:
                cand = dict(best)
                cand[axis] = v
                ms = cost(cand)
                if ms < best_ms * 0.995:
                    best, best_ms, improved = cand, ms, True
        if not improved:
            break
    ctx.set_launch_config(**best)
    table = [{"shape": dict(k), "ms": round(v, 4)} for k, v in sorted(seen.items(), key=lambda kv: kv[1])]
    return {"best": best, "ms": round(best_ms, 4), "default_ms": round(default_ms, 4), "table": table,
            "compute_units": cus}


def load_choice(path: Path, device: str, workload: str) -> dict | None:
    try:
        data = json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return None
    entry = data.get(f"{device}|{workload}")
    return entry.get("best") if isinstance(entry, dict) else None


def store_choice(path: Path, device: str, workload: str, result: dict) -> None:
    p = Path(path)
    try:
        data = json.loads(p.read_text())
    except (OSError, ValueError):
        data = {}
    data[f"{device}|{workload}"] = {"best": result["best"], "ms": result["ms"]}
    p.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")
