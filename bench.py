"""Benchmark: closest-hit throughput of the MI355X traversal on the 1M-triangle
scene (1920x1080 jittered primary rays), plus the path tracer's samples/s.

A step is one closest-hit launch over the 2,073,600 primary rays of the seeded
heightfield + 2,000-sphere scene (1,164,290 instanced triangles), rays and BVH
resident in HBM.  Prints one JSON line.  Only one GPU is used: --gpus above 1
exits with a message.

Other modes (one JSON line each): --workload batch (32 scenes of 10k-2M
triangles at 960x540, 16 spp, one batched wavefront pass per sample, launch
shape from the autotuner), --workload instanced (10,000 instances of 20
meshes, about 19 M triangles, 3840x2160, 4 spp, 6 bounces), --cache-study (time to an RMSE target, cached vs
plain path tracing) and --copy-mib (plain copy kernel, calibrates the memory
counters).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s closest-hit; sec to RMSE<=0.01 vs converged ref, cached vs plain PT"
HBM_PEAK_GBS = 8000.0


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


# Host threads of the CPU baseline: every core this process may run on, up to
# the host-CPU share of one GPU on the GPU boxes (16 threads); nproc and the
# affinity mask are reported beside it.
CPU_SHARE_PER_GPU = 16


def cpu_threads() -> dict:
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "affinity": allowed, "threads": max(1, min(allowed, CPU_SHARE_PER_GPU))}


def cpu_baseline(scene, nodes, rays, wide, wide_tris):
    """The oracle's closest-hit on the host cores over the same rays and BVH:
    the reference order, the binary kernel's pruned order and the wide
    kernel's walk (whose counters give that kernel's algorithmic bytes).
    Timed on a -march=native build of the oracle made on this host, after
    checking that its hits equal the portable build's."""
    from oracle import Oracle
    from oracle.pyoracle import MODE_PRUNED, MODE_REF, MODE_WIDE

    threads = cpu_threads()["threads"]
    o = Oracle.from_scene(scene, nodes, wide, wide_tris, native=True)
    portable = Oracle.from_scene(scene, nodes, wide, wide_tris)
    check = rays[::97].copy()
    if o.trace(check, MODE_REF, threads)[0].tobytes() != portable.trace(check, MODE_REF, threads)[0].tobytes():
        raise RuntimeError("the -march=native oracle build gives other hits than the portable build")
    res = {}
    for name, mode in (("reference_order", MODE_REF), ("pruned", MODE_PRUNED), ("wide", MODE_WIDE)):
        o.trace(rays[:4096], mode, threads)  # warm-up
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            _, st = o.trace(rays, mode, threads)
            times.append(time.perf_counter() - t0)
        t = float(np.median(times))
        res[name] = {"mrays_per_s": rays.size / t / 1e6, "inner_per_ray": st.inner / st.rays,
                     "leaf_per_ray": st.leaf / st.rays, "seconds": t}
    return threads, res


def cpu_pt_baseline(scene, nodes, params, rows: int = 64):
    """The oracle's path tracer on the host cores over a band of the same
    frame (full width, `rows` rows from the middle, the frame's spp)."""
    from oracle import Oracle

    threads = cpu_threads()["threads"]
    o = Oracle.from_scene(scene, nodes, native=True)
    w, h = params.pt.dimension[0], params.pt.dimension[1]
    y0 = max(0, h // 2 - rows // 2)
    t0 = time.perf_counter()
    o.pt_frame(params, 0, y0, w, rows, threads)
    t = time.perf_counter() - t0
    n = w * rows * params.pt.spp
    return {"value": round(n / t / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"{w}x{rows} band of the frame at {params.pt.spp} spp ({n} samples), oracle path tracer "
                      f"(-march=native build)", "cpu_model": cpu_model(), "seconds": round(t, 2)}


def pt_frame_bench(ctx, scene, spp: int, frames: int, d_out: int) -> dict:
    """Frames of `spp` samples per pixel (8 bounces, RR after 3) of the scene's
    view, timed with device events; rays counted by the kernels."""
    import strelka_amd as sa

    w, h = scene.camera.dimension[0], scene.camera.dimension[1]
    p = sa.render_params(scene.camera, spp=spp, max_depth=8, frame=0, sampling=1, sky=scene.sky)
    ctx.render_device(p, d_out)  # warm-up: buffers, graph capture
    ctx.sync()
    rays = [0, 0]
    ctx.record(6)
    for f in range(frames):
        p.pt.frameNumber = f + 1
        ctx.render_device(p, d_out)
    ctx.record(7)
    ctx.sync()
    ms = ctx.elapsed_ms(6, 7) / frames
    rays = ctx.ray_totals()
    return {"frame_ms": round(ms, 3), "spp": spp, "resolution": [w, h],
            "samples_per_s": round(w * h * spp / (ms * 1e-3), 1),
            "mrays_per_s": round((rays[0] + rays[1]) / (ms * 1e-3) / 1e6, 1),
            "closest_rays_per_frame": rays[0], "shadow_rays_per_frame": rays[1]}


def with_graph_state(shape: dict) -> dict:
    """The launch shape plus whether a captured graph is replayed: the library
    replays one only when every pass runs on one stream (pass_streams 1), the
    frame is not one persistent launch (frame_blocks 0) and the radiance cache
    is off."""
    out = dict(shape)
    out["graph_replayed"] = (bool(shape.get("graph")) and int(shape.get("pass_streams", 1)) <= 1
                             and not shape.get("frame_blocks"))
    return out


def choose_launch(ctx, spec: str, run, total_paths: int) -> dict:
    """The path tracer's launch shape: the autotuner's choice on `run`, or a
    fixed shape given as JSON."""
    from strelka_amd import autotune

    if spec == "auto":
        return autotune.tune(ctx, run, total_paths=total_paths, log=lambda m: print(m, file=sys.stderr, flush=True))
    shape = dict(autotune.DEFAULT, **json.loads(spec))
    ctx.set_launch_config(**shape)
    return {"best": shape, "table": [], "ms": None, "default_ms": None}


def batch_bench(args) -> dict:
    """The batched workload: 32 seeded scenes of uneven size (about 10k to 2M
    instanced triangles, log-uniform), each viewed at 960x540, 16 spp, 8
    bounces, all views traced by the same wavefront launches; the launch shape
    is chosen by the autotuner on a 4-spp frame first."""
    import strelka_amd as sa

    t0 = time.perf_counter()
    scenes = [sa.SceneData.generate(3, v) for v in range(32)]
    gen_s = time.perf_counter() - t0
    ctx = sa.Context(0)
    ctx.batch_clear()
    for s in scenes:
        ctx.batch_add(s)
    t0 = time.perf_counter()
    ctx.batch_commit()
    build_s = time.perf_counter() - t0
    views = ctx.batch_views()
    npix = ctx.batch_pixels()
    tris = [int(s.inst["indexCount"].sum()) // 3 for s in scenes]
    print(f"batch: 32 scenes, {sum(tris)} instanced triangles, BVHs built in {build_s:.1f} s",
          file=sys.stderr, flush=True)
    d_out = ctx.alloc(npix * 16)
    p = sa.render_params(scenes[0].camera, spp=4, max_depth=8, frame=0, sampling=1)
    tuned = choose_launch(ctx, args.launch, lambda: ctx.batch_render_device(p, d_out), npix)
    spp = 16
    p = sa.render_params(scenes[0].camera, spp=spp, max_depth=8, frame=1, sampling=1)
    ctx.batch_render_device(p, d_out)
    ctx.sync()
    ctx.record(8)
    for f in range(args.steps):
        p.pt.frameNumber = 2 + f
        ctx.batch_render_device(p, d_out)
    ctx.record(9)
    ctx.sync()
    ms = ctx.elapsed_ms(8, 9) / args.steps
    rays = ctx.ray_totals()
    dev = ctx.device_info()

    # the fastest per-stage shape against the fastest persistent frame launch
    # (one launch per frame, workgroups pulling 256-slot chunks), each timed
    # on the same 16-spp frames
    def timed(shape):
        ctx.set_launch_config(**shape)
        ctx.batch_render_device(p, d_out)
        ctx.sync()
        ctx.record(8)
        for f in range(args.steps):
            p.pt.frameNumber = 2 + f
            ctx.batch_render_device(p, d_out)
        ctx.record(9)
        ctx.sync()
        return round(ctx.elapsed_ms(8, 9) / args.steps, 3)

    table = [r["shape"] for r in tuned["table"]]
    per_stage = next((t for t in table if not t.get("frame_blocks")), dict(tuned["best"], frame_blocks=0))
    # the frame launch's own knobs (workgroups, slots per chunk), swept on the
    # autotuner's 4-spp frame
    from strelka_amd import autotune
    p4 = sa.render_params(scenes[0].camera, spp=4, max_depth=8, frame=0, sampling=1)
    frame_sweep = []
    for bs, per_cu in ((64, 8), (64, 16), (256, 2), (256, 4)):
        for chunk in (256, 1024, 4096):
            shape = dict(per_stage, block_size=bs, frame_blocks=per_cu * dev["compute_units"], frame_chunk=chunk)
            ms4 = autotune.measure(ctx, lambda: ctx.batch_render_device(p4, d_out), shape)
            frame_sweep.append({"shape": shape, "ms_4spp": round(ms4, 3)})
            print(f"frame launch {shape}: {ms4:.3f} ms at 4 spp", file=sys.stderr, flush=True)
    frame = min(frame_sweep, key=lambda r: r["ms_4spp"])["shape"]
    comparison = {"per_stage": {"shape": per_stage, "ms": timed(per_stage)},
                  "frame_launch": {"shape": frame, "ms": timed(frame)}, "frame_sweep_4spp": frame_sweep}
    print(f"launch comparison: {comparison}", file=sys.stderr, flush=True)
    ctx.set_launch_config(**tuned["best"])
    ctx.free(d_out)
    ctx.close()
    samples = npix * spp
    return {"metric": "samples/s, batched path tracing of 32 uneven scenes", "value": round(samples / (ms * 1e-3), 1),
            "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": round(ms, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (32 seeded PCG scenes)",
            "config": {"workload": "32 seeded scenes of 10k-2M instanced triangles, 960x540 each, 16 spp, 8 bounces, "
                                   "one batched wavefront pass per sample", "scenes": 32, "pixels": npix, "spp": spp,
                       "instanced_triangles": tris, "bvh_build_s": round(build_s, 2),
                       "scene_gen_s": round(gen_s, 2)},
            "mrays_per_s": round((rays[0] + rays[1]) / (ms * 1e-3) / 1e6, 1),
            "rays_per_frame": {"closest": rays[0], "shadow": rays[1]},
            "launch_comparison_16spp": comparison,
            "autotune": {"best": with_graph_state(tuned["best"]), "best_4spp_ms": tuned["ms"], "default_4spp_ms": tuned["default_ms"],
                         "table": tuned["table"],
                         "device": dev["name"], "compute_units": dev["compute_units"]}}


def instanced_bench(args) -> dict:
    """The large instanced workload: 10,000 instances of 20 seeded meshes
    (about 19 M triangles), 3840x2160, 4 spp per frame, 6 bounces, RR after 3,
    traced through the two-level structure (a tree per mesh under a tree of
    instances) and, beside it, through the flat tree of all triangles; launch
    shape from the autotuner on a 4-spp frame for each."""
    import strelka_amd as sa

    scene = sa.SceneData.generate(5)
    ntri = int(scene.inst["indexCount"].sum()) // 3
    ctx = sa.Context(0)
    ctx.set_scene(scene)
    w, h = scene.camera.dimension[0], scene.camera.dimension[1]
    d_out = ctx.alloc(w * h * 16)
    cam_rays = scene.jittered_rays(0x5354524B00000005)
    d_rays = ctx.alloc(cam_rays.size * 32)
    d_hits = ctx.alloc(cam_rays.size * 20)
    ctx.upload(d_rays, cam_rays)
    spp = 4

    def measure(structure: str) -> dict:
        method = {"flat_device": "lbvh", "flat_device_ploc": "ploc"}.get(structure)
        if method:
            ctx.build_bvh_gpu(method)  # first call loads the build kernels
            ctx.sync()
        t0 = time.perf_counter()
        if structure in ("two_level", "two_level_world"):
            ctx.build_two_level()
            if structure == "two_level_world":
                ctx.two_level_world_triangles(True)
                ctx.sync()
        elif method:
            ctx.build_bvh_gpu(method)
            ctx.sync()
        else:
            ctx.build_bvh()
        build_s = time.perf_counter() - t0
        print(f"instanced ({structure}): {ntri} triangles, built in {build_s:.2f} s", file=sys.stderr, flush=True)
        # primary-ray closest hit
        ctx.trace_closest_device(d_rays, cam_rays.size, d_hits)
        ctx.sync()
        ctx.record(6)
        for _ in range(5):
            ctx.trace_closest_device(d_rays, cam_rays.size, d_hits)
        ctx.record(7)
        ctx.sync()
        trace_ms = ctx.elapsed_ms(6, 7) / 5
        # the same rays, one wave per 8x8 tile of the frame
        assert cam_rays.size == w * h
        ctx.trace_closest_image_device(d_rays, w, h, d_hits)
        ctx.record(6)
        for _ in range(5):
            ctx.trace_closest_image_device(d_rays, w, h, d_hits)
        ctx.record(7)
        ctx.sync()
        tiles_ms = ctx.elapsed_ms(6, 7) / 5
        p1 = sa.render_params(scene.camera, spp=4, max_depth=6, frame=0, sampling=1, sky=scene.sky)
        tuned = choose_launch(ctx, args.launch, lambda: ctx.render_device(p1, d_out), w * h)
        p = sa.render_params(scene.camera, spp=spp, max_depth=6, frame=1, sampling=1, sky=scene.sky)
        ctx.render_device(p, d_out)
        ctx.sync()
        ctx.record(8)
        for f in range(args.steps):
            p.pt.frameNumber = 2 + f
            ctx.render_device(p, d_out)
        ctx.record(9)
        ctx.sync()
        ms = ctx.elapsed_ms(8, 9) / args.steps
        rays = ctx.ray_totals()
        return {"build_s": round(build_s, 3), "frame_ms": round(ms, 3),
                "samples_per_s": round(w * h * spp / (ms * 1e-3), 1),
                "mrays_per_s": round((rays[0] + rays[1]) / (ms * 1e-3) / 1e6, 1),
                "rays_per_frame": {"closest": rays[0], "shadow": rays[1]},
                "primary_closest_mrays_per_s": round(cam_rays.size / (trace_ms * 1e-3) / 1e6, 1),
                "primary_kernel_ms": round(trace_ms, 4),
                "primary_tiles_mrays_per_s": round(cam_rays.size / (tiles_ms * 1e-3) / 1e6, 1), "launch": with_graph_state(tuned["best"]),
                "autotune_4spp_ms": {"default": tuned.get("default_ms"), "best": tuned.get("ms")}}

    want = set(args.structures.split(","))
    two = measure("two_level") if "two_level" in want else None
    two_world = measure("two_level_world") if "two_level_world" in want else None
    flat = measure("flat") if "flat" in want else None
    flat_dev = measure("flat_device") if "flat_device" in want else None
    flat_ploc = measure("flat_device_ploc") if "flat_device_ploc" in want else None
    head = next(r for r in (two, two_world, flat, flat_dev, flat_ploc) if r is not None)
    for b in (d_out, d_rays, d_hits):
        ctx.free(b)
    ctx.close()
    return {"metric": "samples/s, path tracing of the large instanced scene",
            "value": head["samples_per_s"], "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
            "warmup": 1, "ms_per_step": head["frame_ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded PCG meshes and transforms)",
            "config": {"workload": "10,000 instances of 20 seeded meshes under a tree of instances, 3840x2160, "
                                   "4 spp, 6 bounces", "triangles": ntri, "build_s": head["build_s"]},
            "two_level": two, "two_level_world_triangles": two_world, "flat": flat, "flat_device_built": flat_dev,
            "flat_device_clustered": flat_ploc}


def load_traffic(kernel: str):
    f = ROOT / "profiles" / f"{kernel}_traffic.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            return None
    return None


def load_counters(kernel: str) -> dict | None:
    """Per-launch counter figures of a kernel from its committed profile
    (profiles/<kernel>_counters.json, written from rocprofv3 --pmc runs)."""
    f = ROOT / "profiles" / f"{kernel}_counters.json"
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text())
    except (OSError, ValueError):
        return None


def limiter(counters: dict | None, traffic: float | None, kernel_ms: float) -> dict | None:
    """The resource that binds the kernel, from its counters: the VALU issue
    fraction (SIMD quad-cycles in which at least one VALU instruction issues,
    (SQ_ACTIVE_INST_VALU - SQ_ACTIVE_INST_VALU2) / (4 CUs x GRBM_GUI_ACTIVE per
    XCD / 4)), the measured HBM fraction (measured bytes per launch over this
    run's launch time and the 8 TB/s peak) and the L2 hit rate; the binding
    resource is the one nearest its limit."""
    if not counters:
        return None
    valu = counters.get("valu_issue_quad_cycle_frac")
    hbm = (traffic / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None
    cands = {k: v for k, v in (("valu_issue", valu), ("hbm", hbm)) if v is not None}
    resource = max(cands, key=cands.get) if cands else None
    return {"resource": resource, "frac": round(cands[resource], 4) if resource else None,
            "valu_busy": counters.get("valu_busy"), "valu_dual_issue_share": counters.get("valu_dual_issue_share"),
            "valu_thread_utilization": counters.get("valu_utilization"),
            "hbm_measured_frac": round(hbm, 4) if hbm is not None else None,
            "l2_hit_rate": counters.get("l2_hit_rate"), "l2_requests_per_ray": counters.get("l2_requests_per_ray"),
            "valu_instructions_per_ray": counters.get("valu_insts_per_ray"),
            "source": counters.get("source")}


def cache_study(args) -> dict:
    """Time to reach an RMSE target against a converged plain render, for the
    plain path tracer and the cached one (training included in its time).
    The 1.16M-triangle heightfield + spheres scene at the study resolution,
    1 spp per frame, 8 bounces, RR after 3, ACES-tonemapped RMSE.  A cached
    run starts from a fresh cache: `warmup` frames train on the training
    paths only (their time counts), then the accumulation restarts and every
    frame queries the cache at vertex depth `query_depth` and trains."""
    import strelka_amd as sa

    scene = sa.SceneData.generate(args.study_scene)
    cam = scene.camera
    W, H = args.study_res if args.study_res else (cam.dimension[0], cam.dimension[1])
    cam.dimension[0], cam.dimension[1] = W, H
    cam.invDimension[0], cam.invDimension[1] = 1.0 / W, 1.0 / H
    ctx = sa.Context(0)
    ctx.set_scene(scene)
    if args.study_structure == "two_level":
        ctx.build_two_level()
    elif args.study_structure == "device":
        ctx.build_bvh_gpu()
    else:
        ctx.build_bvh()
    # pass streams: a frame's bounces overlap the previous frame's tail (and,
    # cached, its inference and training, which every cached frame waits for
    # before its own inference)
    streams = getattr(args, "study_streams", 2)
    ctx.set_launch_config(block_size=64, tile=8, pass_streams=streams)
    npix = W * H
    default_paths = max(1024, int(65536 * npix / (1920 * 1080)))
    d_out = ctx.alloc(npix * 16)
    out = np.zeros((H, W, 4), np.float32)

    def params(frame, mode, qd=2, paths=default_paths, min_rough=0.0, steps=1, suffix=0, direct=False):
        return sa.render_params(cam, spp=args.study_spp, max_depth=args.study_depth, frame=frame, sampling=1,
                                sky=scene.sky, tonemap=1,
                                cache_mode=mode, cache_query_depth=qd, cache_train_paths=paths,
                                cache_min_roughness=min_rough, cache_train_steps=steps, cache_train_suffix=suffix,
                                cache_query_direct=direct)

    ctx.reset_accumulation()
    t0 = time.perf_counter()
    ref_frames = (args.study_ref_spp + args.study_spp - 1) // args.study_spp
    for f in range(ref_frames):
        ctx.render_device(params(1_000_000 + f, 0), d_out)
        if f % 64 == 63:
            ctx.sync()
            print(f"reference {(f + 1) * args.study_spp}/{ref_frames * args.study_spp} spp", file=sys.stderr,
                  flush=True)
    ctx.sync()
    ref_s = time.perf_counter() - t0
    ref_tm = ctx.download(d_out, out.copy())[..., :3]
    ref_lin = ctx.accumulation(W, H)[..., :3]

    # every 8 frames from 32 on: the time to the target is interpolated between
    # the two checkpoints that bracket it (sparse checkpoints quantised it by
    # up to 50%)
    checkpoints = sorted({1, 2, 4, 8, 12, 16, 24} | set(range(32, args.study_frames + 1, 8)) | {args.study_frames})

    def run(mode, qd=2, warmup=0, paths=default_paths, min_rough=0.0, steps=1, label="", suffix=0, every=1,
            direct=False):
        # frames are issued back to back and timed between checkpoints (the
        # image read-back at a checkpoint is not timed)
        t_ms = 0.0
        if warmup:
            ctx.sync()
            ctx.record(0)
            for f in range(warmup):
                ctx.render_device(params(500_000 + f, 4, qd, paths, min_rough, steps, 0, direct), d_out)
            ctx.record(1)
            t_ms += ctx.elapsed_ms(0, 1)
        ctx.reset_accumulation()
        curve = []
        ctx.sync()
        ctx.record(0)
        for f in range(args.study_frames):
            # every: a query-and-train frame every `every` frames, query-only frames between
            m = 2 if (mode == 1 and every > 1 and f % every) else mode
            ctx.render_device(params(f, m, qd, paths, min_rough, steps, suffix, direct), d_out)
            if f + 1 in checkpoints:
                ctx.record(1)
                t_ms += ctx.elapsed_ms(0, 1)
                tm = ctx.download(d_out, out.copy())[..., :3]
                lin = ctx.accumulation(W, H)[..., :3]
                curve.append({"frames": f + 1, "ms": round(t_ms, 2),
                              "rmse": round(float(np.sqrt(np.mean((tm - ref_tm) ** 2))), 5),
                              "rmse_linear": round(float(np.sqrt(np.mean((lin - ref_lin) ** 2))), 5)})
                print(f"{label} frame {f + 1}: {curve[-1]}", file=sys.stderr, flush=True)
                ctx.sync()
                ctx.record(0)
        return curve

    def time_to(curve, target):
        # first checkpoint at or below the target, the crossing interpolated
        # log-linearly in RMSE from the checkpoint before it
        for i, p in enumerate(curve):
            if p["rmse"] <= target:
                if i == 0 or curve[i - 1]["rmse"] <= p["rmse"]:
                    return p["ms"] / 1e3
                q = curve[i - 1]
                w = (math.log(q["rmse"]) - math.log(target)) / (math.log(q["rmse"]) - math.log(p["rmse"]))
                return (q["ms"] + w * (p["ms"] - q["ms"])) / 1e3
        return None

    plain = run(0, label="plain")
    tp = time_to(plain, args.study_target)
    variants = json.loads(args.study_variants)
    results = []
    for v in variants:
        ctx.cache_create(log2_table=v.get("log2_table", 19), base_res=16, max_res=v.get("max_res", 2048),
                         lr=v.get("lr", args.study_lr), ema=v.get("ema", args.study_ema),
                         lr_half_life=v.get("lr_half_life", 0.0), min_lr=v.get("min_lr", 0.0))
        curve = run(2 if v.get("freeze") else 1, v.get("query_depth", 2), v.get("warmup", 0),
                    v.get("train_paths", default_paths),
                    v.get("min_rough", 0.0), v.get("train_steps", 1), label=f"cached {v}",
                    suffix=v.get("train_suffix", 0), every=v.get("train_every", 1),
                    direct=bool(v.get("query_direct", 0)))
        tc = time_to(curve, args.study_target)
        results.append({"variant": v, "time_to_target_s": tc, "speedup": (tp / tc) if (tp and tc) else None,
                        "curve": curve})
    ctx.free(d_out)
    ctx.close()
    best = min((r for r in results if r["time_to_target_s"]), key=lambda r: r["time_to_target_s"], default=None)
    return {"study": "cache", "scene": args.study_scene, "structure": args.study_structure,
            "pass_streams": streams, "spp_per_frame": args.study_spp, "max_depth": args.study_depth, "resolution": [W, H], "target_rmse_tonemapped": args.study_target,
            "reference_spp": args.study_ref_spp, "reference_s": round(ref_s, 2),
            "default_train_paths_per_frame": default_paths,
            "time_to_target_s": {"plain": tp, "cached_best": best["time_to_target_s"] if best else None},
            "speedup": best["speedup"] if best else None, "plain": plain, "cached": results}


# The cached estimator the default run times against plain path tracing: the
# cache answers at the second path vertex (query depth 1) after 8 training-only
# frames, one optimiser step per frame over 4,096 training paths that end one
# vertex past the query depth with a query of their own (self-training),
# learning rate 1e-2 halving every 128 steps down to 2e-4, a hash grid of 2^16
# entries per level up to 256 cells across (a cache that stands in only after
# the second vertex needs no finer one, and a smaller table makes inference
# and the optimiser step cheaper).
TIME_TO_RMSE_VARIANT = {"query_depth": 1, "warmup": 8, "train_steps": 1, "train_paths": 4096, "train_suffix": 1,
                        "lr_half_life": 128, "min_lr": 0.0002, "log2_table": 16, "max_res": 256}


# Spatial splits of the headline tree: where an object split's children overlap
# by more than this fraction of the root's area (Stich et al. 2009).
SPATIAL_SPLIT_ALPHA = 1e-5


def time_to_rmse_leg() -> dict:
    """Bounded time-to-RMSE measurement for the default run: the 1M-triangle
    scene at 960x540, 1 spp per frame, 8 bounces, up to 256 frames each of the
    plain and the cached path tracer, ACES-tonemapped RMSE against a 2,048-spp
    plain render (the same protocol as --cache-study, one cached variant)."""
    ns = argparse.Namespace(study_scene=2, study_res=(960, 540), study_structure="flat", study_spp=1, study_depth=8,
                            study_ref_spp=2048, study_frames=256, study_target=0.01, study_lr=1e-2,
                            study_ema=0.99, study_variants=json.dumps([TIME_TO_RMSE_VARIANT]))
    r = cache_study(ns)
    cached = r["cached"][0]
    return {"plain_s": r["time_to_target_s"]["plain"], "cached_s": cached["time_to_target_s"],
            "speedup": cached["speedup"], "floor": cached["curve"][-1]["rmse"],
            "plain_rmse_at_end": r["plain"][-1]["rmse"], "frames": ns.study_frames, "target": ns.study_target,
            "resolution": list(ns.study_res), "reference_spp": ns.study_ref_spp, "variant": TIME_TO_RMSE_VARIANT,
            "estimator": "reference (refCompat 1) for both"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--settle-ms", type=float, default=300.0,
                    help="untimed closest-hit launches for this long before the warm-up steps (clock ramp)")
    ap.add_argument("--pt-frames", type=int, default=2)
    ap.add_argument("--pt-spp", type=int, default=64)
    ap.add_argument("--launch", default="auto",
                    help="launch shape of the path tracer: 'auto' (autotuner) or a JSON object of "
                         "Context.set_launch_config arguments")
    ap.add_argument("--workload", choices=("trace", "batch", "instanced"), default="trace",
                    help="trace: closest-hit on the 1M-triangle scene (+ its 64-spp path-traced frame); "
                         "batch: the 32-scene batched path tracer; instanced: the 19M-triangle instanced scene")
    ap.add_argument("--structures", default="two_level,two_level_world,flat,flat_device,flat_device_ploc",
                    help="instanced workload: the structures to measure (the first one gives value)")
    ap.add_argument("--no-time-to-rmse", action="store_true",
                    help="skip the time-to-RMSE leg (cached against plain path tracing) of the default run")
    ap.add_argument("--ray-order", choices=("tiles", "linear"), default="tiles",
                    help="headline closest-hit launches: the image entry point (one wave per 8x8 tile of "
                         "the row-major frame) or the flat entry point (64 consecutive rays per wave)")
    ap.add_argument("--trace-only", action="store_true",
                    help="only the timed closest-hit launches (for counter profiles of that kernel)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run the plain copy kernel over this many MiB (calibrates memory counters)")
    ap.add_argument("--cache-study", action="store_true",
                    help="only run the time-to-RMSE study of the cached vs plain path tracer")
    ap.add_argument("--study-res", type=int, nargs=2, default=None,
                    help="resolution of the study (default: 960x540 for scene 2, else the scene's own)")
    ap.add_argument("--study-scene", type=int, default=2,
                    help="seeded scene of the study: 2 (the 1M-triangle heightfield + spheres, open sky) or 7 "
                         "(a closed room lit by a small ceiling light)")
    ap.add_argument("--study-frames", type=int, default=256)
    ap.add_argument("--study-spp", type=int, default=1, help="samples per pixel per frame of the study")
    ap.add_argument("--study-depth", type=int, default=8, help="bounces of the study's paths")
    ap.add_argument("--study-streams", type=int, default=2, help="pass streams of both estimators in the study")
    ap.add_argument("--study-structure", choices=("flat", "two_level", "device"), default="flat",
                    help="acceleration structure of the study's scene")
    ap.add_argument("--study-ref-spp", type=int, default=2048)
    ap.add_argument("--study-target", type=float, default=0.01)
    ap.add_argument("--study-lr", type=float, default=1e-2)
    ap.add_argument("--study-ema", type=float, default=0.99)
    ap.add_argument("--hw-queues", type=int, default=8,
                    help="hardware queues HIP gives this process (GPU_MAX_HW_QUEUES, raised to this value before "
                         "HIP starts): the context stream plus up to four pass streams")
    ap.add_argument("--study-variants",
                    default='[{"query_depth": 0, "warmup": 64, "train_steps": 2, "train_paths": 16384, '
                            '"lr_half_life": 128, "min_lr": 0.0002}]',
                    help="JSON list of cached-run settings: query_depth, min_rough, train_steps, warmup, freeze (no "
                         "training after the warm-up), lr, lr_half_life (Adam steps), min_lr, ema, train_paths, "
                         "train_suffix (training paths end with a query that many vertices past the query depth), "
                         "train_every (query-only frames between training frames), query_direct (queries "
                         "cover their vertex's direct light), log2_table, max_res")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        return 2
    # HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 by
    # default); streams beyond that share queues and lose their overlap
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < args.hw_queues:
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)

    import strelka_amd as sa

    if args.cache_study:
        if args.study_res is None and args.study_scene == 2:
            args.study_res = (960, 540)
        print(json.dumps(cache_study(args)))
        return 0

    if args.workload == "batch":
        print(json.dumps(batch_bench(args)))
        return 0
    if args.workload == "instanced":
        print(json.dumps(instanced_bench(args)))
        return 0

    if args.copy_mib:
        ctx = sa.Context(0)
        nbytes = args.copy_mib << 20
        src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
        for _ in range(args.warmup):
            ctx.copy_device(dst, src, nbytes)
        ctx.record(0)
        for _ in range(args.steps):
            ctx.copy_device(dst, src, nbytes)
        ctx.record(1)
        ms = ctx.elapsed_ms(0, 1) / args.steps
        print(json.dumps({"copy_bytes": nbytes, "kernel_ms": round(ms, 4),
                          "read_plus_write_GBs": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}))
        ctx.free(src)
        ctx.free(dst)
        ctx.close()
        return 0

    scene = sa.SceneData.generate(2)
    ntri = int(scene.inst["indexCount"].sum()) // 3
    rays = scene.jittered_rays(0x5354524B00000002)
    n = rays.size
    ctx = sa.Context(0)
    t0 = time.perf_counter()
    ctx.set_scene(scene)
    # binned SAH with spatial splits, as the reference asks of Embree
    ctx.build_bvh(spatial_split_alpha=SPATIAL_SPLIT_ALPHA)
    build_s = time.perf_counter() - t0
    nodes = ctx.get_bvh()
    d_rays = ctx.alloc(n * 32)
    d_hits = ctx.alloc(n * 20)
    ctx.upload(d_rays, rays)

    # the rays are the frame's, row-major: the image entry point walks them
    # one 8x8 tile per wave; the flat one 64 consecutive rays (one row
    # segment) per wave.  Both write every hit at its ray's index, identical.
    width, height = scene.camera.dimension[0], scene.camera.dimension[1]
    assert width * height == n

    def trace_tiles():
        ctx.trace_closest_image_device(d_rays, width, height, d_hits)

    def trace_linear():
        ctx.trace_closest_device(d_rays, n, d_hits)

    trace = trace_tiles if args.ray_order == "tiles" else trace_linear

    # settle: the GPU reaches its sustained clocks only after some tens of ms
    # of load, so the launches run untimed for settle_ms before the warm-up
    # steps (the first launches after a few warm-ups ran 2-5% slower)
    settle_launches = 0
    t_settle = time.perf_counter()
    while (time.perf_counter() - t_settle) * 1e3 < args.settle_ms:
        for _ in range(8):
            trace()
        settle_launches += 8
        ctx.sync()
    for _ in range(args.warmup):
        trace()
    ctx.sync()
    ctx.record(0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        trace()
    ctx.record(1)
    ctx.sync()
    wall = time.perf_counter() - t0
    kernel_ms = ctx.elapsed_ms(0, 1) / args.steps
    mrays = n * args.steps / wall / 1e6
    if args.trace_only:
        print(json.dumps({"metric": METRIC, "value": round(mrays, 2), "unit": "Mrays/s", "kernel_ms": round(kernel_ms, 4),
                          "steps": args.steps, "rays": n, "ray_order": args.ray_order,
                          "walk_stack": ctx.bvh_info().walkStack}))
        ctx.close()
        return 0

    # the other entry point on the same rays (same hits)
    other = trace_linear if args.ray_order == "tiles" else trace_tiles
    other()
    ctx.record(2)
    for _ in range(max(1, args.steps // 2)):
        other()
    ctx.record(3)
    other_kernel_ms = ctx.elapsed_ms(2, 3) / max(1, args.steps // 2)

    # reference-order (unpruned skip-link) GPU kernel on the same rays
    ctx.record(2)
    for _ in range(max(1, args.steps // 10)):
        ctx.trace_closest_device(d_rays, n, d_hits, reference_order=True)
    ctx.record(3)
    ref_kernel_ms = ctx.elapsed_ms(2, 3) / max(1, args.steps // 10)
    # packed binary-BVH walk on the same rays
    ctx.trace_closest_device(d_rays, n, d_hits, binary=True)
    ctx.record(2)
    for _ in range(max(1, args.steps // 2)):
        ctx.trace_closest_device(d_rays, n, d_hits, binary=True)
    ctx.record(3)
    bin_kernel_ms = ctx.elapsed_ms(2, 3) / max(1, args.steps // 2)

    # path tracer on the same scene: 1080p, pt_spp samples per frame, 8 bounces, RR after 3;
    # launch shape from the autotuner on a 4-spp frame (several passes, so
    # that the pass-stream overlap shows)
    d_out = ctx.alloc(n * 16)
    p1 = sa.render_params(scene.camera, spp=4, max_depth=8, frame=0, sampling=1, sky=scene.sky)
    tuned = choose_launch(ctx, args.launch, lambda: ctx.render_device(p1, d_out), n)
    pt = pt_frame_bench(ctx, scene, args.pt_spp, max(1, args.pt_frames), d_out)
    pt["launch"] = with_graph_state(tuned["best"])
    pt["hw_queues"] = int(os.environ["GPU_MAX_HW_QUEUES"])
    pt["autotune_4spp_ms"] = {"default": tuned.get("default_ms"), "best": tuned.get("ms")}
    ctx.free(d_out)

    # the same scene with the tree built on the device (linear BVH, then
    # locally-ordered clustering): build time and closest-hit rate of the wide
    # walk over it
    device_build = {}
    for method in ("lbvh", "ploc"):
        ctx.build_bvh_gpu(method)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.build_bvh_gpu(method)
        ctx.sync()
        dev_build_s = time.perf_counter() - t0
        for _ in range(3):
            ctx.trace_closest_device(d_rays, n, d_hits)
        ctx.sync()
        ctx.record(2)
        for _ in range(args.steps):
            ctx.trace_closest_device(d_rays, n, d_hits)
        ctx.record(3)
        ctx.sync()
        dev_kernel_ms = ctx.elapsed_ms(2, 3) / args.steps
        device_build[method] = {"build_s": round(dev_build_s, 4), "wide_kernel_ms": round(dev_kernel_ms, 4),
                                "mrays_per_s": round(n / (dev_kernel_ms * 1e-3) / 1e6, 2)}
    ctx.free(d_rays)
    ctx.free(d_hits)

    cpu = None
    roof = None
    if not args.no_cpu_baseline:
        # the same deterministic build, host side: the wide layout for the oracle
        hb = sa.HostBvh.build(scene, spatial_split_alpha=SPATIAL_SPLIT_ALPHA)
        if hb.nodes.tobytes() != nodes.tobytes():
            raise RuntimeError("host and context BVH builds differ")
        cores, res = cpu_baseline(scene, nodes, rays, hb.wide, hb.wide_tris)
        wd, pr = res["wide"], res["pruned"]
        # algorithmic bytes of the wide kernel: ray in, hit out, one 80-B node
        # per wide visit, one 48-B triangle per test (counts from the oracle's
        # restatement of the same walk on the same rays)
        bytes_per_ray = 32 + 20 + 80 * wd["inner_per_ray"] + 48 * wd["leaf_per_ray"]
        achieved = n * bytes_per_ray / (kernel_ms * 1e-3) / 1e9
        traffic = load_traffic("trace_closest_config2")
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "limiter": limiter(load_counters("trace_closest_config2"), traffic, kernel_ms),
                "kernel": "k_trace_closest (compressed 8-wide walk, " + args.ray_order + " order)", "kernel_ms": round(kernel_ms, 4),
                "bytes_per_ray": round(bytes_per_ray, 1),
                "wide_node_visits_per_ray": round(wd["inner_per_ray"], 2),
                "triangle_tests_per_ray": round(wd["leaf_per_ray"], 2),
                "binary_walk": {"node_fetches_per_ray": round(pr["inner_per_ray"], 2),
                                "triangle_tests_per_ray": round(pr["leaf_per_ray"], 2),
                                "bytes_per_ray": round(32 + 20 + 64 * pr["inner_per_ray"] + 48 * pr["leaf_per_ray"], 1),
                                "kernel_ms": round(bin_kernel_ms, 4)}}
        cpu = {"value": round(res["reference_order"]["mrays_per_s"], 4), "unit": "Mrays/s", "cores": cores,
               "kind": "port",
               "sample": f"the full {n}-ray 1080p primary frame, oracle closestHit in reference order "
                         f"(skip-link, unpruned), median of 3, -march=native build",
               "cpu_model": cpu_model(), "host_threads": cpu_threads(),
               "pruned_mrays_per_s": round(pr["mrays_per_s"], 4), "wide_mrays_per_s": round(wd["mrays_per_s"], 4)}
        p = sa.render_params(scene.camera, spp=4, max_depth=8, frame=1, sampling=1, sky=scene.sky)
        pt["cpu_baseline"] = cpu_pt_baseline(scene, nodes, p)

    # the main context (its stream, four pass streams and graph) is closed
    # first: with its streams still mapped to hardware queues the launch-bound
    # 960x540 frames of the time-to-RMSE leg ran 1.7x slower
    ctx.close()
    ttr = None if args.no_time_to_rmse else time_to_rmse_leg()

    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded PCG scene and rays)",
        "config": {"workload": "1920x1080 jittered primary rays, closest hit, 1.16M-triangle heightfield + "
                               "2,000-sphere scene",
                   "bvh": f"host binned SAH, spatial splits alpha {SPATIAL_SPLIT_ALPHA}",
                   "ray_order": "8x8 tiles of the row-major frame, one per wave (stk_trace_closest_image_device)"
                                if args.ray_order == "tiles" else
                                "64 consecutive rays per wave (stk_trace_closest_device)",
                   "rays": n, "triangles": ntri, "bvh_nodes": int(nodes.size), "bvh_build_s": round(build_s, 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": {"settle": {"ms": args.settle_ms, "launches": settle_launches},
                  ("linear_order_gpu_mrays_per_s" if args.ray_order == "tiles" else "tile_order_gpu_mrays_per_s"):
                      round(n / (other_kernel_ms * 1e-3) / 1e6, 2),
                  "reference_order_gpu_mrays_per_s": round(n / (ref_kernel_ms * 1e-3) / 1e6, 2),
                  "binary_bvh_gpu_mrays_per_s": round(n / (bin_kernel_ms * 1e-3) / 1e6, 2),
                  "device_built_tree": device_build,
                  "path_tracer": pt,
                  "time_to_rmse": ttr},
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
