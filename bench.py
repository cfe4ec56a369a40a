#!/usr/bin/env python3
"""ReSTIR GI frame benchmark on one MI355X.

Workload (the 1080p ReSTIR GI configuration): 1920x1080, initial sample ->
temporal reuse -> spatial reuse per frame, max depth 4 (numberOfBounces),
temporal M cap 20, 5 spatial neighbours within 30 px, spatial M cap 300,
traced visibility rays, on a seeded terrain scene: a 1024x1024-vertex fBm
heightfield (2,093,058 tris) + 512 icospheres of 1,280 tris + 64 quad lights
= 2,748,546 triangles, camera on a seeded Catmull-Rom path over 128 frames.
A "step" is one frame.  Inputs (scene, BVH, buffers) are resident in HBM
before the timed region; the UBO is passed by value per frame.

Prints one JSON line (see the end of main()).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "ms/frame & Mrays/s @1080p ReSTIR GI; equal-RMSE time, MLP cache vs plain"
HBM_PEAK_GBS = 8000.0

W, H = 1920, 1080
SCENE = dict(seed=1, grid=1024, spheres=512, lights=64)
CAMERA = dict(seed=2, n_ctrl=8, frames=128)
PARAMS = dict(temporal_m_cap=20, spatial_m_cap=300, spatial_samples=5, spatial_radius=30.0)
BOUNCES = 4
FOV = 45.0

# Algorithmic HBM bytes per pixel of each pass (compulsory traffic counted once):
#   initial : sample write 84 (x_v, n_v, x_s, n_s, L, f, matID, seed, p_q) + G-buffer pos/nrm/matUV 48 + motion 4
#   temporal: sample 76 + motion 4 + validity 4 + prev reservoir 84 + reservoir write 84 + output 4
#   spatial : centre x_v, n_v, n_s, matID 40 + centre reservoir 84 + reservoir write 84 + output 4
ALGO_BYTES = {"initial": 136, "temporal": 256, "spatial": 212}

# HBM bytes per launch of each pass's kernel on this workload, from rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE runs of `bench.py --steps 20 --warmup 3
# --no-cpu-baseline --no-mlp-study` (KiB per dispatch, averaged over the 20
# timed frames; the initial pass adds its head kernel and the two tail
# launches of one frame), each scaled by what the counter reads for a plain
# 1 GiB device copy in the same kind of run: FETCH_SIZE x 2.0 (the copy reads
# 524,316 KiB), WRITE_SIZE x 1.0 (1,048,576 KiB).  They are measured, not
# computed here; the committed pmc summary under profiles/ holds the runs.
PMC_FETCH_KIB = {"initial": 25312.8 + 23296.0, "temporal": 258479.1, "spatial": 657584.4}
PMC_WRITE_KIB = {"initial": 380636.0 + 33106.8, "temporal": 340205.8, "spatial": 298874.2}
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0

# The traversal kernels are bound by VALU issue, not by HBM: wave64 VALU
# instructions per frame (SQ_INSTS_VALU in a rocprofv3 --pmc run of the same
# command, averaged over the same frames; the initial pass adds its tail
# launches, which its timing events bracket too) against the issue peak of 1,024
# SIMDs, each retiring one wave64 VALU instruction per 4 cycles at 2.4 GHz.
PMC_VALU_INSTS = {"initial": 440243888.0 + 80125810.0, "spatial": 637689899.0}
VALU_PEAK_GINST_S = 1024 * 2.4 / 4  # 614.4 G wave64 instructions per second


def pmc_traffic(name: str) -> int:
    return int((PMC_FETCH_KIB[name] * FETCH_SCALE + PMC_WRITE_KIB[name] * WRITE_SCALE) * 1024)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline work")
    ap.add_argument("--no-mlp-study", action="store_true")
    ap.add_argument("--study-frames", type=int, default=128)
    ap.add_argument("--conv-spp", type=int, default=65536, help="converged image of the cache study")
    ap.add_argument("--c5-conv-spp", type=int, default=4096,
                    help="converged images of the 3840x2160 workload (three checkpoint views)")
    ap.add_argument("--builder", choices=("lbvh", "ploc", "sah", "sah_host"), default="sah",
                    help="BVH builder (the build itself is outside the timed region): binned SAH on the "
                         "GPU (default: the best trees), the same SAH tree built on host threads, or the "
                         "GPU builders LBVH / PLOC")
    ap.add_argument("--workload", choices=("c2", "c1", "c3", "c5", "scenes"), default="c2",
                    help="c2: 1080p ReSTIR GI (the headline line); c1: 640x360 plain path tracing on the "
                         "Cornell box; c3: 12 batched views of the 1080p scene (HIP graph, autotuned); "
                         "c5: 3840x2160 ReSTIR GI + FP8 MLP cache on the large terrain; scenes: the "
                         "reference's own glTF scenes cb2 and sb2 at 1920x1080 (context for its README)")
    return ap.parse_args()


# MLP FLOPs: 2 * (96*64 + 4*64*64 + 64*3) = 45,440 per query; a training
# sample adds the backward pass (about 2x the forward, plus the output layer):
# 136,320.
MLP_FLOP_QUERY = 45440
MLP_FLOP_TRAIN = 136320
MFMA_BF16_PEAK_TFLOPS = 2500.0
SPREADS = (0.01, 100.0)
# spatial neighbour counts of the study; 0 keeps the spatial pass (the
# centre reservoir's M cap, W and shading) without neighbours
STUDY_NEIGHBOURS = (5, 3, 2, 1, 0)


def mlp_study(scene, camera, args) -> dict:
    """Equal-RMSE comparison of ReSTIR GI with and without the MLP radiance
    cache on the 1080p configuration: the camera runs the workload's seeded
    path (frames 1..study_frames), each run from empty history (with the
    cache, a freshly initialised network trained online, one Adam step per
    frame on 65,536 path suffixes).  The error is the RMSE of the
    pre-tone-mapping radiance of the last frame against an FP64 converged
    image of that frame's view (conv_spp samples per pixel of the estimator
    the ReSTIR output converges to); ms/frame is the sum of the GPU pass
    times (initial, cache, temporal, spatial) over all frames.  The reference
    point is plain ReSTIR GI with the workload's parameters (5 spatial
    neighbours).  The cache form is swept over the spread threshold c (paths
    stop at x_s when its spread exceeds c times the primary footprint) and
    the spatial neighbour count; plain ReSTIR GI with fewer neighbours is run
    too.  The equal-RMSE time ratio is the cheapest cache setting at or below
    the reference's error, divided by the reference's time.  `ratio_floor`
    is what no cache setting can beat: the plain frame's initial-sample and
    temporal passes (camera and x_s rays of every pixel, which ReSTIR GI needs
    whatever L_out's source) over its whole frame.  The comparison is
    repeated with the visibility check as the reference ships it (always
    visible: `reference_visibility`)."""
    import numpy as np

    from xvk_ray_amd import native as N

    frames = args.study_frames
    views = [camera.view(f) for f in range(frames)]
    last = frames - 1
    ctx = N.Context(scene, W, H, builder=builder_id(args))
    ubo_last = N.make_ubo(views[last], views[last - 1] if last else views[last], FOV, W, H, bounces=BOUNCES,
                          frame=frames)
    t0 = time.perf_counter()
    done = 0
    chunk = 64
    while done < args.conv_spp:
        k = min(chunk, args.conv_spp - done)
        ctx.converge(ubo_last, k, (1 << 20) + done)
        ctx.sync()
        done += k
        if done % 4096 == 0 or done == args.conv_spp:
            log(f"converged image {done}/{args.conv_spp} spp ({time.perf_counter() - t0:.1f} s)")
    conv = ctx.converged_mean()
    conv_s = time.perf_counter() - t0

    def run(cache: bool, spread: float, neighbours: int, vis: int = N.XR_VIS_TRACE) -> dict:
        ctx.resize(W, H)  # empty history
        if cache:
            ctx.mlp_reset(1)
        flags = N.XR_FLAG_RADIANCE | (N.XR_FLAG_MLP_CACHE if cache else 0)
        p = N.default_params(visibility=vis, flags=flags, mlp_spread=spread,
                             **dict(PARAMS, spatial_samples=neighbours))
        ctx.reset_counters()
        ctx.reset_pass_times()
        for f in range(frames):
            ubo = N.make_ubo(views[f], views[f - 1] if f else views[f], FOV, W, H, bounces=BOUNCES, frame=f + 1)
            ctx.render(ubo, N.XR_MODE_RESTIR, p)
        rad = ctx.read(N.XR_BUF_RADIANCE)[:, :3].astype(np.float64)
        err = float(np.sqrt(np.mean((rad - conv) ** 2)))
        ctx.sync()
        pass_ms, nfr = ctx.pass_time_totals()
        ext, shadow, valid, queries, train = ctx.counters_all()
        ms = sum(pass_ms) / nfr
        out = {"spatial_samples": neighbours, "ms_per_frame": round(ms, 4),
               "pass_ms": {k: round(v / nfr, 4) for k, v in zip(("initial", "temporal", "spatial", "cache"), pass_ms)},
               "rays_per_frame": (ext + shadow) // nfr, "rmse_last_frame": round(err, 6)}
        if cache:
            flop = (MLP_FLOP_QUERY * queries + MLP_FLOP_TRAIN * train) / nfr
            tflops = flop / (pass_ms[3] / nfr * 1e-3) / 1e12 if pass_ms[3] > 0 else 0.0
            out.update({"spread_c": spread, "queries_per_frame": queries // nfr,
                        "train_samples_per_frame": train // nfr,
                        "mlp": {"gflop_per_frame": round(flop / 1e9, 3), "achieved_TFLOPs": round(tflops, 2),
                                "frac_mfma_bf16": round(tflops / MFMA_BF16_PEAK_TFLOPS, 5)}})
        if vis != N.XR_VIS_TRACE:
            out["visibility"] = "reference"
        log(f"{'cache c=' + str(spread) if cache else 'plain'} K={neighbours} vis={vis}: {ms:.3f} ms/frame "
            f"{out['pass_ms']}, rmse {err:.4f}")
        return out

    plain = run(False, 0.0, PARAMS["spatial_samples"])
    plain_fewer = [run(False, 0.0, k) for k in STUDY_NEIGHBOURS if k != PARAMS["spatial_samples"]]
    sweep = [run(True, c, k) for k in STUDY_NEIGHBOURS for c in SPREADS]
    ok = [r for r in sweep if r["rmse_last_frame"] <= plain["rmse_last_frame"]]
    best = min(ok, key=lambda r: r["ms_per_frame"]) if ok else None
    ref_plain = run(False, 0.0, PARAMS["spatial_samples"], N.XR_VIS_REFERENCE)
    ref_cache = [run(True, SPREADS[0], k, N.XR_VIS_REFERENCE) for k in STUDY_NEIGHBOURS]
    ref_ok = [r for r in ref_cache if r["rmse_last_frame"] <= ref_plain["rmse_last_frame"]]
    ref_best = min(ref_ok, key=lambda r: r["ms_per_frame"]) if ref_ok else None
    ctx.close()
    pm = plain["pass_ms"]
    return {
        "camera": f"the workload's seeded path, frames 1-{frames}", "frames": frames,
        "converged_spp": args.conv_spp, "converged_seconds": round(conv_s, 2),
        "error": f"RMSE of frame {frames}'s radiance against the converged image of its view",
        "plain": plain, "plain_fewer_neighbours": plain_fewer, "cache": sweep,
        "time_ratio_at_equal_rmse": (round(best["ms_per_frame"] / plain["ms_per_frame"], 4) if best else None),
        "equal_rmse_setting": ({"spread_c": best["spread_c"], "spatial_samples": best["spatial_samples"]}
                               if best else None),
        "target_ratio": 0.6,
        "ratio_floor": round((pm["initial"] + pm["temporal"]) / plain["ms_per_frame"], 4),
        "reference_visibility": {
            "plain": ref_plain, "cache": ref_cache,
            "time_ratio_at_equal_rmse": (round(ref_best["ms_per_frame"] / ref_plain["ms_per_frame"], 4)
                                         if ref_best else None),
        },
    }


def host_info() -> dict:
    """The host the CPU baseline runs on: logical CPUs, the CPUs this process
    may run on, the CPU model, the OpenMP thread setting, and whether the
    tools for the lavapipe baseline (dxc, a lavapipe Vulkan ICD) exist."""
    import glob
    import shutil
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    icds = glob.glob("/usr/share/vulkan/icd.d/*lvp*") + glob.glob("/etc/vulkan/icd.d/*lvp*")
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "cpu_model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "dxc": shutil.which("dxc") is not None,
            "lavapipe_icd": bool(icds), "vulkan_loader": bool(glob.glob("/usr/lib/x86_64-linux-gnu/libvulkan.so*"))}


def baseline_threads() -> int:
    """Every CPU this process may run on, unless OMP_NUM_THREADS asks for fewer
    (the GPU hosts give one GPU's process a share of their cores that way)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def cpu_baseline(scene, app, ctx, target_s: float) -> dict:
    """The oracle (C++ restatement of the shaders, OpenMP over the host cores)
    on a bounded pixel sample of the same frame, scaled to a full frame."""
    import numpy as np

    from oracle.binding import Oracle, OracleScene, temporal
    from xvk_ray_amd import native as N

    L = Oracle.lib()
    threads = baseline_threads()
    L.orc_set_threads(threads)
    v, i, m, o = scene.arrays()
    orc = OracleScene(v, i, m, o)
    ubo = app.last_ubo()
    ub = N.ubo_to_bytes(ubo)
    frame = ubo.current_frame
    samples = ctx.read(N.XR_BUF_SAMPLES).copy()
    motion = ctx.read(N.XR_BUF_MOTION).copy()
    gpos = np.ascontiguousarray(ctx.read(N.XR_BUF_GBUF_POS))
    tres = ctx.read(N.XR_BUF_TEMPORAL).copy()
    sres = ctx.read(N.XR_BUF_SPATIAL).copy()
    out = ctx.read(N.XR_BUF_OUTPUT).copy()
    n = W * H
    rng = np.random.default_rng(5)

    def run(npix):
        px = np.sort(rng.choice(n, size=npix, replace=False)).astype(np.uint32)
        t0 = time.perf_counter()
        r = orc.initial_sample(ub, W, H, pixels=px)
        t1 = time.perf_counter()
        temporal(W, H, frame, PARAMS["temporal_m_cap"], samples, motion, gpos, tres, sres, out, pixels=px)
        t2 = time.perf_counter()
        shadow = orc.spatial(W, H, frame, 1, PARAMS["spatial_samples"], PARAMS["spatial_radius"],
                             PARAMS["spatial_m_cap"], 1, samples, tres, sres, out, pixels=px)
        t3 = time.perf_counter()
        return t3 - t0, r["rays"] + shadow, (t1 - t0, t2 - t1, t3 - t2)

    dt, _, _ = run(4096)
    npix = int(min(n, max(4096, 4096 * target_s / max(dt, 1e-3))))
    # Below a full frame: one run over npix random pixels.  A full frame that
    # finishes early is repeated until the target time is reached.
    total, rays, reps, parts = 0.0, 0, 0, [0.0, 0.0, 0.0]
    while True:
        dt, r, p = run(npix)
        total += dt
        rays += r
        reps += 1
        parts = [a + b for a, b in zip(parts, p)]
        if npix < n or total >= target_s or reps >= 64:
            break
    ms_frame = total * 1e3 * n / (npix * reps)
    what = (f"{npix} random pixels of one 1920x1080 frame" if npix < n
            else f"{reps} full 1920x1080 frames")
    return {
        "value": round(rays / total / 1e6, 3),
        "unit": "Mrays/s",
        "ms_per_frame": round(ms_frame, 1),
        "cores": threads,
        "host": host_info(),
        "kind": "port",
        "sample": f"{what} (all three passes, traced visibility), {total:.1f} s; "
                  f"C++ restatement of the shaders, not lavapipe",
        "pass_seconds": [round(x, 3) for x in parts],
    }


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def builder_id(args) -> int:
    from xvk_ray_amd import native as N
    return {"lbvh": N.XR_BUILDER_LBVH, "ploc": N.XR_BUILDER_PLOC, "sah": N.XR_BUILDER_SAH,
            "sah_host": N.XR_BUILDER_SAH_HOST}[args.builder]


def base_result(args, metric, value, unit, ms, dtype, config) -> dict:
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic", "config": config}


def run_c1(args) -> dict:
    """Plain 1-spp path tracing, 640x360, Cornell box (24 quad tris + 1,280-tri
    icosphere + ceiling light), pinhole 60 degrees at (0.5, 0.5, -1.2), max
    depth 3, progressive accumulation into the RGBA32F image.  The sample cap
    (maxSpp) is set above warm-up + timed frames so every frame traces (the
    driver stops tracing once the cap is reached, pathTracer.cpp:71-78)."""
    from xvk_ray_amd import native as N

    Wc, Hc = 640, 360
    frames = args.warmup + args.steps
    scene = N.Scene.cornell(1)
    camera = N.Camera.static((0.5, 0.5, -1.2), (0.5, 0.5, 0.5))
    app = N.App(scene, camera, Wc, Hc, mode=N.XR_MODE_PLAIN, spp=1, bounces=3, max_spp=frames + 1, fov=60.0,
                accumulate=True, sky=False)
    ctx = app.context()
    ctx.build_accel(builder_id(args))
    params = N.default_params(plain_max_depth=3)
    for _ in range(args.warmup):
        app.draw_frame(params)
    ctx.sync()
    steps = args.steps
    ctx.reset_counters()
    assert args.steps >= 1
    ctx.reset_pass_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        app.draw_frame(params)
    ctx.sync()
    wall = time.perf_counter() - t0
    ext, _, _ = ctx.counters3()
    pass_ms, nfr = ctx.pass_time_totals()
    ms = wall * 1e3 / steps
    res = base_result(args, "ms/frame & Mrays/s @640x360 plain path tracing", round(ext / wall / 1e6, 2), "Mrays/s",
                      round(ms, 4), "f32",
                      {"workload": "plain_pt_640x360_cornell", "width": Wc, "height": Hc,
                       "triangles": scene.triangle_count, "max_depth": 3, "spp": 1, "accumulated_frames": frames,
                       "parallelism": "single_gpu"})
    res.update({"ms_per_frame": round(ms, 4), "kernel_ms_per_frame": round(pass_ms[0] / nfr, 4),
                "rays_per_frame": ext // steps})
    log(f"c1: {ms:.4f} ms/frame (kernel {pass_ms[0] / nfr:.4f} ms), {ext // steps} rays/frame")
    if not args.no_cpu_baseline:
        import numpy as np

        from oracle.binding import Oracle, OracleScene
        L = Oracle.lib()
        threads = baseline_threads()
        L.orc_set_threads(threads)
        orc = OracleScene(*scene.arrays())
        ub = N.ubo_to_bytes(app.last_ubo())
        accum = np.zeros((Wc * Hc, 4), np.float32)
        out = np.zeros(Wc * Hc, np.uint32)
        total, rays, reps = 0.0, 0, 0
        while total < args.cpu_seconds and reps < 1000:
            t = time.perf_counter()
            rays += orc.plain(ub, Wc, Hc, 3, accum, out)
            total += time.perf_counter() - t
            reps += 1
        res["cpu_baseline"] = {"value": round(rays / total / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                               "host": host_info(), "kind": "port", "ms_per_frame": round(total * 1e3 / reps, 2),
                               "sample": f"{reps} full 640x360 frames, {total:.1f} s; C++ restatement of the "
                                         f"shaders, not lavapipe"}
    return res


C3_SIZES = [(256, 256)] * 4 + [(960, 540)] * 4 + [(1920, 1080)] * 2 + [(1280, 720)] * 2


def run_c3(args) -> dict:
    """12 views of uneven size (4x 256x256, 4x 960x540, 2x 1920x1080,
    2x 1280x720, 8,326,144 pixels) of the 1080p terrain scene, each on its own
    seeded camera path (orbit seeds 2..13, 64 frames) with its own reservoir
    buffers, ReSTIR GI as in the headline workload; one launch per pass for
    all views, the frame replayed as a HIP graph; tile shape, LDS stack
    depth and launch form (a workgroup per tile, or persistent workgroups
    looping over per-XCD tile queues) chosen by the autotuner on the first
    frame's views."""
    from xvk_ray_amd import native as N

    frames = 64
    scene = N.Scene.terrain(**SCENE)
    ctx = N.Context(scene, 8, 8, builder=builder_id(args))
    cams = [N.Camera.orbit(2 + v, 8, frames) for v in range(len(C3_SIZES))]
    views = [[c.view(f) for f in range(frames)] for c in cams]
    params = N.default_params(visibility=N.XR_VIS_TRACE, **PARAMS)

    def ubos(f):
        return [N.make_ubo(views[v][f], views[v][f - 1] if f > 0 else views[v][f], FOV, w, h, bounces=BOUNCES,
                           frame=f + 1) for v, (w, h) in enumerate(C3_SIZES)]

    all_ubos = [ubos(f) for f in range(frames)]
    batch = N.Batch(ctx, C3_SIZES)
    t = time.perf_counter()
    tune = batch.autotune(all_ubos[0], params, frames=4)
    cfg = batch.config
    log(f"c3 autotune ({time.perf_counter() - t:.1f} s): {[round(x, 3) for x in tune]} -> config {cfg} "
        f"{N.Batch.config_info(cfg)} persistent {N.Batch.config_persistent(cfg)}")
    warm = min(args.warmup, frames)
    for f in range(warm):
        batch.render(all_ubos[f], params)
    ctx.sync()
    batch.reset_counters()
    t0 = time.perf_counter()
    for f in range(warm, frames):
        batch.render(all_ubos[f], params)
    ctx.sync()
    wall = time.perf_counter() - t0
    timed = frames - warm
    ext, shadow = batch.counters()[:2]
    ms = wall * 1e3 / timed
    pixels = sum(w * h for w, h in C3_SIZES)
    tw, th, lds = N.Batch.config_info(cfg)
    res = base_result(args, "ms/frame & Mrays/s, 12 batched ReSTIR GI views", round((ext + shadow) / wall / 1e6, 2),
                      "Mrays/s", round(ms, 4), "f32",
                      {"workload": "restir_gi_12_views_batched", "views": [f"{w}x{h}" for w, h in C3_SIZES],
                       "pixels": pixels, "triangles": scene.triangle_count, "max_depth": BOUNCES,
                       "frames": frames, "graph": True, "parallelism": "single_gpu"})
    res["steps"], res["warmup"] = timed, warm
    # the same views and frames with the BF16 radiance cache: each view chooses
    # its own training samples, one training step covers all of them
    batch.reset()
    cparams = N.default_params(visibility=N.XR_VIS_TRACE, flags=N.XR_FLAG_MLP_CACHE, **PARAMS)
    ctx.mlp_reset(1)
    for f in range(warm):
        batch.render(all_ubos[f], cparams)
    ctx.sync()
    batch.reset_counters()
    t0 = time.perf_counter()
    for f in range(warm, frames):
        batch.render(all_ubos[f], cparams)
    ctx.sync()
    cwall = time.perf_counter() - t0
    cext, cshadow, _, cq, ct = batch.counters()
    cms = cwall * 1e3 / timed
    log(f"c3 with the BF16 cache: {cms:.3f} ms/frame, {ct // timed} training samples and {cq // timed} queries "
        f"per frame")
    res.update({"ms_per_frame": round(ms, 4), "ns_per_pixel": round(ms * 1e6 / pixels, 3),
                "mlp_cache_bf16": {"ms_per_frame": round(cms, 4), "mrays_per_s": round((cext + cshadow) / cwall / 1e6, 1),
                                   "train_samples_per_frame": ct // timed, "queries_per_frame": cq // timed,
                                   "train_target_per_view": cparams.mlp_train_pixels},
                "rays_per_frame": (ext + shadow) // timed,
                "autotune": {"ms_per_config": [round(x, 4) for x in tune],
                             "configs": [list(N.Batch.config_info(k)) + [N.Batch.config_persistent(k)]
                                         for k in range(N.Batch.config_count())],
                             "config_fields": ["tile_w", "tile_h", "lds_stack_depth", "persistent"],
                             "chosen": {"tile": f"{tw}x{th}", "lds_stack_depth": lds,
                                        "persistent": N.Batch.config_persistent(cfg)}}})
    batch.close()
    ctx.close()
    return res


def run_c5(args) -> dict:
    """3840x2160 ReSTIR GI, max depth 6, on the large terrain (2048x2048-vertex
    fBm heightfield + 4,096 icospheres of 1,280 tris + 256 quad lights, seed
    1), camera on the seeded 128-frame path; plain ReSTIR GI vs ReSTIR GI with
    the MLP radiance cache in FP8 E4M3 weights, with the workload's 5 spatial
    neighbours, with 1 and without spatial neighbours (the cache form's
    equal-error setting at 1080p).  RMSE of each method's radiance against an FP64
    converged image of the same camera at frames 1, 16 and 128."""
    import numpy as np

    from xvk_ray_amd import native as N

    W5, H5, depth = 3840, 2160, 6
    frames = 128
    checkpoints = (1, 16, 128)
    t = time.perf_counter()
    scene = N.Scene.terrain(1, 2048, 4096, 256)
    camera = N.Camera.orbit(2, 8, frames)
    log(f"c5 scene {scene.triangle_count} tris ({time.perf_counter() - t:.1f} s)")
    saved = {}

    def run(cache: bool, neighbours: int = PARAMS["spatial_samples"]) -> dict:
        key = (cache, neighbours)
        app = N.App(scene, camera, W5, H5, mode=N.XR_MODE_RESTIR, spp=1, bounces=depth, fov=FOV, sky=False,
                    reuse=True)
        ctx = app.context()
        ctx.build_accel(builder_id(args))
        info = ctx.accel_info()
        flags = N.XR_FLAG_RADIANCE | ((N.XR_FLAG_MLP_CACHE | N.XR_FLAG_MLP_FP8) if cache else 0)
        p = N.default_params(visibility=N.XR_VIS_TRACE, flags=flags, **dict(PARAMS, spatial_samples=neighbours))
        ctx.reset_counters()
        ctx.reset_pass_times()
        t0 = time.perf_counter()
        for f in range(1, frames + 1):
            app.draw_frame(p)
            if f in checkpoints:
                saved[(key, f)] = (ctx.read(N.XR_BUF_RADIANCE)[:, :3].astype(np.float64), app.last_ubo())
                log(f"c5 {'cache' if cache else 'plain'} K={neighbours} frame {f}")
        ctx.sync()
        wall = time.perf_counter() - t0
        pass_ms, nfr = ctx.pass_time_totals()
        ext, shadow, valid, queries, train = ctx.counters_all()
        out = {"spatial_samples": neighbours, "ms_per_frame_gpu": round(sum(pass_ms) / nfr, 4),
               "pass_ms": {k: round(v / nfr, 4) for k, v in zip(("initial", "temporal", "spatial", "cache"), pass_ms)},
               "rays_per_frame": (ext + shadow) // nfr, "mrays_per_s": round((ext + shadow) / wall / 1e6, 1),
               "bvh": {"nodes": info.node_count, "depth": info.max_depth, "build_ms": round(info.build_ms, 1)}}
        if cache:
            out.update({"queries_per_frame": queries // nfr, "train_samples_per_frame": train // nfr,
                        "mlp_gflop_per_frame": round((MLP_FLOP_QUERY * queries + MLP_FLOP_TRAIN * train)
                                                     / nfr / 1e9, 3)})
        app.close()
        return out

    k5 = PARAMS["spatial_samples"]
    runs = {(False, k5): run(False), (True, k5): run(True), (True, 1): run(True, 1), (True, 0): run(True, 0)}
    # converged images of the checkpoint cameras
    ctx = N.Context(scene, W5, H5, builder=builder_id(args))
    rmse = {key: {} for key in runs}
    for f in checkpoints:
        ubo = saved[((False, k5), f)][1]
        ctx.converge_reset()
        done = 0
        while done < args.c5_conv_spp:
            k = min(32, args.c5_conv_spp - done)
            ctx.converge(ubo, k, (1 << 20) + done)
            ctx.sync()
            done += k
        log(f"c5 converged image for frame {f}: {done} spp")
        conv = ctx.converged_mean()
        for key in runs:
            rad = saved[(key, f)][0]
            rmse[key][f] = round(float(np.sqrt(np.mean((rad - conv) ** 2))), 6)
    ctx.close()
    for key, r in runs.items():
        r["rmse"] = rmse[key]
    res_plain, res_cache, res_cache1, res_cache0 = runs[(False, k5)], runs[(True, k5)], runs[(True, 1)], runs[(True, 0)]
    ms = res_cache["ms_per_frame_gpu"]
    res = base_result(args, "ms/frame & Mrays/s @3840x2160 ReSTIR GI + FP8 MLP radiance cache",
                      res_cache["mrays_per_s"], "Mrays/s", ms, "f32+fp8",
                      {"workload": "restir_gi_mlp_fp8_4k_large_terrain", "width": W5, "height": H5,
                       "triangles": scene.triangle_count, "max_depth": depth, "frames": frames,
                       "camera": "seeded 128-frame orbit", "parallelism": "single_gpu"})
    res["steps"], res["warmup"] = frames, 0
    res.update({"plain_restir": res_plain, "mlp_cache_fp8": res_cache,
                "mlp_cache_fp8_fewer_neighbours": [res_cache1, res_cache0], "converged_spp": args.c5_conv_spp,
                "time_ratio_cache_vs_plain": round(ms / res_plain["ms_per_frame_gpu"], 4),
                "time_ratio_cache_fewer_neighbours_vs_plain": [round(r["ms_per_frame_gpu"]
                                                                     / res_plain["ms_per_frame_gpu"], 4)
                                                               for r in (res_cache1, res_cache0)]})
    return res


GLTF_DIR = Path(__file__).resolve().parent / "tests" / "golden" / "gltf"
# The reference's README (README.md:7-9): frame cost of three unnamed scenes
# on unstated hardware, CPU wall time between frames; the ReSTIR GI columns are
# the UI value, which divides by 3.
README_MS = {"scene1": {"pt_1spp": 2.8, "restir_temporal": 2.9, "restir_temporal_spatial": 3.3},
             "scene2": {"pt_1spp": 1.2, "restir_temporal": 1.7, "restir_temporal_spatial": 2.3},
             "scene3": {"pt_1spp": 2.6, "restir_temporal": 2.8, "restir_temporal_spatial": 2.8}}


def run_scenes(args) -> dict:
    """The reference's own scenes through the glTF loader at 1920x1080 with the
    reference app's settings (src/main.cpp:41-50: 10 bounces, fov 45, no
    accumulation, no sky) at its default view (Scene.cpp:29): plain 1-spp path
    tracing (unbounded depth, rgen.hlsl) and ReSTIR GI in reference mode
    (M caps 30 / 300, 6 neighbours in 5 px, CheckVisibility always true).
    cb2: 32 triangles, untextured; sb2: 75,277 triangles, PNG textures.  Wall
    time per frame with the scene resident; the README's numbers are context
    only (unnamed scenes, unstated hardware, CPU frame deltas)."""
    from xvk_ray_amd import native as N

    out = {}
    total_rays, total_s = 0, 0.0
    for name in ("cb2", "sb2"):
        scene = N.Scene.gltf(GLTF_DIR / name / "scene.gltf")
        camera = N.Camera.static((0.0, 1.0, 3.0), (0.0, 1.0, 0.0))
        row = {"triangles": scene.triangle_count, "textures": len(scene.textures())}
        for mode_name, mode in (("pt_1spp", N.XR_MODE_PLAIN), ("restir", N.XR_MODE_RESTIR)):
            app = N.App(scene, camera, W, H, mode=mode, spp=1, bounces=10, fov=45.0, accumulate=False, sky=False,
                        reuse=True)
            ctx = app.context()
            ctx.build_accel(builder_id(args))
            params = N.default_params()
            for _ in range(args.warmup):
                app.draw_frame(params)
            ctx.sync()
            ctx.reset_counters()
            ctx.reset_pass_times()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                app.draw_frame(params)
            ctx.sync()
            wall = time.perf_counter() - t0
            ext, shadow, _ = ctx.counters3()
            pass_ms, nfr = ctx.pass_time_totals()
            ms = wall * 1e3 / args.steps
            entry = {"ms_per_frame": round(ms, 4), "rays_per_frame": (ext + shadow) // args.steps}
            if mode == N.XR_MODE_RESTIR:
                entry["pass_ms"] = {k: round(v / nfr, 4) for k, v in zip(("initial", "temporal", "spatial", "cache"),
                                                                            pass_ms)}
            info = ctx.accel_info()
            row["bvh"] = {"builder": args.builder, "nodes": info.node_count, "depth": info.max_depth}
            row[mode_name] = entry
            total_rays += ext + shadow
            total_s += wall
            log(f"scenes: {name} {mode_name} {ms:.4f} ms/frame")
            del ctx, app
        out[name] = row
    res = base_result(args, "ms/frame on the reference's glTF scenes @1920x1080 (context)",
                      round(total_rays / total_s / 1e6, 2), "Mrays/s", round(total_s * 1e3 / (4 * args.steps), 4),
                      "f32", {"workload": "reference_scenes_1080p", "width": W, "height": H, "scenes": ["cb2", "sb2"],
                              "bounces": 10, "parallelism": "single_gpu"})
    res.update({"scenes": out, "readme_ms_context": README_MS})
    return res


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "the frame library runs on one GPU only; --gpus must be 1",
                          "requested_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "c1":
        print(json.dumps(run_c1(args)))
        return
    if args.workload == "c5":
        print(json.dumps(run_c5(args)))
        return
    if args.workload == "c3":
        print(json.dumps(run_c3(args)))
        return
    if args.workload == "scenes":
        print(json.dumps(run_scenes(args)))
        return

    from xvk_ray_amd import native as N

    t_setup = time.perf_counter()
    scene = N.Scene.terrain(**SCENE)
    camera = N.Camera.orbit(**CAMERA)
    app = N.App(scene, camera, W, H, mode=N.XR_MODE_RESTIR, spp=1, bounces=BOUNCES, fov=FOV, sky=False, reuse=True)
    ctx = app.context()
    ctx.build_accel(builder_id(args))
    params = N.default_params(visibility=N.XR_VIS_TRACE, **PARAMS)
    info = ctx.accel_info()
    setup_s = time.perf_counter() - t_setup
    log(f"scene {info.triangle_count} tris, BVH {info.node_count} nodes depth {info.max_depth} "
        f"built in {info.build_ms:.1f} ms; setup {setup_s:.1f} s")

    for _ in range(args.warmup):
        app.draw_frame(params)
    ctx.sync()
    log(f"{args.warmup} warm-up frames done")
    ctx.reset_counters()
    ctx.reset_pass_times()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        app.draw_frame(params)
    ctx.sync()
    wall = time.perf_counter() - t0
    ext, shadow, valid_px = ctx.counters3()
    pass_ms, frames = ctx.pass_time_totals()
    assert frames == args.steps, (frames, args.steps)

    ms_per_step = wall * 1e3 / args.steps
    log(f"{args.steps} timed frames: {ms_per_step:.3f} ms/frame, passes {[round(x / frames, 3) for x in pass_ms]}")
    rays = ext + shadow
    mrays = rays / wall / 1e6
    names = ["initial", "temporal", "spatial"]
    per_pass = {}
    for k, name in enumerate(names):
        ms = pass_ms[k] / frames
        gbs = ALGO_BYTES[name] * W * H / (ms * 1e-3) / 1e9
        per_pass[name] = {"ms": round(ms, 4), "algo_bytes_per_px": ALGO_BYTES[name], "achieved_GBs": round(gbs, 1),
                          "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                          "hbm_GBs_measured": round(pmc_traffic(name) / (ms * 1e-3) / 1e9, 1)}
    dom = max(names, key=lambda k: per_pass[k]["ms"])
    roofline = {
        "kernel": {"initial": "k_initial_sample", "temporal": "k_temporal", "spatial": "k_spatial"}[dom],
        "bound": "hbm",
        "achieved": per_pass[dom]["achieved_GBs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": per_pass[dom]["frac_hbm"],
        "traffic": pmc_traffic(dom),
        "algo_bytes_per_launch": ALGO_BYTES[dom] * W * H,
    }
    # head kernel and spatial kernel against the VALU issue roofline
    issue = {}
    for name, kern in (("initial", "k_initial_sample+k_tail_bounce"), ("spatial", "k_spatial")):
        ms = pass_ms[names.index(name)] / frames
        ginst = PMC_VALU_INSTS[name] / (ms * 1e-3) / 1e9
        issue[kern] = {"bound": "valu", "achieved": round(ginst, 1), "peak": VALU_PEAK_GINST_S,
                       "unit": "G wave64 VALU instructions/s", "frac": round(ginst / VALU_PEAK_GINST_S, 4),
                       "valu_insts_per_frame": int(PMC_VALU_INSTS[name])}
    result = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "restir_gi_1080p_terrain",
            "width": W, "height": H, "triangles": info.triangle_count, "max_depth": BOUNCES,
            "temporal_m_cap": PARAMS["temporal_m_cap"], "spatial_samples": PARAMS["spatial_samples"],
            "spatial_radius_px": PARAMS["spatial_radius"], "visibility": "traced",
            "bvh": {"builder": args.builder, "nodes": info.node_count, "depth": info.max_depth,
                    "build_ms": round(info.build_ms, 2)},
            "parallelism": "single_gpu",
        },
        "ms_per_frame": round(ms_per_step, 3),
        "rays_per_frame": rays // args.steps,
        "extend_rays_per_frame": ext // args.steps,
        "shadow_rays_per_frame": shadow // args.steps,
        "valid_primary_frac": round(valid_px / (args.steps * W * H), 4),
        "passes": per_pass,
        "temporal_plus_spatial_ms": round(per_pass["temporal"]["ms"] + per_pass["spatial"]["ms"], 4),
        "roofline": roofline,
        "issue_roofline": issue,
        # registers, scratch, LDS and resident waves of the traversal kernels
        "kernels": {k["name"]: {f: k[f] for f in ("vgprs", "scratch_bytes", "lds_bytes", "waves_per_simd")}
                    for k in N.kernel_info() if k["name"] in ("k_initial_sample", "k_tail_bounce", "k_spatial")},
        "setup_s": round(setup_s, 2),
    }
    if not args.no_mlp_study:
        log("MLP radiance cache equal-RMSE study")
        result["mlp_cache_study"] = mlp_study(scene, camera, args)
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle on a pixel sample)")
        result["cpu_baseline"] = cpu_baseline(scene, app, ctx, args.cpu_seconds)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
