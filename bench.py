"""Benchmark: Mpaths/s of the per-sample radiance path on one MI355X.

Workload (one step = one full render): the 100,820-triangle displaced
geodesic icosphere on a ground plane with 8 quad lights, 1920 x 1080 at
64 spp, path depth 6, balance-heuristic MIS of one light and one BSDF sample
per vertex, FP32 shading with FP64 per-pixel accumulation.  Scene, BVH and
accumulator are resident in HBM before the timed region.

Prints one JSON line:
metric/value/unit, roofline of the dominant kernel (closest-hit traversal)
measured with HIP events on the render stream over the timed region, and a
CPU baseline (the FP64 oracle on the host cores, a bounded row sample).
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

NODE_BYTES = 64       # one quantized 4-wide node (chk_bvh.h Node4Q)
TRI_BYTES = 48        # one leaf triangle record (three float4)
RAY_BYTES = 32 + 16   # chk_ray in + hit record out
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=6)
    ap.add_argument("--freq", type=int, default=71, help="geodesic frequency (71 -> 100,820 triangles)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 4 times (the known-size kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters), then exit")
    ap.add_argument("--cpu-row-step", type=int, default=1)
    ap.add_argument("--cpu-spp", type=int, default=64)
    ap.add_argument("--traffic", type=str,
                    default=str(ROOT / "profiles" / "closest_hit_traffic_icosphere100k_1080p.json"),
                    help="JSON with the measured HBM bytes per closest-hit launch (from the "
                         "FETCH_SIZE / WRITE_SIZE runs recorded in that file)")
    # time-to-RMSE leg: learned proposal vs plain MIS (config 5 scene)
    ap.add_argument("--no-rmse", action="store_true")
    ap.add_argument("--rmse-width", type=int, default=1920)
    ap.add_argument("--rmse-height", type=int, default=1080)
    ap.add_argument("--rmse-row-step", type=int, default=16,
                    help="the reference (and the RMSE) covers every N-th row of the frame")
    ap.add_argument("--rmse-full-reference", action="store_true",
                    help="also render the reference for every row and report the RMSE of the "
                         "final images over the whole frame (about 4 minutes at 1080p)")
    ap.add_argument("--rmse-tris", type=int, default=2_000_000)
    ap.add_argument("--rmse-lights", type=int, default=32)
    ap.add_argument("--rmse-depth", type=int, default=4)
    ap.add_argument("--rmse-ref-spp", type=int, default=65536)
    ap.add_argument("--rmse-pass-spp", type=int, default=32)
    ap.add_argument("--rmse-max-spp", type=int, default=8192)
    ap.add_argument("--rmse-target", type=float, default=1e-2)
    ap.add_argument("--rmse-check-spp", type=int, default=64)
    ap.add_argument("--train-spp", type=int, default=32)
    ap.add_argument("--train-batch", type=int, default=16384)
    ap.add_argument("--train-steps", type=int, default=64, help="Adam steps per bounce")
    ap.add_argument("--train-lr", type=float, default=1e-3)
    # batching leg (config 4) and LTC leg (config 3)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-requests", type=int, default=64)
    ap.add_argument("--batch-scenes", type=int, default=16)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--no-ltc", action="store_true")
    ap.add_argument("--no-mlp", action="store_true")
    ap.add_argument("--mlp-queries", type=int, default=480 * 270 * 32)
    ap.add_argument("--no-direct", action="store_true")
    ap.add_argument("--direct-spp", type=int, default=1)
    ap.add_argument("--direct-cpu-row-step", type=int, default=2)
    ap.add_argument("--no-sources", action="store_true")
    ap.add_argument("--no-calibration", action="store_true")
    ap.add_argument("--calib-rays", type=int, default=20000)
    ap.add_argument("--calib-points", type=int, default=4000)
    ap.add_argument("--sources-spp", type=int, default=16)
    ap.add_argument("--sources-cpu-row-step", type=int, default=1)
    ap.add_argument("--ltc-points", type=int, default=4 * 1024 * 1024)
    ap.add_argument("--ltc-quads", type=int, default=16)
    return ap.parse_args()


def batch_leg(args, ctx):
    """Config 4: requests of log-uniform size 64..1024 per side, spp from
    {1, 4, 16, 64, 256}, scenes from the Cornell generator with log-uniform
    1k..500k triangles and 1..16 lights (a pool of batch_scenes scenes the
    requests draw from).  Autotune, then time the batch."""
    import dataclasses

    import numpy as np

    from chakumrender_amd import abi, host, scenes

    t0 = time.perf_counter()
    pool_sc, reqs = scenes.batch_workload(n_requests=args.batch_requests, n_scenes=args.batch_scenes)
    pool = [abi.DeviceScene(ctx, sc) for sc in pool_sc]
    items = []
    for q in reqs:
        ds = pool[q["scene"]]
        cam = dataclasses.replace(ds.scene.camera, width=q["width"], height=q["height"])
        items.append(dict(ds=ds, camera=cam, params=dict(spp=q["spp"], max_depth=q["max_depth"],
                                                         seed=q["seed"])))
    setup_s = time.perf_counter() - t0
    t1 = time.perf_counter()
    best, tune_ms = host.batch_autotune(ctx, items)
    tune_s = time.perf_counter() - t1
    b = host.Batch(ctx, items)
    b.run(best)  # warm-up (and graph capture when chosen)
    abi.sync()
    lats, runs = [], []
    t2 = time.perf_counter()
    for _ in range(args.batch_reps):
        lat, st = b.run(best)
        lats.append(lat)
        runs.append(st)
    wall = (time.perf_counter() - t2) / args.batch_reps
    paths = runs[-1]["paths"]
    lat = np.array(lats[-1])
    dev_ms = float(np.mean([r["batch_ms"] for r in runs]))
    # the same tuning through captured graphs (capture on the first call,
    # replays after)
    gtun = dict(best, use_graph=1)
    gb = host.Batch(ctx, items)
    gb.run(gtun)
    abi.sync()
    glats, gruns = [], []
    for _ in range(args.batch_reps):
        gl, gst = gb.run(gtun)
        glats.append(gl)
        gruns.append(gst)
    glat = np.array(glats[-1])
    del gb
    # the same requests one chk_render call each, in the batch's order
    # (shortest first), into their own accumulators
    imgs = [host.Image(it["camera"].width, it["camera"].height) for it in items]
    order = sorted(range(len(items)), key=lambda k: items[k]["camera"].width *
                   items[k]["camera"].height * items[k]["params"]["spp"])
    for k in order:  # warm-up (buffers)
        host.render(items[k]["ds"], items[k]["params"], camera=items[k]["camera"], image=imgs[k],
                    stats=False, download=False)
    abi.sync()
    t3 = time.perf_counter()
    for k in order:
        host.render(items[k]["ds"], items[k]["params"], camera=items[k]["camera"], image=imgs[k],
                    stats=False, download=False)
    abi.sync()
    seq_ms = (time.perf_counter() - t3) * 1e3
    del imgs
    out = {
        "workload": f"{args.batch_requests} requests, sizes 64..1024 log-uniform, spp in "
                    f"{{1,4,16,64,256}}, depth 4, {args.batch_scenes} Cornell-generator scenes "
                    f"1k..500k triangles, 1..16 lights",
        "paths_per_batch": int(paths),
        "mpaths_per_s": round(paths / (dev_ms / 1e3) / 1e6, 2),
        "wall_mpaths_per_s": round(paths / wall / 1e6, 2),
        "batch_ms": round(dev_ms, 3),
        "sequential_ms": round(seq_ms, 3),
        "sequential_mpaths_per_s": round(paths / (seq_ms / 1e3) / 1e6, 2),
        "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
        "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
        "tuning": best,
        "graph": {"batch_ms": round(float(np.mean([r["batch_ms"] for r in gruns])), 3),
                  "mpaths_per_s": round(paths / (float(np.mean([r["batch_ms"] for r in gruns])) / 1e3)
                                        / 1e6, 2),
                  "latency_ms_p50": round(float(np.percentile(glat, 50)), 3),
                  "latency_ms_p99": round(float(np.percentile(glat, 99)), 3),
                  "graph_cached": gruns[-1]["graph_cached"]},
        "autotune_s": round(tune_s, 2),
        "chunks": runs[-1]["chunks"],
        "segments": runs[-1]["segments"],
        "graph_cached": runs[-1]["graph_cached"],
        "setup_s": round(setup_s, 2),
    }
    for ds in pool:
        ds.close()
    return out


def ltc_leg(args, ctx):
    """Config 3: chk_ltc_eval on ltc_points points x ltc_quads quads
    (ks = 1, the LTC part), FP64 arithmetic; pairs/s and the FP64 VALU
    roofline at 180 FLOP per (point, quad) pair (unclipped quad: corner
    transforms 72, normalisation 36, edge terms 60, combine 4, per-point
    setup ~6; add/mul/compare 1, FMA 2, div/sqrt/acos/sin 1)."""
    import numpy as np

    from chakumrender_amd import abi, host, scenes

    P = scenes.ltc_points(args.ltc_points, n_quads=args.ltc_quads)
    tris, rad = scenes.quads_to_tris(P["quads"], P["radiance"])
    ls = host.LightSource(ctx, tris, rad)
    k = args.ltc_points
    bufs = [abi.DevBuf.from_array(np.ascontiguousarray(a, np.float32)) for a in
            (P["x"], P["n"], P["v"], P["alpha"], np.ones(k, np.float32))]
    rgb = abi.DevBuf(12 * k)
    ls.eval_dev(bufs, k, rgb)
    abi.sync()

    def trial(reps=5):
        t0 = time.perf_counter()
        for _ in range(reps):
            ls.eval_dev(bufs, k, rgb)
        abi.sync()
        return (time.perf_counter() - t0) / reps

    trial(20)  # untimed: the clock ramps up over the first launches
    trials = [trial() for _ in range(5)]
    dt = min(trials)
    pairs = k * args.ltc_quads
    flop = 180.0 * pairs
    return {
        "workload": f"{k} points x {args.ltc_quads} quads ({2 * args.ltc_quads} triangles), "
                    "GGX LTC part (ks = 1), FP64 arithmetic, FP32 in/out",
        "ms": round(dt * 1e3, 3),
        "timing": "best of 5 trials of 5 launches (wall clock, synchronised) after 20 untimed "
                  "launches (the clock ramps up over the first)",
        "trials_ms": [round(t * 1e3, 3) for t in trials],
        "gpairs_per_s": round(pairs / dt / 1e9, 3),
        "roofline": {"bound": "fp64 valu", "achieved": round(flop / dt / 1e12, 3),
                     "peak": 78.6, "unit": "TFLOP/s", "frac": round(flop / dt / 1e12 / 78.6, 4),
                     "work_model": "180 FLOP per (point, quad) pair"},
    }


def mlp_leg(args, ctx):
    """The proposal network's inference (chk_proposal_eval -> k_mlp_eval) on
    mlp_queries vertices, the live vertices of one bounce of the
    time-to-RMSE leg's 32-spp pass at 480x270.  Work model: 28,800 FLOP per
    query (2 x (3 x 64 x 64 + 64 x 33) multiply-adds of the four layers)
    against the dense BF16 MFMA peak; HBM: 28 B of inputs in, 128 B of pmf
    and 4 B of mixing weight out per query."""
    import numpy as np

    from chakumrender_amd import abi, host

    k = args.mlp_queries
    pr = host.Proposal(ctx, 32, seed=5)
    s = np.random.default_rng(7).random((k, 7), dtype=np.float32)
    sd = abi.DevBuf.from_array(s)
    probs, mix = abi.DevBuf(128 * k), abi.DevBuf(4 * k)
    L = pr.L

    def run():
        abi.check(L.chk_proposal_eval(pr.h, sd.ptr, k, probs.ptr, mix.ptr, None, None),
                  "chk_proposal_eval")

    run()
    abi.sync()

    def trial(reps=10):
        t0 = time.perf_counter()
        for _ in range(reps):
            run()
        abi.sync()
        return (time.perf_counter() - t0) / reps

    t_w = time.perf_counter()  # untimed: the clock ramps up over the first 0.1 s
    while time.perf_counter() - t_w < 0.1:
        trial(20)
    trials = [trial() for _ in range(5)]
    dt = min(trials)
    flop = 28800.0 * k
    return {
        "workload": f"{k} queries, 7 inputs -> 64-wide encoding -> 3 x 64 ReLU -> 33 outputs, "
                    "BF16 MFMA with FP32 accumulation, 32-light softmax + mixing sigmoid",
        "ms": round(dt * 1e3, 4),
        "timing": "best of 5 trials of 10 launches (wall clock, synchronised) after 0.1 s of "
                  "untimed launches",
        "trials_ms": [round(t * 1e3, 4) for t in trials],
        "gqueries_per_s": round(k / dt / 1e9, 3),
        "hbm_gbs": round(160.0 * k / dt / 1e9, 1),
        "roofline": {"bound": "mfma", "achieved": round(flop / dt / 1e12, 2), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(flop / dt / 1e12 / 2500.0, 4),
                     "work_model": "28,800 FLOP per query"},
    }


def direct_leg(args, ctx, sc, ds):
    """The reference's shaders through chk_render_shade on the main scene:
    its shipped shade() (shadeDirectOnly with
    getLightFromMeshSource_Analytic_MISCV, shader.cpp:100-109,339-363), the
    same with getLightFromAllSources_MIS, shadeExplicit (MISCV at every
    vertex, shader.cpp:111-207) and shadeImplicit (shader.cpp:276-337).  The
    direct-lighting estimators take 10 area samples per mesh light + 60 BSDF
    samples per shading point, FP64.  The oracle's FP64 restatement of the
    same render runs beside each on every direct_cpu_row_step-th row."""
    import numpy as np

    from chakumrender_amd import abi, host

    out = {"workload": f"icosphere {sc.n_tris} tris, {sc.lights.shape[0]} mesh lights, "
                       f"{args.width}x{args.height} {args.direct_spp} spp, shadeDirectOnly; "
                       f"{10 * sc.lights.shape[0] + 60} MIS samples per camera hit, FP64"}
    params = dict(spp=args.direct_spp, seed_lo=0x5EED0009)
    image = host.Image(args.width, args.height)
    rows_ = (("miscv", abi.SHADE_DIRECT_ONLY, abi.DIRECT_MISCV),
             ("mis", abi.SHADE_DIRECT_ONLY, abi.DIRECT_MIS),
             ("explicit_miscv", abi.SHADE_EXPLICIT, abi.DIRECT_MISCV),
             ("implicit", abi.SHADE_IMPLICIT, abi.DIRECT_MIS))
    for name, shader, est in rows_:
        dp = host.direct_params(est)
        host.render_shade(ds, params, shader, dp, image=image, download=False)  # warm-up
        abi.sync()
        reps = 2
        t0 = time.perf_counter()
        for _ in range(reps):
            _, st = host.render_shade(ds, params, shader, dp, image=image, download=False)
        abi.sync()
        dt = (time.perf_counter() - t0) / reps
        paths = args.width * args.height * args.direct_spp
        row = {"ms": round(dt * 1e3, 2), "mpaths_per_s": round(paths / dt / 1e6, 3),
               "shadow_rays_per_s_g": round(st["rays_shadow"] / dt / 1e9, 3),
               "trace_launches": st["extend_launches"], "chunks": st["chunks"]}
        if shader == abi.SHADE_DIRECT_ONLY:
            row["samples_per_s_g"] = round(paths * (10 * sc.lights.shape[0] + 60) / dt / 1e9, 3)
        if not args.no_cpu_baseline:
            import oracle

            D = oracle.DirectScene(sc)
            threads = cpu_threads()
            t1 = time.perf_counter()
            D.render_shade(shader, params, dp, row_step=args.direct_cpu_row_step, threads=threads)
            ct = time.perf_counter() - t1
            rows = len(range(0, args.height, args.direct_cpu_row_step))
            cpaths = rows * args.width * args.direct_spp
            row["cpu_baseline"] = {"value": round(cpaths / ct / 1e6, 4), "unit": "Mpaths/s",
                                   "cores": threads, "kind": "port",
                                   "sample": f"every {args.direct_cpu_row_step}th row "
                                             f"({rows} rows x {args.width} px x {args.direct_spp} spp)",
                                   "seconds": round(ct, 2)}
            row["vs_cpu"] = round(row["mpaths_per_s"] / row["cpu_baseline"]["value"], 1)
        out[name] = row
    return out


def sources_leg(args, ctx):
    """Spheres and the other light sources through chk_render_shade with
    getLightFromAllSources_MIS (shadeDirectOnly, shader.cpp:339-363): the
    reference's own scene1.xml (sphere walls and balls, a sphere source;
    cone sampling) at its 512 x 384, and a 1080p room holding every kind
    (spheres, a sphere source, a triangle light, point sources, mesh lights
    and an 8-bit sky environment map).  The oracle's FP64 render of the same
    runs beside each on a bounded sample of rows."""
    from chakumrender_amd import abi, host, scene_xml, scenes

    import numpy as np

    G = np.load(ROOT / "tests" / "golden" / "scene_xml.npz")
    s1 = scene_xml.load_scene_xml(str(G["scene1_xml"]))
    room = scenes.sources_room(width=args.width, height=args.height)
    room.env = scenes.EnvMap(scenes.sky_image(), 1, (1.0, 0.9, 0.8), 2.0)
    legs = (("scene1_mis", s1, args.sources_spp, args.sources_cpu_row_step,
             f"scene1.xml (reference): {len(s1.spheres)} spheres incl. the 1e5-radius walls, "
             f"1 sphere source, {s1.camera.width}x{s1.camera.height} {args.sources_spp} spp, "
             "shadeDirectOnly + getLightFromAllSources_MIS (10 cone + 60 BSDF samples), FP64"),
            ("all_sources_mis", room, 1, args.sources_cpu_row_step * 2,
             f"room: 3 spheres, 1 sphere source, 1 triangle light, 2 point sources, 2 mesh "
             f"lights, 64x32 sky map; {args.width}x{args.height} 1 spp, shadeDirectOnly + "
             "getLightFromAllSources_MIS (110 samples + 2 point lights per hit), FP64"))
    out = {}
    dp = host.direct_params(abi.DIRECT_MIS)
    for name, sc, spp, row_step, workload in legs:
        t0 = time.perf_counter()
        ds = abi.DeviceScene(ctx, sc)
        setup = time.perf_counter() - t0
        W, H = sc.camera.width, sc.camera.height
        params = dict(spp=spp, seed_lo=0x5EED0013)
        image = host.Image(W, H)
        host.render_shade(ds, params, abi.SHADE_DIRECT_ONLY, dp, image=image, download=False)
        abi.sync()
        reps = 2
        t0 = time.perf_counter()
        for _ in range(reps):
            _, st = host.render_shade(ds, params, abi.SHADE_DIRECT_ONLY, dp, image=image,
                                      download=False)
        abi.sync()
        dt = (time.perf_counter() - t0) / reps
        paths = W * H * spp
        row = {"workload": workload, "ms": round(dt * 1e3, 2),
               "mpaths_per_s": round(paths / dt / 1e6, 3),
               "shadow_rays_per_s_g": round(st["rays_shadow"] / dt / 1e9, 3),
               "scene_setup_s": round(setup, 3)}
        if not args.no_cpu_baseline:
            import oracle

            D = oracle.DirectScene(sc)
            threads = cpu_threads()
            t1 = time.perf_counter()
            D.render_shade(abi.SHADE_DIRECT_ONLY, params, dp, row_step=row_step, threads=threads)
            ct = time.perf_counter() - t1
            rows = len(range(0, H, row_step))
            row["cpu_baseline"] = {"value": round(rows * W * spp / ct / 1e6, 4), "unit": "Mpaths/s",
                                   "cores": threads, "kind": "port",
                                   "sample": f"every {row_step}th row ({rows} rows x {W} px x {spp} spp)",
                                   "seconds": round(ct, 2)}
            row["vs_cpu"] = round(row["mpaths_per_s"] / row["cpu_baseline"]["value"], 1)
        ds.close()
        out[name] = row
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Threads for the CPU baselines: the CPUs this process may run on
    (sched_getaffinity), or fewer when OMP_NUM_THREADS sets the host share
    (the GPU boxes allot each GPU 16 CPUs and preset it to that)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def cpu_info():
    """The host the CPU baselines run on, from /proc/cpuinfo and the
    scheduler: model, logical CPUs, physical cores, those this process may
    use, and the threads the baselines use."""
    info = {"cpu_model": cpu_model(), "logical_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)), "threads_used": cpu_threads(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        cores, phys = set(), None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "physical id":
                phys = v
            elif k == "core id":
                cores.add((phys, v))
            elif k == "flags" and "flags" not in info:
                info["avx512"] = "avx512f" in v.split()
        info["physical_cores"] = len(cores) or None
    except OSError:
        pass
    return info


def calibration_leg(args, sc):
    """The compiled reference (oracle/_ref, built from its own sources)
    against the FP64 restatement on identical inputs, one core each:
    BvhAccel::intersectT over the reference's midpoint BVH (the restatement
    without t-pruning, as the reference) on camera rays of the main scene,
    and getLightFromMeshSource_Analytic at points on it.  The ratio says how
    the port's CPU baseline compares with the reference's own code."""
    import numpy as np

    import oracle
    from chakumrender_amd import abi, host, scenes

    R = oracle.ref_lib()
    if R is None:
        return {"skipped": "oracle/_ref/libchkref.so not built"}
    L = oracle.lib()
    L.orc_set_threads(1)
    out = {"cores": 1, "cpu_model": cpu_model()}
    tris = np.ascontiguousarray(sc.tris, np.float64)
    pos, d0, cx, cy = sc.camera.basis()
    rng = np.random.default_rng(11)
    m = args.calib_rays
    u = rng.random((m, 2))
    dirs = cx * (u[:, :1] - 0.5) + cy * (u[:, 1:] - 0.5) + d0
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rays = np.ascontiguousarray(np.c_[np.repeat(pos[None], m, 0), dirs])
    ids_r, t_r, n_r = np.empty(m, np.int32), np.empty(m), np.empty((m, 3))
    t0 = time.perf_counter()
    R.ref_bvh_closest(tris, tris.shape[0], rays, 0, ids_r, t_r, n_r)  # the build alone
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    R.ref_bvh_closest(tris, tris.shape[0], rays, m, ids_r, t_r, n_r)
    t_ref = time.perf_counter() - t0 - t_build
    bvh = oracle.RefBvh(tris)
    t0 = time.perf_counter()
    ids_p = bvh.closest(rays, prune=False)["id"]
    t_port = time.perf_counter() - t0
    out["intersectT"] = {"rays": m, "triangles": int(tris.shape[0]),
                         "reference_us_per_ray": round(t_ref / m * 1e6, 3),
                         "port_us_per_ray": round(t_port / m * 1e6, 3),
                         "reference_over_port": round(t_ref / t_port, 3),
                         "ids_equal": float((ids_r == ids_p).mean())}
    x, n, d, prim = scenes.shading_points(sc, args.calib_points, seed=12,
                                          eye=tuple(float(v) for v in pos))
    Rd = oracle.RefDirect(sc)
    t0 = time.perf_counter()
    ref_v, _ = Rd.stats(0, x, n, d, prim, reps=1, batches=1)
    t_ref = time.perf_counter() - t0
    D = oracle.DirectScene(sc)
    t0 = time.perf_counter()
    port_v = D.direct_light(host.direct_params(abi.DIRECT_ANALYTIC), x, n, d, prim, threads=1)
    t_port = time.perf_counter() - t0
    out["getLightFromMeshSource_Analytic"] = {
        "points": int(len(prim)), "mesh_lights": int(sc.lights.shape[0]),
        "reference_us_per_point": round(t_ref / len(prim) * 1e6, 3),
        "port_us_per_point": round(t_port / len(prim) * 1e6, 3),
        "reference_over_port": round(t_ref / t_port, 3),
        "max_rel_diff": float(np.max(np.abs(ref_v - port_v) / np.maximum(np.abs(ref_v), 1e-12)))}
    L.orc_set_threads(cpu_threads())
    return out


def progressive(render_pass, ref, rows, pass_spp, max_spp, target, sync, W, H):
    """Render whole-frame passes of pass_spp until the RMSE against ref (the
    reference's rows `rows`) reaches target or max_spp is reached; returns
    ([(gpu seconds so far, spp, rmse)], the final image's sums).  Only the
    render calls are timed (the RMSE readback is not)."""
    from chakumrender_amd import host

    img = host.Image(W, H)
    t = 0.0
    curve = []
    a = None
    for k in range(max_spp // pass_spp):
        sync()
        t0 = time.perf_counter()
        render_pass(k * pass_spp, img)
        sync()
        t += time.perf_counter() - t0
        n = (k + 1) * pass_spp
        a, _ = img.get()
        err = float(((a[rows] / n - ref) ** 2).mean() ** 0.5)
        curve.append((t, n, err))
        if err <= target:
            break
    return curve, a


def time_to(curve, target):
    """Seconds at which the curve crosses target (log-linear interpolation
    between passes), or None."""
    import math

    prev = None
    for t, _, e in curve:
        if e <= target:
            if prev is None:
                return t
            t0, e0 = prev
            f = (math.log(e0) - math.log(target)) / (math.log(e0) - math.log(e))
            return t0 + f * (t - t0)
        prev = (t, e)
    return None


def rmse_at(curve, t_query):
    """RMSE of the curve at time t_query (log-linear interpolation)."""
    import math

    if not curve or t_query < curve[0][0]:
        return None
    prev = curve[0]
    for p in curve[1:]:
        if p[0] >= t_query:
            f = (t_query - prev[0]) / max(p[0] - prev[0], 1e-12)
            return math.exp(math.log(prev[2]) + f * (math.log(p[2]) - math.log(prev[2])))
        prev = p
    return None


def rmse_leg(args, ctx):
    """Time-to-RMSE of plain MIS and of the learned proposal on the config-5
    scene (2M triangles, 32 quad lights) at the config's 1920 x 1080.

    The reference is plain MIS at rmse_ref_spp samples with FP64
    accumulation, rendered for every rmse_row_step-th row of the frame (the
    render's window: each row exactly as in a whole-frame render); the
    RMSE of the whole-frame passes is taken over those rows.  A systematic
    sample of 1/16 of the rows is an unbiased estimate of the whole-frame
    RMSE; --rmse-full-reference renders every row to measure both."""
    import numpy as np

    from chakumrender_amd import abi, host, scenes

    W, H = args.rmse_width, args.rmse_height
    sc = scenes.cornell_large(total_tris=args.rmse_tris, n_lights=args.rmse_lights, width=W, height=H)
    ds = abi.DeviceScene(ctx, sc)
    nl = ds.info()["n_lights"]
    base = dict(max_depth=args.rmse_depth)
    step = max(1, args.rmse_row_step)
    rows = np.arange(step // 2, H, step)

    def reference(rlist):
        img = host.Image(W, H)
        for r in rlist:
            host.render(ds, dict(base, spp=args.rmse_ref_spp, seed=0x5EED0F00,
                                 window=(0, int(r), W, 1)), image=img, stats=False, download=False)
        return img.get()[0] / args.rmse_ref_spp

    t0 = time.perf_counter()
    ref_full = reference(rows)
    ref = ref_full[rows]
    ref_s = time.perf_counter() - t0
    print(f"# rmse leg: {sc.n_tris} tris, {nl} lights, {W}x{H}, reference {args.rmse_ref_spp} spp "
          f"over {len(rows)} rows in {ref_s:.1f} s, mean {ref.mean():.4f}", file=sys.stderr, flush=True)
    ref_check = None
    if not args.no_cpu_baseline:
        # the reference rows near the middle of the frame against the FP64
        # oracle's own render of the same rows with independent keys:
        # per-pixel z from the oracle's sample variance (the GPU reference's
        # is 1 / 1024 of it at most)
        import oracle

        mid = rows[(rows >= H // 2 - 4 * step) & (rows < H // 2 + 4 * step)]
        n = args.rmse_check_spp
        t1 = time.perf_counter()
        acc, sq, _, _ = oracle.render(sc, dict(base, spp=n, seed=0x5EED0F77),
                                      rows=(int(mid[0]), int(mid[-1]) + 1), row_step=step,
                                      threads=cpu_threads(), sumsq=True)
        m = acc[mid] / n
        var = np.maximum(sq[mid] / n - m * m, 0) / n
        z = np.abs(ref_full[mid] - m) / np.sqrt(var + 1e-30)
        band = ref_full[mid]
        ref_check = {"rows": [int(v) for v in mid], "cpu_spp": n,
                     "within_3sigma": round(float((z <= 3).mean()), 4),
                     "band_mean_gpu_ref": round(float(band.mean()), 6),
                     "band_mean_cpu": round(float(m.mean()), 6),
                     "band_mean_z": round(float(abs(band.mean() - m.mean()) /
                                                np.sqrt(var.sum() / var.size ** 2)), 2),
                     "cpu_seconds": round(time.perf_counter() - t1, 2)}

    def plain_pass(off, img):
        host.render(ds, dict(base, spp=args.rmse_pass_spp, sample_offset=off, seed=0x5EED0005),
                    image=img, stats=False, download=False)

    prop = host.Proposal(ctx, nl, seed=0x5EED0005)
    train = dict(train_spp=args.train_spp, batch=args.train_batch,
                 steps_per_bounce=args.train_steps, lr=args.train_lr)

    def learned_pass(off, img):
        host.render_learned(ds, dict(base, spp=args.rmse_pass_spp, sample_offset=off,
                                     seed=0x5EED0005), prop, train=train, image=img,
                            stats=False, download=False)

    cp, img_p = progressive(plain_pass, ref, rows, args.rmse_pass_spp, args.rmse_max_spp,
                            args.rmse_target, abi.sync, W, H)
    cl, img_l = progressive(learned_pass, ref, rows, args.rmse_pass_spp, args.rmse_max_spp,
                            args.rmse_target, abi.sync, W, H)
    tp, tl = time_to(cp, args.rmse_target), time_to(cl, args.rmse_target)
    # equal time: the RMSE each reaches at the time the faster one stops
    t_eq = min(cp[-1][0], cl[-1][0])
    out = {
        "workload": f"cornell-style {sc.n_tris} tris, {nl} quad lights, {W}x{H}, depth "
                    f"{args.rmse_depth}; reference plain MIS {args.rmse_ref_spp} spp FP64",
        "target": args.rmse_target,
        "rmse": f"per-channel linear radiance over every {step}th row of the frame "
                f"({len(rows)} rows x {W} px), whole-frame passes of {args.rmse_pass_spp} spp",
        "plain_s": None if tp is None else round(tp, 4),
        "learned_s": None if tl is None else round(tl, 4),
        "ratio_learned_over_plain": None if (tp is None or tl is None) else round(tl / tp, 4),
        "plain_spp_at_target": cp[-1][1] if tp is not None else None,
        "learned_spp_at_target": cl[-1][1] if tl is not None else None,
        "equal_time": {"seconds": round(t_eq, 4), "plain_rmse": rmse_at(cp, t_eq),
                       "learned_rmse": rmse_at(cl, t_eq)},
        "plain_mpaths_per_s": round(W * H * cp[-1][1] / cp[-1][0] / 1e6, 2),
        "learned_mpaths_per_s": round(W * H * cl[-1][1] / cl[-1][0] / 1e6, 2),
        "train": dict(train, pass_spp=args.rmse_pass_spp),
        "reference_rows": int(len(rows)),
        "reference_seconds": round(ref_s, 2),
        "reference_check": ref_check,
    }
    if args.rmse_full_reference:
        # every row of the reference: the final images' RMSE over the whole
        # frame against the row-sample estimate
        t1 = time.perf_counter()
        full = reference(range(H))
        out["full_reference"] = {
            "seconds": round(time.perf_counter() - t1, 1),
            "plain_rmse_frame": float(((img_p / cp[-1][1] - full) ** 2).mean() ** 0.5),
            "plain_rmse_rows": cp[-1][2],
            "learned_rmse_frame": float(((img_l / cl[-1][1] - full) ** 2).mean() ** 0.5),
            "learned_rmse_rows": cl[-1][2],
            "reference_rows_equal": bool(np.array_equal(full[rows], ref)),
        }
    prop.close()
    ds.close()
    return out


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus above 1 is not supported "
                                   "(a render shards by pixel or sample range across GPUs)"}))
        sys.exit(2)

    from chakumrender_amd import abi, host, scenes

    if args.calibrate_copy:
        nb = args.calibrate_copy << 20
        a, b = abi.DevBuf.zeros(nb), abi.DevBuf(nb)
        for _ in range(4):
            abi.check(abi.load().chk_memcpy(b.ptr, a.ptr, nb, 2, None), "chk_memcpy d2d")
        abi.sync()
        print(json.dumps({"copy_bytes_per_dispatch": nb, "dispatches": 4}))
        return

    t0 = time.time()
    sc = scenes.icosphere_scene(freq=args.freq, width=args.width, height=args.height)
    ctx = abi.Context(0)
    ds = abi.DeviceScene(ctx, sc)
    info = ds.info()
    params = dict(spp=args.spp, max_depth=args.depth, seed=0x5EED0002)
    image = host.Image(args.width, args.height)
    print(f"# scene {sc.n_tris} tris, {info['n_nodes']} nodes, stack {info['stack_size']}, "
          f"build {info['build_ms']:.0f} ms, setup {time.time() - t0:.1f} s", file=sys.stderr, flush=True)

    # work model: node visits / triangle tests of this exact workload (untimed)
    _, cnt = host.render(ds, params, image=image, count=True, download=False)
    print(f"# counted: {cnt['rays_extend']} extend rays, {cnt['rays_shadow']} shadow rays, "
          f"{cnt['node_visits']} node visits, {cnt['tri_tests']} tri tests", file=sys.stderr, flush=True)

    prof = dict(params, flags=abi.RENDER_PROFILE)
    for _ in range(args.warmup):
        host.render(ds, params, image=image, stats=False, download=False)
    abi.sync()
    ext_ms = shade_ms = shadow_ms = other_ms = 0.0
    launches = 0
    shadow_launches = cnt["extend_launches"] - cnt["chunks"]  # no shadow launch after the last bounce
    t_start = time.perf_counter()
    for _ in range(args.steps):
        _, st = host.render(ds, prof, image=image, download=False)
        ext_ms += st["extend_ms"]
        shade_ms += st["shade_ms"]
        shadow_ms += st["shadow_ms"]
        other_ms += st["other_ms"]
        launches += st["extend_launches"]
    abi.sync()
    elapsed = time.perf_counter() - t_start
    # after the timed steps: one frame with its chunks one after another, so
    # each launch has the GPU to itself (the timed steps overlap consecutive
    # chunks on two streams, and a launch then shares the GPU)
    _, st1 = host.render(ds, dict(params, flags=abi.RENDER_PROFILE | abi.RENDER_ONE_STREAM),
                         image=image, download=False)
    abi.sync()

    paths = args.width * args.height * args.spp
    ms_per_step = elapsed * 1e3 / args.steps
    value = paths * args.steps / elapsed / 1e6
    rays_total = cnt["rays_extend"] + cnt["rays_shadow"]
    mrays = rays_total * args.steps / elapsed / 1e6

    # roofline of the closest-hit kernel: algorithmic bytes per launch over
    # its average launch duration with the GPU to itself (HIP events around
    # each launch of the one-stream frame, on its stream).  In the timed
    # steps two chunks overlap, and a launch's event span then also holds
    # the wait for the other stream's kernels.
    alg_bytes_total = (cnt["node_visits"] * NODE_BYTES + cnt["tri_tests"] * TRI_BYTES
                       + cnt["rays_extend"] * RAY_BYTES)
    alg_bytes_launch = alg_bytes_total / cnt["extend_launches"]
    avg_launch_s = st1["extend_ms"] / 1e3 / st1["extend_launches"]
    span_launch_s = ext_ms / 1e3 / launches
    achieved = alg_bytes_launch / avg_launch_s / 1e9
    traffic, meas = None, {}
    if args.traffic and Path(args.traffic).exists():
        meas = json.loads(Path(args.traffic).read_text())
        traffic = meas.get("bytes_per_launch")
    # any-hit: the same work model over the shadow rays
    sh_bytes_launch = ((cnt["shadow_node_visits"] * NODE_BYTES + cnt["shadow_tri_tests"] * TRI_BYTES
                        + cnt["rays_shadow"] * RAY_BYTES) / max(1, shadow_launches))
    sh_launch_s = st1["shadow_ms"] / 1e3 / max(1, st1["extend_launches"] - st1["chunks"])

    line = {
        "metric": "Mpaths/s",
        "value": round(value, 3),
        "unit": "Mpaths/s",
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
            "workload": f"icosphere {sc.n_tris} tris (freq {args.freq}, 3-octave noise 0.1), ground, "
                        f"8 quad lights; {args.width}x{args.height} {args.spp} spp, depth {args.depth}, "
                        "balance-heuristic MIS (1 light + 1 BSDF sample/vertex), FP64 accumulation",
            "width": args.width, "height": args.height, "spp": args.spp, "max_depth": args.depth,
            "n_tris": sc.n_tris, "paths_per_step": paths,
        },
        "mrays_per_s": round(mrays, 3),
        "rays_per_path": round(rays_total / paths, 4),
        "kernel_ms_per_step": {"closest_hit": round(ext_ms / args.steps, 3),
                               "shade": round(shade_ms / args.steps, 3),
                               "any_hit": round(shadow_ms / args.steps, 3),
                               "camera_accumulate": round(other_ms / args.steps, 3),
                               "note": "HIP-event spans of each phase on its chunk's stream; "
                                       "consecutive chunks overlap on two streams, so the spans "
                                       "add up to more than ms_per_step"},
        "kernel_ms_one_stream": {"closest_hit": round(st1["extend_ms"], 3),
                                 "shade": round(st1["shade_ms"], 3),
                                 "any_hit": round(st1["shadow_ms"], 3),
                                 "camera_accumulate": round(st1["other_ms"], 3),
                                 "frame_ms": round(st1["extend_ms"] + st1["shade_ms"] +
                                                   st1["shadow_ms"] + st1["other_ms"], 3),
                                 "note": "one frame after the timed steps with its chunks one "
                                         "after another (CHK_RENDER_ONE_STREAM)"},
        "roofline": {
            "kernel": "closest-hit traversal: k_trace<closest> for the camera rays, "
                      "k_trace_stream<closest> (ray replacement) for the later bounces "
                      "(chakumrender_amd/csrc/chk_trace.hip)",
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "bytes_per_launch": int(alg_bytes_launch),
            "avg_launch_ms": round(avg_launch_s * 1e3, 4),
            "timing": f"HIP events around each of the {st1['extend_launches']} launches of one frame "
                      "rendered right after the timed steps with its chunks one after another "
                      "(CHK_RENDER_ONE_STREAM), so each launch has the GPU to itself",
            "work_model": f"{NODE_BYTES} B/node visit + {TRI_BYTES} B/triangle test + {RAY_BYTES} B/ray",
            "what_the_frac_is": "algorithmic node/triangle/ray bytes per second against the HBM peak; "
                                "the caches serve almost all of them (measured below)",
            "timed_steps": {
                "avg_launch_span_ms": round(span_launch_s * 1e3, 4),
                "note": "HIP-event span per closest-hit launch in the timed steps, where two chunks "
                        "overlap on two streams: it holds the wait for the other stream's kernels "
                        "as well as the kernel"},
            "measured": None if not meas else {
                "hbm_bytes_per_launch": traffic,
                "hbm_gbs": round(traffic / avg_launch_s / 1e9, 1),
                "hbm_frac": round(traffic / avg_launch_s / 1e9 / HBM_PEAK_GBS, 4),
                "l1_hit_rate": meas.get("l1_hit_rate"), "l2_hit_rate": meas.get("l2_hit_rate"),
                "valu_busy_pct": meas.get("valu_busy_pct"),
                "valu_utilization_pct": meas.get("valu_utilization_pct"),
                "waves_per_cu": meas.get("waves_per_cu"),
                "valu_issue": None if not meas.get("valu_wave_insts_per_launch") else {
                    "wave_insts_per_launch": meas["valu_wave_insts_per_launch"],
                    "achieved_ginst_s": round(meas["valu_wave_insts_per_launch"] / avg_launch_s / 1e9, 1),
                    "peak_ginst_s": meas["valu_issue_peak_ginst_s"],
                    "frac": round(meas["valu_wave_insts_per_launch"] / avg_launch_s / 1e9
                                  / meas["valu_issue_peak_ginst_s"], 3),
                    "peak_note": meas.get("valu_issue_peak_note")},
                "bound_by": "divergent VALU issue and dependent-fetch latency with L1/L2-resident "
                            "nodes (profiles/traversal_counters_icosphere100k_1080p.md)"},
        },
        "any_hit": {
            "kernel": "k_trace_stream<any> (ray replacement)",
            "rays_per_frame": cnt["rays_shadow"],
            "node_visits_per_ray": round(cnt["shadow_node_visits"] / max(1, cnt["rays_shadow"]), 2),
            "tri_tests_per_ray": round(cnt["shadow_tri_tests"] / max(1, cnt["rays_shadow"]), 2),
            "closest_node_visits_per_ray": round(cnt["node_visits"] / max(1, cnt["rays_extend"]), 2),
            "closest_tri_tests_per_ray": round(cnt["tri_tests"] / max(1, cnt["rays_extend"]), 2),
            "avg_launch_ms": round(sh_launch_s * 1e3, 4),
            "achieved_gbs": round(sh_bytes_launch / sh_launch_s / 1e9, 1) if sh_launch_s > 0 else None,
            "grays_per_s": round(cnt["rays_shadow"] / (st1["shadow_ms"] / 1e3) / 1e9, 3)
                           if st1["shadow_ms"] > 0 else None,
            "timing": "the one-stream frame, as the closest-hit roofline",
        },
    }

    if not args.no_rmse:
        line["time_to_rmse"] = rmse_leg(args, ctx)
    if not args.no_batch:
        line["batch"] = batch_leg(args, ctx)
    if not args.no_ltc:
        line["ltc"] = ltc_leg(args, ctx)
    if not args.no_mlp:
        line["mlp"] = mlp_leg(args, ctx)
    if not args.no_direct:
        line["direct"] = direct_leg(args, ctx, sc, ds)
    if not args.no_sources:
        line["sources"] = sources_leg(args, ctx)
    if not args.no_cpu_baseline and not args.no_calibration:
        line["cpu_calibration"] = calibration_leg(args, sc)

    if not args.no_cpu_baseline:
        import numpy as np

        import oracle

        threads = cpu_threads()
        rows = (0, args.height)
        # independent keys: the oracle's frame is also the statistical check
        # of the GPU's headline frame
        t1 = time.perf_counter()
        acc_c, sq_c, cpaths, crays = oracle.render(sc, dict(params, spp=args.cpu_spp,
                                                            seed=params["seed"] ^ 0xC0FFEE),
                                                   rows=rows, row_step=args.cpu_row_step,
                                                   threads=threads, sumsq=True)
        dt = time.perf_counter() - t1
        n_rows = len(range(0, args.height, args.cpu_row_step))
        # the headline frame again with sums of squares (untimed), same keys
        # as the timed renders
        acc_g, sq_g, _ = host.render(ds, dict(params, flags=abi.RENDER_SUMSQ))
        r = slice(0, args.height, args.cpu_row_step)
        ng, nc = args.spp, args.cpu_spp
        mg, mc = acc_g[r] / ng, acc_c[r] / nc
        se = np.sqrt(np.maximum(sq_g[r] / ng - mg * mg, 0) / ng +
                     np.maximum(sq_c[r] / nc - mc * mc, 0) / nc)
        with np.errstate(all="ignore"):
            z = np.where(se > 0, np.abs(mg - mc) / se, np.where(np.abs(mg - mc) <= 1e-6, 0.0, np.inf))
        line["parity_vs_oracle"] = {
            "what": "headline frame (the timed render's keys) against the FP64 oracle's frame with "
                    "independent keys, per pixel channel, sigma from both sample variances",
            "pixel_channels": int(z.size),
            "within_3sigma": round(float(np.mean(z <= 3)), 5),
            "max_z": round(float(z.max()), 2),
            "image_mean_gpu": round(float(mg.mean()), 6),
            "image_mean_oracle": round(float(mc.mean()), 6),
        }
        line["cpu_baseline"] = {
            "value": round(cpaths / dt / 1e6, 4),
            "unit": "Mpaths/s",
            "cores": threads,
            "kind": "port",
            "sample": f"FP64 oracle, every {args.cpu_row_step}th row ({n_rows} rows x {args.width}) "
                      f"at {args.cpu_spp} spp = {cpaths} paths, {crays} rays, {dt:.1f} s",
            "host": cpu_info(),
        }
        line["speedup_vs_cpu"] = round(value / line["cpu_baseline"]["value"], 2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
