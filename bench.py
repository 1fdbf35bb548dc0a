#!/usr/bin/env python3
"""Benchmark of the gfx950 Radiance path tracer on one MI355X.

Workload (default): the large scene -- 40 seeded instances of a displaced
geodesic icosphere (25,920 triangles each, 1,036,800 in all) with random
rotations and scales on a 20 x 20 ground plane, 4 quad area lights, a 50/50
mix of Lambert and Beckmann materials; 1920x1080, 64 spp, maxBounce 8, FP32
traversal/shading, FP64 per-pixel accumulation, material-sorted shading
queues. One step = one full frame (132,710,400 camera paths). Scene data
(geometry, BVH, light tables) is resident in HBM before the timed region;
the accumulators stay on the device.

Prints one JSON line. `value` is the plain path's paths/s. `config.cached`
is the wall-clock to a fixed RMSE against the FP64 oracle fixture, plain
against the radiance cache: first in the mode the workload specifies (paths
end at their second diffuse or rough vertex, 2^16 training samples per frame
from 1/16 of paths, Adam lr 1e-2), then in the first-vertex mode, against
both the clamped and the unclamped RMSE of the plain 64-spp frame.
`roofline` is for the closest-hit traversal of one bounce's ray queue (a
k_trace launch for camera rays, a k_trace_tail launch and its follow-ups for
later bounces): algorithmic bytes per query = 44 B of queue I/O + node
visits x node bytes + triangle tests x 48 B, with the visit counts measured
by the same kernels on a 4-spp sample of the frame, divided by the mean
duration of one queue's traversal from HIP events recorded around its
launches on the stream they run on, over the timed steps (four streams
share the GPU). `roofline.one_stream` repeats the measurement on frames
whose chunks run on one stream, after the timed region. `cpu_baseline` is the oracle's FP32 restatement of the
reference CPU path (CPUMulticore thread scheme, reference BVH and traversal,
release flags) on a bounded sample of the same frame on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cuda-raytracer_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "paths/s + wall-clock to fixed RMSE vs FP64 oracle (plain vs cached) on 1 MI355X"
HBM_PEAK_GBS = 8000.0
QUEUE_BYTES_PER_QUERY = 44  # o, d in (24 B); t, u, v, hit code out (16 B); bucket entry (4 B)
TRI_BYTES = 48


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(scene, target_seconds):
    import numpy as np
    import oracle_lib as ol
    orc = ol.OracleScene(scene, "f32", traversal=0)
    p = scene.params(seed=1)
    w, h = p.width, p.height
    # the reference's hardware_concurrency() - 1 tracer threads
    # (src/core/tracer/cpuMulticore.cu:19-21), over the CPUs this process may
    # use (at most 16: one GPU's share of the host)
    ncpu = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 2))
    nthreads = max(1, ncpu - 1)
    # probe on 8 rows at 1 spp, then size the sample (rows, then samples per
    # pixel of the frame's own strata) to ~target_seconds
    rows = np.arange(0, h, max(1, h // 8))[:8]
    pix = (rows[:, None] * w + np.arange(w)[None, :]).ravel()
    _, _, st = orc.render(p, 0, 1, nthreads, pix)
    rate = pix.size / max(st["seconds"], 1e-6)
    want = target_seconds * rate
    nrows = int(min(h, max(8, want / w)))
    spp = int(min(p.spp, max(1, want / (nrows * w))))
    rows = np.linspace(0, h - 1, nrows).astype(np.int64)
    pix = (rows[:, None] * w + np.arange(w)[None, :]).ravel()
    _, _, st = orc.render(p, 0, spp, nthreads, pix)
    orc.free()
    paths = pix.size * spp
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": paths / st["seconds"],
        "unit": "paths/s",
        "cores": nthreads,
        "kind": "port",
        "sample": f"{nrows} of {h} rows x {w} px x {spp} spp ({paths} paths, {st['seconds']:.1f} s) of the "
                  f"same frame; FP32 restatement of CPUMulticoreTracer + RadianceKernel, reference BVH and "
                  f"traversal, -O3 -ffast-math -fno-finite-math-only -funroll-loops; "
                  f"{ncpu} logical CPUs usable, {nthreads} tracer threads",
        "cpu_model": model,
        "hardware_concurrency": ncpu,
        "host_logical_cpus": os.cpu_count(),
        "threads": nthreads,
    }


def oracle_fixtures(workload):
    """The committed converged FP64 oracle images of a workload: {"subset":
    a fixed 4096-pixel subset, "crop": the 64 x 64 block centred in the frame,
    "full": every pixel}, those that exist (tests/golden/oracle_<workload>
    [_crop|_full].npz), each with the workload's lens strata and 65536 samples
    per pixel."""
    out = {}
    for key, suffix in (("subset", ""), ("crop", "_crop"), ("full", "_full")):
        path = os.path.join(ROOT, "tests", "golden", f"oracle_{workload}{suffix}.npz")
        if os.path.exists(path):
            out[key] = path
    return out


def rmse_vs_oracle(workload, rgb, valid, width):
    """RMSE of the rendered frame against the committed converged FP64 oracle
    image (the full frame where its fixture exists, else the fixed pixel
    subset; same lens strata, 65536 samples per pixel), on linear radiance,
    unclamped and clamped to [0, 1] like the reference's 8-bit output."""
    import numpy as np
    fx = oracle_fixtures(workload)
    path = fx.get("full", fx.get("subset"))
    if path is None:
        return None
    f = np.load(path)
    pix = f["pixels"].astype(np.int64)
    ref = f["mean"]
    mine = rgb[pix] / np.maximum(valid[pix], 1)[:, None]
    d = mine - ref
    dc = np.clip(mine, 0, 1) - np.clip(ref, 0, 1)
    return {"unclamped": float(np.sqrt((d * d).mean())), "clamped": float(np.sqrt((dc * dc).mean())),
            "pixels": int(pix.size), "oracle_samples_per_pixel": int(f["seeds"]) * int(f["spp"])}


def fit_noise_bias(ms, mse, oracle_frames):
    """Least-squares fit of mse(m) = b^2 + sigma^2 / m over the frame counts
    ms; the oracle's own noise (sigma^2 / oracle_frames, an average of that
    many frames) is taken off b^2. Returns the signed bias b, sigma of one
    frame, the fitted b^2 and the oracle noise^2."""
    import numpy as np
    A = np.stack([np.ones(len(ms)), 1.0 / np.asarray(ms, float)], axis=1)
    (b2, s2), *_ = np.linalg.lstsq(A, np.asarray(mse, float), rcond=None)
    b2t = b2 - s2 / oracle_frames
    return {"bias": float(np.sign(b2t) * np.sqrt(abs(b2t))), "sigma_per_frame": float(np.sqrt(max(s2, 0.0))),
            "bias_squared_fit": float(b2), "oracle_noise_squared": float(s2 / oracle_frames)}


def fit_text(fit):
    return ", ".join(f"{k}: b {v['clamped']['bias']:.4f} / sigma {v['clamped']['sigma_per_frame']:.4f} clamped, "
                     f"b {v['unclamped']['bias']:.3f} / sigma {v['unclamped']['sigma_per_frame']:.3f}"
                     for k, v in fit.items() if isinstance(v, dict))


def rmse_fit(kr, scene, gs, workload, spp, nrc=None, frames=16):
    """Noise and bias against the FP64 oracle, separated by fitting
    RMSE^2(m) = b^2 + sigma^2 / m, where m independent frames of the
    workload's spp (seeds 101, 102, ...) are averaged: m = 1, 2, 4, 8, 16,
    i.e. 64 ... 1024 samples per pixel at 64 spp, every group of m frames
    disjoint from the others and RMSE^2(m) their mean. The reference's lens
    places sample k of a pixel at a fixed point of a ceil(sqrt(spp))^2 grid
    (no jitter; src/core/misc/lens.hpp:19,61-65), so renders at other spp
    converge to other expectations:
    only whole frames of the oracle's own spp share its expectation. The
    oracle image averages 1024 such frames, so its own noise sigma^2 / 1024
    is taken off b^2. Per oracle fixture (subset, crop, full), unclamped and
    clamped to [0, 1]; nrc: the cached path with that (trained) cache."""
    import numpy as np
    fx = {k: np.load(v) for k, v in oracle_fixtures(workload).items()}
    if not fx:
        return None
    cam = scene.camera_struct()
    imgs = {k: [] for k in fx}
    for i in range(frames):
        r = kr.Render(gs, cam, scene.params(seed=101 + i, spp=spp))
        if nrc is not None:
            r.attach_nrc(nrc)
        r.run(0, spp)
        r.sync()
        rgb, val = r.read()
        r.free()
        for k, f in fx.items():
            pix = f["pixels"].astype(np.int64)
            imgs[k].append(rgb[pix] / np.maximum(val[pix], 1)[:, None])
    out = {"frames": frames, "spp_per_frame": spp}
    ms = [m for m in (1, 2, 4, 8, 16, 32) if m <= frames]
    for k, f in fx.items():
        nseeds = int(f["seeds"])
        res = {"pixels": int(f["pixels"].size)}
        for key, g in (("unclamped", lambda x: x), ("clamped", lambda x: np.clip(x, 0.0, 1.0))):
            ref = g(f["mean"])
            im = np.stack([g(x) for x in imgs[k]])
            y = []
            for m in ms:
                groups = im[: (frames // m) * m].reshape(frames // m, m, *im.shape[1:]).mean(axis=1)
                y.append(float(((groups - ref) ** 2).mean()))
            res[key] = {"samples_per_pixel": [m * spp for m in ms], "rmse": [float(np.sqrt(v)) for v in y],
                        **fit_noise_bias(ms, y, nseeds)}
            # the noise of the mean of all the frames: a b^2 well below it is
            # not resolved by this many frames
            res[key]["noise_squared_of_all_frames"] = res[key]["sigma_per_frame"] ** 2 / frames
        out[k] = res
    return out


# The cache's two ways of ending paths (kr_nrc_config.flags) and their
# settings. "spec" is the workload's own statement of the cache: paths end at
# their second diffuse or rough vertex (GlossyBSDF roughness >= 0.3 counts as
# rough) and the cache stands for everything that vertex scatters (Mueller et
# al. 2021); Adam lr 1e-2; 4 steps of 16,384 samples = 2^16 training samples
# per frame from 1/16 of the paths, traced unbiased to maxBounce. "first"
# (KR_NRC_INDIRECT) deviates from it: paths end at their first diffuse or
# rough vertex after its light sample and the cache stands for what the rest
# of the path brings; 8 steps (2^17 samples) per frame and rough threshold
# 0.6 (glossy lobes below it keep their paths going). The first bounce
# carries most of the noise of these scenes, and on the large scene the
# second vertex is often never reached, so "first" removes more work.
CACHE_MODES = {
    "spec": {"flags": 0, "steps_per_frame": 4, "rough_threshold": 0.3, "learning_rate": 1e-2},
    "first": {"flags": 0x2, "steps_per_frame": 8, "rough_threshold": 0.6,  # flags: KR_NRC_INDIRECT
              "learning_rate": 1e-2},
}
CACHE_MODE_TEXT = {
    "spec": "paths end at their second diffuse or rough vertex (Beckmann roughness >= 0.3), which queries the "
            "cache for everything it scatters; 2^16 training samples per frame from 1/16 of paths traced "
            "unbiased to maxBounce; Adam lr 1e-2",
    "first": "deviating mode: paths end at their first diffuse or rough vertex (Beckmann roughness >= 0.6) "
             "after its light sample, and the cache stands for the rest of the path; 2^17 training samples "
             "per frame; Adam lr 1e-2",
}


def bench_cache(args, kr, scene, gs, p, dev):
    """Radiance cache on the workload against the plain path. The cache
    (1/16 of paths train, Adam steps of 16384 samples after each frame,
    renders query the moving average of decay 0.99 per step, targets clamped
    at the brightest emitter's radiance; ending mode, steps, rough threshold
    and learning rate per CACHE_MODES) is trained for a number of 4-spp
    warm-up frames, then frames of s = 4, 8, 16, 24, 32, 48, 64 spp (scaled to
    the workload's spp) are timed for both paths, with RMSE against the FP64
    oracle over 3 seeds. The targets are the plain path's RMSE at the
    workload's spp, clamped to [0, 1] (what the reference's 8-bit output
    shows) and unclamped; the cached wall-clock to a target is the time of
    the cached frame at the smallest spp from which on every cached frame
    reaches it (null when the cache's bias lies above it)."""
    import numpy as np
    pos = np.asarray(scene.positions, np.float64)
    lo, hi = pos.min(axis=0) - 0.01, pos.max(axis=0) + 0.01
    # training targets clamped at the brightest emitter's radiance: expected
    # reflected radiance stays below it, larger single-path targets are noise
    emax = max([max(c.get("color", [1, 1, 1])) * c.get("strength", 2.0)
                for t, c in ((v[0], v[1] if len(v) > 1 else {}) for v in scene.shaders.values()) if t == "Emission"]
               + [max(col) * k for _, col, k in scene.point_lights] + [1.0])
    cam = scene.camera_struct()
    modes = ["spec", "first"] if args.cache_mode == "both" else [args.cache_mode]

    fx = np.load(os.path.join(ROOT, "tests", "golden", f"oracle_{args.workload}.npz"))
    fpix, fref = fx["pixels"].astype(np.int64), fx["mean"]

    def frame(r, spp):
        r.reset()
        t0 = time.perf_counter()
        r.run(0, spp)
        r.sync()
        dt = time.perf_counter() - t0
        rgb, val = r.read()
        return dt, rgb[fpix] / np.maximum(val[fpix], 1)[:, None]

    def over_seeds(nrc, s):
        """Best time, RMSE against the oracle over 3 independent seeds (as
        MSE), and the bias: the error of the 3-seed mean minus its own noise,
        b^2 = |mean - ref|^2 - var / 3 (var from the spread of the seeds),
        unclamped and clamped to [0, 1]."""
        dts, imgs = [], []
        for seed in (11, 12, 13):
            r = kr.Render(gs, cam, scene.params(seed=seed, spp=s))
            if nrc is not None:
                r.attach_nrc(nrc)
            frame(r, s)
            dt, img = frame(r, s)
            queries = r.stats().cache_queries
            r.free()
            dts.append(dt)
            imgs.append(img)
        out = {"seconds": min(dts)}
        if nrc is not None:
            out["cache_queries"] = queries
        for key, f in (("", lambda x: x), ("_clamped", lambda x: np.clip(x, 0.0, 1.0))):
            im = np.stack([f(x) for x in imgs])
            ref = f(fref)
            out["rmse" + key] = float(np.sqrt(((im - ref) ** 2).mean()))
            var = im.var(axis=0, ddof=1).mean()
            b2 = ((im.mean(axis=0) - ref) ** 2).mean() - var / 3
            out["bias" + key] = float(np.sign(b2) * np.sqrt(abs(b2)))
        return out

    ladder = sorted({max(1, p.spp * k // 64) for k in (4, 8, 16, 24, 32, 48, 64)})
    if args.cache_ladder:
        ladder = sorted({int(x) for x in args.cache_ladder.split(",") if 1 <= int(x) <= p.spp})
    plain = {s: over_seeds(None, s) for s in ladder}
    for s in ladder:
        log(f"[bench] plain spp {s}: {plain[s]['seconds'] * 1e3:.1f} ms rmse {plain[s]['rmse']:.4f} / "
            f"{plain[s]['rmse_clamped']:.4f} clamped")
    top = plain[ladder[-1]]
    targets = {"clamped": top["rmse_clamped"], "unclamped": top["rmse"]}
    plain_time = top["seconds"]

    def first_reaching(rows, key, target):
        # the first spp from which every larger spp also stays at or below the
        # target (single frames are noisy, and the cache keeps training)
        for k, r in enumerate(rows):
            if all(x["cached"][key] <= target for x in rows[k:]):
                return r["spp"], r["cached"]["seconds"]
        return None, None

    def cached_ladder(m):
        cfg = CACHE_MODES[m]
        nrc = kr.Nrc(dev, lo, hi, train_log2=4, batch=16384, target_clamp=emax, ema_decay=0.99, **cfg)
        warm = 0
        t_warm = time.perf_counter()
        # warm-up frames of 4 spp, each a different sample range of one
        # render: with counter-based random numbers, re-rendering one range
        # would train on the same paths every frame
        rw = kr.Render(gs, cam, scene.params(seed=7, spp=4 * args.cache_warmup))
        rw.attach_nrc(nrc)
        for warm in range(1, args.cache_warmup + 1):
            rw.run(4 * (warm - 1), 4)
        rw.sync()
        t_warm = time.perf_counter() - t_warm
        rw.free()
        log(f"[bench] cache ({m}) warm-up: {warm} frames of 4 spp in {t_warm:.2f}s, loss {nrc.stats()[1]:.3f}")
        rows = []
        for s in ladder:
            res = {"spp": s, "cached": over_seeds(nrc, s)}
            log(f"[bench] {m} spp {s}: cached {res['cached']['seconds'] * 1e3:.1f} ms rmse "
                f"{res['cached']['rmse']:.4f} / {res['cached']['rmse_clamped']:.4f} clamped")
            rows.append(res)
        out = {"mode": m, "description": CACHE_MODE_TEXT[m],
               "settings": {**cfg, "train_fraction": 1 / 16, "batch": 16384, "ema_decay": 0.99},
               "samples_per_frame": cfg["steps_per_frame"] * 16384,
               "warmup_frames": warm, "warmup_seconds": t_warm}
        for name, key in (("clamped", "rmse_clamped"), ("unclamped", "rmse")):
            spp, sec = first_reaching(rows, key, targets[name])
            out[name] = {"target_rmse": targets[name], "cached_spp": spp, "cached_seconds": sec,
                         "speedup": (plain_time / sec) if sec else None}
        out["bias_at_full_spp"] = {"unclamped": rows[-1]["cached"]["bias"],
                                   "clamped": rows[-1]["cached"]["bias_clamped"]}
        # cached + batched: the 3 seeds of the frame that reached the clamped
        # target as one batch (one graph launch per run, the cache attached
        # to all three jobs and trained once per run), after one untimed run
        # that captures the graphs
        sb = out["clamped"]["cached_spp"]
        if sb:
            jobs = [(0, cam, scene.params(seed=seed, spp=sb)) for seed in (11, 12, 13)]
            bt = kr.Batch(dev, [scene], jobs)
            for j in range(len(jobs)):
                bt.attach_nrc(j, nrc)
            bt.run()
            bt.sync()
            bt.reset()
            t0 = time.perf_counter()
            bt.run()
            bt.sync()
            dt = (time.perf_counter() - t0) / len(jobs)
            sq = []
            for j in range(len(jobs)):
                rgb, val = bt.read(j)
                img = np.clip(rgb[fpix] / np.maximum(val[fpix], 1)[:, None], 0.0, 1.0)
                sq.append(((img - np.clip(fref, 0.0, 1.0)) ** 2).mean())
            bt.free()
            rm = float(np.sqrt(np.mean(sq)))
            out["cached_batched"] = {"spp": sb, "jobs": len(jobs), "seconds_per_frame": dt, "rmse_clamped": rm,
                                     "reaches_target": rm <= targets["clamped"],
                                     "speedup": plain_time / dt if rm <= targets["clamped"] else None}
            log(f"[bench] {m} cached + batched, {len(jobs)} frames of {sb} spp in one launch: "
                f"{dt * 1e3:.1f} ms per frame, rmse {rm:.4f} clamped")
        if not args.no_fit:
            out["rmse_fit"] = rmse_fit(kr, scene, gs, args.workload, p.spp, nrc)
            log(f"[bench] {m} bias fit: " + fit_text(out["rmse_fit"]))
        nrc.free()
        out["by_spp"] = [{"spp": r["spp"], "seconds": r["cached"]["seconds"], "rmse": r["cached"]["rmse"],
                          "rmse_clamped": r["cached"]["rmse_clamped"],
                          "cache_queries": r["cached"].get("cache_queries")} for r in rows]
        return out

    out = {
        "plain_spp": ladder[-1], "plain_seconds": plain_time, "targets": targets, "target_clamp": emax,
        "plain_bias_at_full_spp": {"unclamped": top["bias"], "clamped": top["bias_clamped"]},
        "plain_by_spp": [{"spp": s, "seconds": plain[s]["seconds"], "rmse": plain[s]["rmse"],
                          "rmse_clamped": plain[s]["rmse_clamped"]} for s in ladder],
        "modes": {m: cached_ladder(m) for m in modes},
    }
    if "spec" in out["modes"]:
        sp = out["modes"]["spec"]
        out["spec_speedup_clamped"] = sp["clamped"]["speedup"]
        out["spec_speedup_unclamped"] = sp["unclamped"]["speedup"]
    return out


def bench_batch(args, kr, sg, barrier):
    """Batched uneven workload: 16 seeded scenes of 10k-2M triangles, mixed
    resolutions and 16-256 spp, rendered as one wavefront from one captured
    HIP graph after the autotuner picked BVH width, chunk size and blocks
    per CU. One step = every sample of every job."""
    import numpy as np
    log("[bench] building batch workload")
    scenes, specs = sg.batch_workload()
    jobs = [(sp["scene"], scenes[sp["scene"]].camera_struct(),
             kr.Params(sp["width"], sp["height"], sp["spp"], sp["max_bounce"], sp["seed"], 0, 0)) for sp in specs]
    total = sum(p.width * p.height * p.spp for _, _, p in jobs)
    ntris = sum(int(np.asarray(s.indices).shape[0]) for s in scenes)
    dev = kr.Device(0)
    b = kr.Batch(dev, scenes, jobs)
    t0 = time.perf_counter()
    tun = b.autotune()
    tune_s = time.perf_counter() - t0
    log(f"[bench] autotune {tune_s:.1f}s: BVH{tun.bvh_width}, {tun.chunk_paths} paths/chunk, "
        f"{tun.blocks_per_cu} blocks/CU ({tun.paths_per_second / 1e6:.0f} M paths/s in trials)")
    for i in range(args.warmup):
        b.reset()
        b.run()
        b.sync()
        log(f"[bench] warmup {i + 1}/{args.warmup} done ({b.info().seconds:.3f}s device)")
    info = b.info()
    log(f"[bench] graph: {info.graph_nodes} kernel nodes over {info.chunks} chunks, captured in "
        f"{info.capture_seconds:.2f}s")
    barrier()
    elapsed = 0.0
    for i in range(args.steps):
        b.reset()
        barrier()
        t1 = time.perf_counter()
        b.run()
        b.sync()
        barrier()
        elapsed += time.perf_counter() - t1
        log(f"[bench] step {i + 1}/{args.steps}: {b.info().seconds:.3f}s device")
    value = total * args.steps / elapsed

    # the same jobs one by one (plain renders at their default settings, BVH4),
    # with per-kernel events and a 1-spp counters pass for the traversal roofline
    log("[bench] jobs one by one ...")
    seq = 0.0
    trace_ms = trace_busy = trace_launches = closest = 0
    nodes = tris = cl_cnt = 0
    node_bytes = 0
    ups = [kr.GpuScene(dev, s.to_desc(bvh_width=4)) for s in scenes]
    for k, cam, p in jobs:
        r = kr.Render(ups[k], cam, p)
        r.set_profiling(kr.KR_PROFILE_COUNTERS)
        r.run(0, 1)
        r.sync()
        st = r.stats()
        nodes += st.node_visits
        tris += st.tri_tests
        cl_cnt += st.closest_rays
        node_bytes = ups[k].info().node_bytes
        r.reset()
        r.set_profiling(kr.KR_PROFILE_EVENTS)
        barrier()
        t1 = time.perf_counter()
        r.run(0, p.spp)
        r.sync()
        barrier()
        seq += time.perf_counter() - t1
        st = r.stats()
        trace_ms += st.trace_ms
        trace_busy += st.trace_busy_ms
        trace_launches += st.trace_launches
        closest += st.closest_rays
        r.free()
    # and one by one as a caller without a batch would: a render created,
    # run and freed per job (the scenes stay uploaded)
    barrier()
    t1 = time.perf_counter()
    for k, cam, p in jobs:
        r = kr.Render(ups[k], cam, p)
        r.run(0, p.spp)
        r.sync()
        r.free()
    barrier()
    seq_setup = time.perf_counter() - t1
    for u in ups:
        u.free()
    bytes_per_query = QUEUE_BYTES_PER_QUERY + (nodes * node_bytes + tris * TRI_BYTES) / max(1, cl_cnt)
    achieved = closest * bytes_per_query / (trace_busy * 1e-3) / 1e9 if trace_busy > 0 else 0.0
    log(f"[bench] one by one: {seq:.3f}s ({total / seq / 1e6:.0f} M paths/s), {seq_setup:.3f}s with a render "
        f"created per job; batched {elapsed / args.steps:.3f}s")

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(scenes[specs[0]["scene"]], args.cpu_seconds)
        cpu["sample"] = "job 0 of the batch: " + cpu["sample"]
        log(f"[bench] cpu baseline {cpu['value']:.0f} paths/s on {cpu['cores']} threads")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched uneven workload: 16 seeded scenes (10k-2M triangles, 6.25 M in total), "
                        "resolutions 256^2..1920x1080, 16-256 spp, maxBounce 8, one graph launch",
            "jobs": len(jobs), "triangles": ntris, "paths_per_step": total,
            "tuning": tun.as_dict(), "autotune_seconds": tune_s,
            "graph_nodes": info.graph_nodes, "chunks": info.chunks, "capture_seconds": info.capture_seconds,
            "one_by_one_seconds": seq, "speedup_vs_one_by_one": seq / (elapsed / args.steps),
            "one_by_one_with_render_setup_seconds": seq_setup,
            "speedup_vs_one_by_one_with_setup": seq_setup / (elapsed / args.steps),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "closest-hit wide-BVH traversal of one bounce's ray queue (k_trace, or k_trace_tail and its "
                      "follow-up launches), timed on the jobs rendered one by one",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_query": bytes_per_query,
            "queries_per_launch": closest / max(1, trace_launches),
            "avg_launch_ms": trace_busy / max(1, trace_launches),
            "launch_time": "union of the traversal intervals per queue (four streams per render)",
            "avg_overlapped_launch_ms": trace_ms / max(1, trace_launches),
        },
        "cpu_baseline": cpu,
    }
    b.free()
    dev.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="large", choices=["mesh", "spheres", "large", "room", "batch"])
    ap.add_argument("--bvh-width", type=int, default=0)
    ap.add_argument("--chunk-paths", type=int, default=0)
    ap.add_argument("--leaf-size", type=int, default=0,
                    help="most triangles per BVH leaf (0: the library's default)")
    ap.add_argument("--full-nodes", action="store_true",
                    help="full FP32 BVH child boxes instead of the default 8-bit quantised ones")
    ap.add_argument("--no-cache", action="store_true",
                    help="skip the radiance-cache leg (plain vs cached wall-clock to the plain frame's RMSE "
                         "against the FP64 oracle), which runs by default on workloads with an oracle fixture")
    ap.add_argument("--cache-warmup", type=int, default=64,
                    help="4-spp training frames (distinct sample ranges) before the cached frames are timed")
    ap.add_argument("--cache-ladder", default="",
                    help="comma-separated spp of the timed frames (default 4, 8, 16, 24, 32, 48, 64 scaled to the "
                         "workload's spp); the targets are the plain frame's RMSE at the largest")
    ap.add_argument("--cache-mode", default="both", choices=["both", "spec", "first"],
                    help="where cached paths end (CACHE_MODES): spec = the workload's second diffuse or rough "
                         "vertex, first = the first; both measures spec, then first")
    ap.add_argument("--no-material-sort", action="store_true",
                    help="shade every hit from one mixed-material queue")
    ap.add_argument("--one-stream", action="store_true",
                    help="run every chunk on one stream (KR_RENDER_ONE_STREAM): kernels do not overlap, so "
                         "their launch durations are their own")
    ap.add_argument("--persistent-traversal", action="store_true",
                    help="traversal kernels refill finished lanes from the queue")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fit", action="store_true",
                    help="skip the noise / bias fit RMSE^2(m) = b^2 + sigma^2 / m over m = 1..16 frames")
    ap.add_argument("--no-torch", action="store_true", help="skip torch.cuda.synchronize (profiling runs)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: there is no multi-GPU path (paths are independent and would "
              "shard by sample range with one FP64 sum-reduce, which is not built). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from koishi_amd import kr, scenegen as sg

    torch = None
    if not args.no_torch:
        import torch as _t
        torch = _t

    def barrier():
        if torch is not None:
            torch.cuda.synchronize()

    if args.workload == "batch":
        return bench_batch(args, kr, sg, barrier)

    log(f"[bench] building workload {args.workload}")
    scene = sg.WORKLOADS[args.workload]()
    dev = kr.Device(0)
    gs = kr.GpuScene(dev, scene.to_desc(bvh_width=args.bvh_width, leaf_size=args.leaf_size,
                                        flags=kr.KR_SCENE_FULL_NODES if args.full_nodes else 0))
    info = gs.info()
    log(f"[bench] scene: {info.num_triangles} tris, {info.num_spheres} spheres, {info.num_lights} lights, "
        f"BVH{info.bvh_width} {info.bvh_nodes} nodes depth {info.bvh_depth}, build {info.build_seconds:.2f}s")
    p = scene.params(seed=1, chunk_paths=args.chunk_paths,
                     flags=(kr.KR_RENDER_PERSISTENT_TRAVERSAL if args.persistent_traversal else 0)
                     | (kr.KR_RENDER_NO_MATERIAL_SORT if args.no_material_sort else 0)
                     | (kr.KR_RENDER_ONE_STREAM if args.one_stream else 0))
    rend = kr.Render(gs, scene.camera_struct(), p)
    paths_per_step = p.width * p.height * p.spp

    # traversal work per query, counted by the same kernel on a 4-spp sample
    rend.set_profiling(kr.KR_PROFILE_COUNTERS)
    rend.run(0, min(4, p.spp))
    rend.sync()
    st = rend.stats()
    nodes_per_ray = st.node_visits / max(1, st.closest_rays)
    tris_per_ray = st.tri_tests / max(1, st.closest_rays)
    bytes_per_query = QUEUE_BYTES_PER_QUERY + nodes_per_ray * info.node_bytes + tris_per_ray * TRI_BYTES
    log(f"[bench] per closest-hit query: {nodes_per_ray:.1f} nodes, {tris_per_ray:.1f} tri tests, "
        f"{bytes_per_query:.0f} algorithmic bytes")

    # the timed frames record HIP events around the traversal launches only
    # (events around every launch cost 2.5 % of the mesh frame); the other
    # kinds' times come from more frames after the timed region. The warm-up
    # frames record the same events, so that the timed ones find the events
    # already created.
    rend.set_profiling(kr.KR_PROFILE_TRACE_EVENTS)
    for i in range(args.warmup):
        rend.run(0, p.spp)
        rend.sync()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    rend.reset()
    barrier()
    rend.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        rend.run(0, p.spp)
    rend.sync()
    barrier()
    elapsed = time.perf_counter() - t0
    st = rend.stats()
    log(f"[bench] {args.steps} steps in {elapsed:.3f}s")
    rgb, valid = rend.read()
    value = paths_per_step * args.steps / elapsed
    # the closest-hit traversal of one bounce's queue: one k_trace launch for
    # the camera rays, a k_trace_tail launch and its follow-up launches for
    # the later bounces. HIP events around it, on the stream it runs on, over
    # the timed steps (a frame's chunks run on four streams, so it shares the
    # GPU with the other streams' kernels for part of that time)
    avg_launch_ms = st.trace_ms / max(1, st.trace_launches)
    queries_per_launch = st.closest_rays / max(1, st.trace_launches)
    achieved = queries_per_launch * bytes_per_query / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else 0.0

    rmse = rmse_vs_oracle(args.workload, rgb, valid, p.width)
    fit = None
    if not args.no_fit:
        rend.set_profiling(0)
        fit = rmse_fit(kr, scene, gs, args.workload, p.spp)
        if fit:
            log("[bench] plain bias fit: " + fit_text(fit))
    # every launch timed, over as many frames as the timed region, after one
    # frame that creates the events
    rend.set_profiling(kr.KR_PROFILE_EVENTS)
    rend.reset()
    rend.run(0, p.spp)
    rend.sync()
    rend.reset()
    for _ in range(args.steps):
        rend.run(0, p.spp)
    rend.sync()
    kst = rend.stats()
    rend.set_profiling(0)

    # the same traversal with the chunks on one stream, where a queue's
    # launches have the GPU to themselves: one untimed frame, then up to 3
    # frames with HIP events around the traversal launches (after the timed
    # region)
    one_stream = None
    if not args.one_stream:
        p1 = scene.params(seed=1, chunk_paths=args.chunk_paths, flags=p.flags | kr.KR_RENDER_ONE_STREAM)
        r1 = kr.Render(gs, scene.camera_struct(), p1)
        r1.set_profiling(kr.KR_PROFILE_TRACE_EVENTS)
        r1.run(0, p.spp)
        r1.sync()
        r1.reset()
        n1 = min(3, args.steps)
        for _ in range(n1):
            r1.run(0, p.spp)
        r1.sync()
        s1 = r1.stats()
        ms1 = s1.trace_ms / max(1, s1.trace_launches)
        a1 = (s1.closest_rays / max(1, s1.trace_launches)) * bytes_per_query / (ms1 * 1e-3) / 1e9 if ms1 > 0 else 0.0
        one_stream = {"avg_launch_ms": ms1, "achieved": a1, "frac": a1 / HBM_PEAK_GBS, "frames": n1,
                      "launches": int(s1.trace_launches)}
        r1.free()

    cached = None
    if not args.no_cache and os.path.exists(os.path.join(ROOT, "tests", "golden", f"oracle_{args.workload}.npz")):
        rend.set_profiling(0)
        cached = bench_cache(args, kr, scene, gs, p, dev)

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(scene, args.cpu_seconds)
        log(f"[bench] cpu baseline {cpu['value']:.0f} paths/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": {"mesh": "triangle-mesh BVH scene: displaced icosphere 100,820 tris + ground + quad "
                                 "light, 1024x512, 64 spp, maxBounce 8, plain wavefront path",
                         "spheres": "sphere scene: 8 spheres + ground sphere + quad light, 256x256, 16 spp, "
                                    "maxBounce 5",
                         "large": "large scene: 40 displaced icosphere instances 1,036,800 tris, 4 quad "
                                  "lights, 1920x1080, 64 spp, maxBounce 8",
                         "room": "closed room: 6 walls, displaced icosphere 8,000 tris, ceiling light, "
                                 "1024x768, 64 spp, maxBounce 8"}[args.workload],
            "scene_data": "seeded synthetic scene generated in-process; the plain steps run without a cache",
            "width": p.width, "height": p.height, "spp": p.spp, "max_bounce": p.max_bounce,
            "triangles": info.num_triangles, "bvh_width": info.bvh_width, "bvh_node_bytes": info.node_bytes,
            "bvh_nodes": info.bvh_nodes, "bvh_leaves": info.bvh_leaves,
            "paths_per_step": paths_per_step,
            "mrays_per_s": (st.closest_rays + st.shadow_rays) / elapsed / 1e6,
            # the timed steps' frames accumulated (steps x spp samples per pixel)
            "rmse_vs_fp64_oracle": rmse,
            "rmse_samples_per_pixel": p.spp * args.steps,
            # noise and bias of the plain path against the oracle fixtures
            "rmse_fit": fit,
            "wall_clock_per_frame_s": elapsed / args.steps,
            # frames after the timed region, every launch timed: per kernel kind
            # the busy time (union of its launches' intervals over the four
            # streams) and the summed launch durations per frame
            "kernel_busy_ms_per_step": {"trace": kst.trace_busy_ms / args.steps,
                                        "shade": kst.shade_busy_ms / args.steps,
                                        "shadow": kst.shadow_busy_ms / args.steps},
            "kernel_ms_per_step_summed": {"trace": kst.trace_ms / args.steps, "shade": kst.shade_ms / args.steps,
                                          "shadow": kst.shadow_ms / args.steps, "other": kst.other_ms / args.steps},
            "kernel_launches_per_step": {"trace": kst.trace_launches / args.steps,
                                         "shade": kst.shade_launches / args.steps,
                                         "shadow": kst.shadow_launches / args.steps},
            "cached": cached,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "closest-hit wide-BVH traversal of one bounce's ray queue (k_trace for camera rays; "
                      "k_trace_tail and its follow-up launches for later bounces)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            # HBM bytes per launch come from rocprofv3 FETCH_SIZE / WRITE_SIZE
            # runs of this command (committed summaries), not from this run
            "traffic": None,
            "bytes_per_query": bytes_per_query,
            "queries_per_launch": queries_per_launch,
            "avg_launch_ms": avg_launch_ms,
            "launches": int(st.trace_launches),
            "launch_time": "mean duration of the traversal of one queue (its first launch to its last follow-up), "
                           "HIP events recorded around it on the stream it runs on over the timed steps (four "
                           "streams per frame share the GPU)",
            # the same figures with the frame's chunks on one stream (frames
            # after the timed region): the traversal's own speed
            "one_stream": one_stream,
            # the HBM roofline is the reference line, not the limiter: the
            # first k_trace_tail launch of a bounce moves 2.08 GB of HBM + MALL
            # traffic against 8.94 GB of algorithmic bytes, issues VALU
            # instructions in 87 % of the SIMDs' cycles with 41 % of their
            # lanes active, and keeps the texture data unit 91 % busy
            # (rocprofv3 FETCH_SIZE / WRITE_SIZE, SQ and TA / TD counters of this frame)
            "limiter": "VALU issue and the vector memory data path (the first traversal launch of a bounce issues "
                       "VALU in 87 % of SIMD cycles with 41 % of lanes active, the texture data unit 91 % busy); "
                       "~68 % of the algorithmic bytes hit in L2",
        },
        "cpu_baseline": cpu,
    }
    rend.free()
    gs.free()
    dev.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
