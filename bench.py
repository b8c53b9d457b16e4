"""Benchmark of the hot path on one MI355X.

Workload: incoherent extend rays (8,388,608 rays, origins uniform in the scene box, directions uniform
on the sphere, seed 0) against the 1024 x 1024-cell procedural heightfield (2,097,152 triangles).
One step = one closest-hit launch over all rays, rays and BVH resident in HBM. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "incoherent Mrays/s (2M tris) + equal-time relMSE vs 4096-spp ref, plain vs cache"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name as lscpu prints it (falls back to /proc/cpuinfo)."""
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.strip().startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# Host cores one GPU's job may use on the benchmark machines: the CPU share of a one-GPU slot.
HOST_CORE_SHARE = 16


def cpu_baseline(scene, rays, seconds_target=2.0, runs=5):
    """The oracle's restatement of the reference CPU path (two-level BVH without tmax culling,
    Moller-Trumbore, one task per chunk over the host threads), timed on a bounded sample: median of
    `runs` timed runs after a warm-up, each about `seconds_target` long, on every core of the
    process's CPU share (the affinity mask, capped at the one-GPU slot's share of 16 cores), and on
    one thread. BVH build and upload are excluded."""
    from oracle import refcpu

    affinity = len(os.sched_getaffinity(0))
    threads = min(affinity, HOST_CORE_SHARE)
    o = refcpu.OracleScene(scene)

    def rate(nthreads, mode, budget):
        sample = rays[: 1 << 12]
        t0 = time.perf_counter()
        o.trace(sample, mode, nthreads)  # warm-up and size probe
        dt = max(time.perf_counter() - t0, 1e-4)
        n = int(min(rays.shape[0], max(1 << 12, (1 << 12) * budget / dt)))
        sample = rays[:n]
        times = []
        for _ in range(runs):
            t0 = time.perf_counter()
            o.trace(sample, mode, nthreads)
            times.append(time.perf_counter() - t0)
        return n / float(np.median(times)) / 1e6, n

    mt_rate, n = rate(threads, refcpu.TRACE_REF_MT, seconds_target)
    mt1_rate, n1 = rate(1, refcpu.TRACE_REF_MT, seconds_target / 2)
    wt_rate, nw = rate(threads, refcpu.TRACE_WT, seconds_target / 2)
    model = cpu_model()
    return {"value": round(mt_rate, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"first {n} of the 8,388,608 incoherent rays, reference Moller-Trumbore two-level BVH "
                      f"without tmax culling, median of {runs} runs on {threads} threads ({model}; affinity "
                      f"mask {affinity} CPUs, capped at the one-GPU slot's share of {HOST_CORE_SHARE}); "
                      f"1 thread: {mt1_rate:.3f} Mrays/s on {n1} rays; watertight FP32 variant "
                      f"{wt_rate:.3f} Mrays/s on {nw} rays",
            "one_thread_value": round(mt1_rate, 3),
            "watertight_value": round(wt_rate, 3),
            "affinity_cpus": affinity,
            "cpu_model": model}


def oracle_check(scene, rays, hits, stride=16, fp64_stride=64):
    """Checks the timed call's hits against the oracle: every `stride`-th ray bit for bit against the
    watertight twin, every `fp64_stride`-th against FP64 (ids on unflagged rays, t within 1e-6
    relative), and counts the sampled rays on which the reference's Moller-Trumbore test with its
    |det| < 1e-6 cutoff (MeshRenderProxy.cpp:143-150) reports a different closest hit. Raises if the
    device differs from the twin or from FP64."""
    from oracle import refcpu

    miss = 0xFFFFFFFF
    o = refcpu.OracleScene(scene)
    threads = min(len(os.sched_getaffinity(0)), HOST_CORE_SHARE)
    sub = np.arange(0, rays.shape[0], stride)
    wt = o.trace(rays[sub], refcpu.TRACE_WT, threads)
    same = (hits["prim"][sub] == wt["prim"]) & (hits["t"][sub].view(np.uint32) == wt["t"].view(np.uint32))
    if not same.all():
        raise RuntimeError(f"{int((~same).sum())} of {sub.size} sampled hits differ from the oracle's twin")
    sub64 = np.arange(0, rays.shape[0], fp64_stride)
    h64, flags, t64 = o.trace(rays[sub64], refcpu.TRACE_FP64, threads)
    clean = flags == 0
    if not np.array_equal(hits["prim"][sub64][clean], h64["prim"][clean]):
        raise RuntimeError("hit ids differ from FP64 on unflagged rays")
    m = clean & (h64["prim"] != miss)
    rel = float((np.abs(hits["t"][sub64][m].astype(np.float64) - t64[m]) / t64[m]).max()) if m.any() else 0.0
    if rel > 1e-6:
        raise RuntimeError(f"hit t differs from FP64 by {rel:.3g} relative")
    mt = o.trace(rays[sub], refcpu.TRACE_REF_MT, threads)
    wt_hit, mt_hit = wt["prim"] != miss, mt["prim"] != miss
    differ = mt["prim"] != wt["prim"]
    return {"rays_checked_bit_exact": int(sub.size), "stride": stride, "bit_exact": True,
            "fp64_rays": int(sub64.size), "fp64_unflagged": int(clean.sum()), "fp64_ids_equal": True,
            "fp64_max_rel_t": rel,
            "reference_mt_disagreement": {
                "rays": int(sub.size), "differ": int(differ.sum()), "fraction": round(float(differ.mean()), 4),
                "twin_hits_mt_misses": int((wt_hit & ~mt_hit).sum()),
                "mt_hits_twin_misses": int((mt_hit & ~wt_hit).sum()),
                "both_hit_other_triangle": int((wt_hit & mt_hit & differ).sum()),
                "twin_hits": int(wt_hit.sum())}}


def load_traffic(config):
    """HBM-side bytes per launch and L2 hit rate from the committed rocprofv3 counter summary, for the
    kernel configuration this run uses (the summary holds one `config:` block per profiled
    configuration)."""
    path = os.path.join(ROOT, "profiles", "trace_incoherent_2m_pmc.md")
    if not os.path.exists(path):
        return None, None, None
    blocks, cur = {}, None
    for line in open(path):
        if ":" not in line:
            continue
        k, v = (x.strip() for x in line.split(":", 1))
        if k == "config":
            cur = blocks.setdefault(v, {})
        elif k in ("hbm_bytes_per_launch", "l2_hit_rate") and cur is not None:
            cur[k] = v
    fields = blocks.get(config)
    if not fields or "hbm_bytes_per_launch" not in fields:
        return None, None, f"profiles/trace_incoherent_2m_pmc.md has no counters for config {config} " \
                           f"(profiled: {', '.join(sorted(blocks))})"
    l2 = float(fields["l2_hit_rate"]) if "l2_hit_rate" in fields else None
    return int(float(fields["hbm_bytes_per_launch"])), l2, "profiles/trace_incoherent_2m_pmc.md"


def spread(ms):
    """Median, 90th percentile and maximum of per-frame times, and how many frames took over twice
    the median (a stall of the GPU or the host shows here)."""
    if not ms:
        return None
    a = np.asarray(ms, np.float64)
    med = float(np.median(a))
    return {"median": round(med, 3), "p90": round(float(np.percentile(a, 90)), 3), "max": round(float(a.max()), 3),
            "frames_over_2x_median": int((a > 2 * med).sum()), "frames": int(a.size)}


def relmse(img, ref, eps=1e-2):
    """Relative MSE over pixels and channels: mean((x - r)^2 / (r^2 + eps))."""
    return float(((img - ref) ** 2 / (ref ** 2 + eps)).mean())


def relmse_experiment(spk, scenes, stream, width, height, spp, ref_spp, max_bounce=8, warm_frac=0.1,
                      budgets=(1.0, 4.0)):
    """Equal-time comparison of estimators on the many-instance scene (about 5.5M triangles, 3 area
    lights) against a ref_spp reference.

    plain        the reference's estimator (BSDF sampling only): spp frames (seeds 0..spp-1)
                 accumulated by spk_accumulate; its wall time T is the unit of the time budgets
    plain_nee    plain + next-event estimation
    cache        learned radiance cache (terminate at bounce 2, online training every frame)
    cache_nee    cache + next-event estimation
    cache1       cache from bounce 1 (a path ends at its first non-metal vertex after the camera hit)
    cache1_nee   cache from bounce 1 + next-event estimation
    For each budget b*T every estimator renders frames until b*T has passed (plain at b = 1 renders
    exactly spp frames); the cache estimators start untrained (reset after their graph-capture frame)
    and accumulate only the frames after warm_frac of the budget. Each estimator runs on the faster of
    the launch schedule spk_render_autotune picks for it and the default schedule (median over 8
    frames each, before any timed region). The reference is ref_spp frames of
    the plain FP32 tracer (the reference's estimator), seeds from 1,000,000 on, summed in float64 on the
    device; its own noise adds about relMSE(plain at spp) * spp / ref_spp to every figure (reported as
    reference_noise_relmse). All frames use the 64-bit seed composition: at
    1080p the reference's uint32 j*W + i + f*W*H wraps from frame 2072 on. No per-sample clamp
    anywhere: clamped estimators converge to different (biased) images, so an equal-time comparison
    against one reference would measure the clamp.
    """
    import torch

    log(f"relMSE experiment: generating the many-instance scene, {width}x{height}")
    sc, cam = scenes.many_instance_scene(seed=0, aspect=width / height)
    ctx = spk.Context(0)
    ctx.set_stream(stream.cuda_stream)
    t_build = time.perf_counter()
    ctx.upload_scene(sc)
    t_build = time.perf_counter() - t_build
    st = ctx.stats()
    ctx.set_camera(cam)
    npx = width * height
    rgba = torch.zeros(npx * 4, dtype=torch.float32, device="cuda")
    acc = torch.zeros(npx * 4, dtype=torch.float32, device="cuda")

    def frame(seed, flags=0):
        f = spk.FrameDesc()
        f.width, f.height, f.frame_seed, f.max_bounce, f.debug_mode = width, height, seed, max_bounce, 0
        f.sky_enabled, f.flags, f.output_limit = 0, flags | spk.FRAME_SEED64, 1e30
        return f

    variants = [("plain", 0, None), ("plain_nee", spk.FRAME_NEE, None), ("cache", spk.FRAME_CACHE, 2),
                ("cache_nee", spk.FRAME_CACHE | spk.FRAME_NEE, 2), ("cache1", spk.FRAME_CACHE, 1),
                ("cache1_nee", spk.FRAME_CACHE | spk.FRAME_NEE, 1)]
    def steady_ms(flags, sched, frames=8):
        """Median frame time of a schedule over several frame seeds (after a capture frame)."""
        ctx.set_schedule(*sched)
        ctx.render_frame(frame(999_980, flags), rgba)
        return float(np.median([ctx.render_frame(frame(999_981 + k, flags), rgba).milliseconds
                                for k in range(frames)]))

    # each estimator on the faster (median over 8 frames) of its autotuned schedule and the default
    # schedule: a single tuning pass can pick a schedule several percent off, which would shift the
    # equal-time budget between estimators
    default_schedule = (8, 0, 0, 1)
    schedules, schedule_ms = {}, {}
    for name, flags, bounce in variants:
        if bounce is not None:
            ctx.cache_create(spk.cache_config(terminate_bounce=bounce))
        tuned, _ = ctx.autotune_render(frame(999_990, flags), rgba)
        timed = sorted((steady_ms(flags, sc), sc) for sc in {tuned, default_schedule})
        schedule_ms[name], schedules[name] = round(timed[0][0], 3), timed[0][1]
        if bounce is not None:
            ctx.cache_destroy()
    log(f"relMSE experiment: schedules {schedules}")

    log(f"relMSE experiment: reference at {ref_spp} spp")
    ctx.set_schedule(*schedules["plain"])
    ctx.render_frame(frame(999_997), rgba)  # graph capture
    ref = torch.zeros(npx * 4, dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    rays_ref = 0
    ref_gpu_ms, ref_wall_ms = [], []
    for k in range(ref_spp):
        tf = time.perf_counter()
        fs = ctx.render_frame(frame(1_000_000 + k), rgba)
        rays_ref += fs.rays_traced
        ref += rgba.double()
        ref_gpu_ms.append(fs.milliseconds)
        ref_wall_ms.append((time.perf_counter() - tf) * 1e3)
        if (k + 1) % 512 == 0:
            log(f"  reference {k + 1}/{ref_spp} spp, {time.perf_counter() - t0:.1f} s "
                f"(frame kernels: median {np.median(ref_gpu_ms[-512:]):.2f} ms, max {max(ref_gpu_ms[-512:]):.2f} ms; "
                f"wall per frame: median {np.median(ref_wall_ms[-512:]):.2f} ms, max {max(ref_wall_ms[-512:]):.2f} ms)")
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t0
    ref = (ref / ref_spp).view(height, width, 4)[..., :3]

    log(f"relMSE experiment: plain at {spp} spp")
    ctx.set_schedule(*schedules["plain"])
    ctx.render_frame(frame(999_999), rgba)  # graph capture for the plain flags
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(spp):
        ctx.render_frame(frame(k), rgba)
        ctx.accumulate(rgba, acc, npx, k, 1)
    torch.cuda.synchronize()
    T = time.perf_counter() - t0
    img = acc.view(height, width, 4)[..., :3].double()
    first = {"plain": {"frames": spp, "accumulated_spp": spp, "seconds": round(T, 3),
                       "relmse": relmse(img, ref), "mean": float(img.mean())}}

    by_budget = []
    for b in budgets:
        budget = b * T
        results = dict(first) if b == 1.0 else {}
        for name, flags, bounce in variants:
            if name in results:
                continue
            log(f"relMSE experiment: {name} for {budget:.2f} s")
            ctx.set_schedule(*schedules[name])
            if bounce is not None:
                ctx.cache_create(spk.cache_config(terminate_bounce=bounce))
            ctx.render_frame(frame(999_998, flags), rgba)  # graph capture
            if bounce is not None:
                ctx.cache_reset()
            torch.cuda.synchronize()
            acc.zero_()
            frames = used = 0
            warm = warm_frac * budget if bounce is not None else 0.0
            t0 = time.perf_counter()
            wall_ms = []
            while True:
                el = time.perf_counter() - t0
                if el >= budget:
                    break
                tf = time.perf_counter()
                ctx.render_frame(frame(frames, flags), rgba)
                wall_ms.append((time.perf_counter() - tf) * 1e3)
                frames += 1
                if el >= warm:
                    ctx.accumulate(rgba, acc, npx, used, 1)
                    used += 1
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            img = acc.view(height, width, 4)[..., :3].double()
            r = {"frames": frames, "accumulated_spp": used, "seconds": round(dt, 3),
                 "relmse": relmse(img, ref) if used else None, "mean": float(img.mean()),
                 "frame_wall_ms": spread(wall_ms)}
            if bounce is not None:
                r["terminate_bounce"] = bounce
                r["last_loss"] = ctx.cache_info()["last_loss"]
                ctx.cache_destroy()
            results[name] = r
        e_plain, e_nee = results["plain"]["relmse"], results["plain_nee"]["relmse"]
        # the reference's own noise adds about the same relMSE to every estimator; taking it off
        # estimates each one's error against the converged image
        noise = first["plain"]["relmse"] * spp / ref_spp
        by_budget.append({
            "seconds": round(budget, 3), "estimators": results,
            "ratio_plain_over": {k: round(e_plain / v["relmse"], 3) for k, v in results.items()
                                 if k != "plain" and v["relmse"]},
            "ratio_plain_nee_over": {k: round(e_nee / v["relmse"], 3) for k, v in results.items()
                                     if k.startswith("cache") and v["relmse"]},
            "ratio_plain_over_less_reference_noise": {
                k: round((e_plain - noise) / (v["relmse"] - noise), 3) for k, v in results.items()
                if k != "plain" and v["relmse"] and v["relmse"] > noise},
            "ratio_plain_nee_over_less_reference_noise": {
                k: round((e_nee - noise) / (v["relmse"] - noise), 3) for k, v in results.items()
                if k.startswith("cache") and v["relmse"] and v["relmse"] > noise}})

    out = {
        "scene": f"many-instance generator seed 0: {st.num_triangles} triangles, 3 area lights, {width}x{height}, "
                 f"max {max_bounce} bounces, Russian roulette from bounce 3",
        "relmse_definition": "mean over pixels and channels of (I - R)^2 / (R^2 + 1e-2)",
        "reference_spp": ref_spp,
        "reference_estimator": "plain (BSDF sampling only, the reference's estimator), FP32",
        "reference_noise_relmse": round(first["plain"]["relmse"] * spp / ref_spp, 5),
        "reference_mean": float(ref.mean()),
        "per_sample_clamp": None,
        "reference_s": round(t_ref, 2),
        "reference_frame_kernels_ms": spread(ref_gpu_ms),
        "reference_frame_wall_ms": spread(ref_wall_ms),
        "equal_time_s": round(T, 3),
        "budgets": by_budget,
        "schedules": {k: dict(zip(("blocks_per_cu", "material_sort", "refill", "tail_bounce"), v))
                      for k, v in schedules.items()},
        "schedule_median_frame_ms": schedule_ms,
        "reference_pass_samples_per_s": round(npx * ref_spp / t_ref, 1),
        "reference_pass_mrays_per_s": round(rays_ref / t_ref / 1e6, 1),
        "bvh_build_s": round(t_build, 2),
    }
    ctx.close()
    return out


MLP_FLOP_FORWARD = 2 * (64 * 64 + 3 * 64 * 64 + 64 * 3)  # 33,152 per query (64 inputs: encoding + attributes)
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 MFMA


def cache_kernel_bench(spk, scenes, stream, n_infer=1920 * 1080, n_train=1 << 16, reps=20):
    """The learned radiance cache's kernels alone: inference on n_infer points (one 1080p frame of
    queries) and training steps (forward, backward, Adam) on n_train samples, uniform points in the
    unit cube with random surface attributes, timed with HIP events on the context's stream.
    Algorithmic FLOPs: 33,152 per query forward, 3x that per training sample (forward, input
    gradients, weight gradients)."""
    import torch

    ctx = spk.Context(0)
    ctx.set_stream(stream.cuda_stream)
    ctx.upload_scene(scenes.cornell_scene(0))  # the cache's grid is sized from a scene box
    ctx.cache_create(spk.cache_config())
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    pts = torch.rand(n_infer * 3, device="cuda", generator=g)
    out = torch.empty(n_infer * 3, device="cuda")
    tp = torch.rand(n_train * 3, device="cuda", generator=g)
    tt = torch.rand(n_train * 3, device="cuda", generator=g)

    def attrs(n):  # unit normals, albedo, roughness, 1: what the path tracer records per query
        nrm = torch.nn.functional.normalize(torch.randn(n, 3, device="cuda", generator=g), dim=1)
        return torch.cat([nrm, 0.2 + 0.6 * torch.rand(n, 4, device="cuda", generator=g)[:, :3],
                          torch.rand(n, 1, device="cuda", generator=g), torch.ones(n, 1, device="cuda")], 1).contiguous()

    pa, ta = attrs(n_infer), attrs(n_train)
    for _ in range(3):
        ctx.cache_infer(pts, out, n_infer, attrs_dev=pa)
        ctx.cache_train(tp, tt, n_train, attrs_dev=ta)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        ctx.cache_infer(pts, out, n_infer, attrs_dev=pa)
    b.record(stream)
    torch.cuda.synchronize()
    infer_ms = a.elapsed_time(b) / reps
    a.record(stream)
    for _ in range(reps):
        ctx.cache_train(tp, tt, n_train, attrs_dev=ta)
    b.record(stream)
    torch.cuda.synchronize()
    train_ms = a.elapsed_time(b) / reps
    ctx.close()
    infer_tf = MLP_FLOP_FORWARD * n_infer / (infer_ms / 1e3) / 1e12
    train_tf = 3 * MLP_FLOP_FORWARD * n_train / (train_ms / 1e3) / 1e12
    return {"infer_queries": n_infer, "infer_ms": round(infer_ms, 4),
            "infer_mqueries_per_s": round(n_infer / (infer_ms / 1e3) / 1e6, 1),
            "infer_tflops_algorithmic": round(infer_tf, 2), "infer_frac_bf16_peak": round(infer_tf / BF16_PEAK_TFLOPS, 4),
            "train_samples": n_train, "train_step_ms": round(train_ms, 4),
            "train_tflops_algorithmic": round(train_tf, 2),
            "train_frac_bf16_peak": round(train_tf / BF16_PEAK_TFLOPS, 4),
            "bound": "hash-grid gathers (16 levels x 8 corners x 4 B per query) and table-gradient atomics"}


RENDER_MODES = {"plain": (0, None), "plain_nee": ("NEE", None), "cache_nee": ("NEE", 2), "cache1_nee": ("NEE", 1),
                "cache": (0, 2), "cache1": (0, 1)}


def render_experiment(spk, scenes, stream, width=1920, height=1080, spp=64, max_bounce=8, autotune=True,
                      mode="plain", schedule=None):
    """The full path trace of the many-instance scene (about 5.5M triangles, 400 instances,
    3 area lights): spp frames of one path per pixel, max_bounce bounces, Russian roulette from
    bounce 3, material-sorted queues, the frame graph replayed; schedule from spk_render_autotune.
    mode "plain" is the reference's estimator; the others add next-event estimation and/or the
    learned cache (terminating at bounce 2, or 1 for cache1 and cache1_nee, trained every frame)."""
    import torch

    nee, cache_bounce = RENDER_MODES[mode]
    flags = (spk.FRAME_NEE if nee else 0) | (spk.FRAME_CACHE if cache_bounce is not None else 0)
    log(f"render experiment: many-instance scene {width}x{height}, {spp} spp, {mode}")
    sc, cam = scenes.many_instance_scene(seed=0, aspect=width / height)
    ctx = spk.Context(0)
    ctx.set_stream(stream.cuda_stream)
    t0 = time.perf_counter()
    st = ctx.upload_scene(sc)
    t_build = time.perf_counter() - t0
    ctx.set_camera(cam)
    npx = width * height
    rgba = torch.zeros(npx * 4, dtype=torch.float32, device="cuda")
    acc = torch.zeros(npx * 4, dtype=torch.float32, device="cuda")

    def frame(seed):
        f = spk.FrameDesc()
        f.width, f.height, f.frame_seed, f.max_bounce = width, height, seed, max_bounce
        f.debug_mode, f.sky_enabled, f.flags, f.output_limit = 0, 0, flags, 0.0
        return f

    if cache_bounce is not None:
        ctx.cache_create(spk.cache_config(terminate_bounce=cache_bounce))
    sched = ctx.schedule()
    if schedule is not None:
        ctx.set_schedule(*schedule)
        sched = tuple(schedule)
    elif autotune:
        sched, _ = ctx.autotune_render(frame(10 ** 6), rgba)
    ctx.render_frame(frame(10 ** 6 + 1), rgba)
    torch.cuda.synchronize()
    rays = 0
    t0 = time.perf_counter()
    for k in range(spp):
        rays += ctx.render_frame(frame(k), rgba).rays_traced
        ctx.accumulate(rgba, acc, npx, k, 1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ctx.close()
    return {"scene": f"many-instance generator seed 0: {st.num_triangles} triangles, {width}x{height}, "
                     f"max {max_bounce} bounces", "mode": mode, "spp": spp, "seconds": round(dt, 3),
            "ms_per_frame": round(dt / spp * 1e3, 3), "samples_per_s": round(npx * spp / dt, 1),
            "mrays_per_s": round(rays / dt / 1e6, 1), "rays_per_sample": round(rays / (npx * spp), 3),
            "schedule": {"blocks_per_cu": sched[0], "material_sort": bool(sched[1]), "refill": sched[2],
                         "tail_bounce": sched[3]},
            "bvh": {"width": st.bvh_width, "depth": st.depth, "nodes": st.num_nodes, "build_s": round(t_build, 2)}}


def two_level_experiment(spk, scenes, stream, width=1920, height=1080, spp=16, n_rays=1 << 22, max_bounce=8):
    """Instanced form of the many-instance scene (icospheres and tori as shared object-space
    prototypes, displaced boxes as their own meshes, each placed by a transform): uploaded flattened
    and two-level (TLAS over the instances, one BLAS per mesh); for each, the BVH build time, device
    bytes, closest-hit rate on n_rays incoherent rays in the scene box (autotuned variant) and the
    plain path-trace frame time (autotuned schedule); for the two-level scene, the host time of
    spk_scene_update_instances moving every instance."""
    import torch

    log("two-level experiment: instanced many-instance scene")
    sc, cam = scenes.many_instance_scene(seed=0, aspect=width / height, shared=True)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=n_rays, seed=0)
    d_rays = torch.from_numpy(rays.view(np.uint8).reshape(-1)).cuda()
    d_hits = torch.empty(n_rays * 8, dtype=torch.uint8, device="cuda")
    npx = width * height
    rgba = torch.zeros(npx * 4, dtype=torch.float32, device="cuda")
    out = {"scene": f"many-instance generator seed 0, instanced: {len(sc.meshes)} meshes "
                    f"({sum(m['indices'].shape[0] for m in sc.meshes)} object-space triangles), "
                    f"{len(sc.instances)} instances, {sc.num_triangles} world triangles",
           "rays": n_rays}
    hits = {}
    for two in (False, True):
        ctx = spk.Context(0)
        ctx.set_stream(stream.cuda_stream)
        t0 = time.perf_counter()
        st = ctx.upload_scene(sc, two_level=two)
        up = time.perf_counter() - t0
        cfg, _ = ctx.autotune(d_rays, min(n_rays, 1 << 20))
        st = ctx.stats()
        for _ in range(2):
            ctx.trace(d_rays, d_hits, n_rays, sync=False)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(10):
            ctx.trace(d_rays, d_hits, n_rays, sync=False)
        b.record(stream)
        torch.cuda.synchronize()
        trace_ms = a.elapsed_time(b) / 10
        hits[two] = d_hits.cpu().numpy().view(spk.HIT_DTYPE)["prim"].copy()
        ctx.set_camera(cam)

        def frame(seed):
            f = spk.FrameDesc()
            f.width, f.height, f.frame_seed, f.max_bounce = width, height, seed, max_bounce
            f.debug_mode, f.sky_enabled, f.flags, f.output_limit = 0, 0, 0, 0.0
            return f

        sched, _ = ctx.autotune_render(frame(10 ** 6), rgba)
        ctx.render_frame(frame(10 ** 6 + 1), rgba)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(spp):
            ctx.render_frame(frame(k), rgba)
        torch.cuda.synchronize()
        frame_ms = (time.perf_counter() - t0) / spp * 1e3
        r = {"upload_s": round(up, 3), "build_s": round(st.build_seconds, 3), "nodes": st.num_nodes,
             "depth": st.depth, "device_bytes": st.device_bytes,
             "trace_config": {"bvh_width": cfg[0], "stack": cfg[1], "block": cfg[2], "refill": cfg[3]},
             "trace_ms": round(trace_ms, 4), "trace_mrays_per_s": round(n_rays / trace_ms / 1e3, 1),
             "frame_ms": round(frame_ms, 3), "samples_per_s": round(npx / frame_ms * 1e3, 1),
             "schedule": dict(zip(("blocks_per_cu", "material_sort", "refill", "tail_bounce"), sched))}
        if two:
            xfs = np.stack([i["M"] for i in sc.instances]).astype(np.float32)
            times = []
            for k in range(3):
                moved = xfs.copy()
                moved[:, 2, 3] += np.float32(0.01 * (k + 1))  # every instance moves up
                t0 = time.perf_counter()
                ctx.update_instances(moved)
                times.append(time.perf_counter() - t0)
            ctx.update_instances(xfs)
            r["update_all_instances_s"] = round(float(np.median(times)), 4)
        ctx.close()
        out["two_level" if two else "flattened"] = r
    out["same_ids_fraction"] = float((hits[False] == hits[True]).mean())
    return out


def c1_experiment(spk, scenes, stream, width=256, height=256, spp=16, max_bounce=4):
    """The reference's own CPU-runnable configuration: the Cornell room (1,316 triangles) at 256x256,
    16 spp, 4 bounces, rendered on the device and by the oracle's restatement of the reference CPU
    path on the host's threads (same per-pixel seeds), both timed; samples/s and the relMSE between
    the two images."""
    import torch
    from oracle import refcpu

    sc = scenes.cornell_scene(0)
    cam = scenes.cornell_camera(width, height)
    ctx = spk.Context(0)
    ctx.set_stream(stream.cuda_stream)
    ctx.upload_scene(sc)
    ctx.set_camera(cam)
    npx = width * height
    rgba = torch.zeros(npx * 4, dtype=torch.float32, device="cuda")
    acc = torch.zeros(npx * 4, dtype=torch.float32, device="cuda")

    def frame(seed):
        f = spk.FrameDesc()
        f.width, f.height, f.frame_seed, f.max_bounce = width, height, seed, max_bounce
        return f

    ctx.render_frame(frame(10 ** 6), rgba)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(spp):
        ctx.render_frame(frame(k), rgba)
        ctx.accumulate(rgba, acc, npx, k, 1)
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    gimg = acc.view(height, width, 4)[..., :3].double().cpu().numpy()
    ctx.close()
    threads = min(len(os.sched_getaffinity(0)), HOST_CORE_SHARE)
    o = refcpu.OracleScene(sc)
    cimg = np.zeros((height, width, 4), np.float32)
    t0 = time.perf_counter()
    for k in range(spp):
        img, _ = o.render_frame(cam, width, height, frame_seed=k, max_bounce=max_bounce, threads=threads)
        refcpu.accumulate(img, cimg, k, 1)
    t_cpu = time.perf_counter() - t0
    return {"scene": "Cornell room 1,316 triangles, 256x256, 16 spp, 4 bounces",
            "gpu_s": round(t_gpu, 4), "gpu_samples_per_s": round(npx * spp / t_gpu, 1),
            "cpu_s": round(t_cpu, 3), "cpu_samples_per_s": round(npx * spp / t_cpu, 1), "cpu_threads": threads,
            "cpu_kind": "port (the oracle's restatement of CPURenderer, same seeds)",
            "relmse_gpu_vs_cpu": relmse(gimg, cimg[..., :3].astype(np.float64))}


def batch_experiment(spk, scenes, stream, count=64, spp=32, max_bounce=6, schedule=None):
    """Many scenes of uneven size in one launch: `count` many-instance scenes (triangle budgets
    log-uniform in [1k, 500k], images 64^2..512^2), `spp` frames of `max_bounce` bounces each, all
    scenes' paths in the same persistent kernels; schedule picked by spk_batch_autotune, or the
    given (blocks_per_cu, material_sort, refill, tail_bounce)."""
    import torch

    log(f"batch experiment: generating {count} scenes")
    items = scenes.batch_scenes(seed=0, count=count)
    ctx = spk.Context(0)
    ctx.set_stream(stream.cuda_stream)
    t0 = time.perf_counter()
    offs = ctx.upload_batch(items)
    t_build = time.perf_counter() - t0
    npx = offs[-1]
    rgba = torch.zeros(npx * 4, dtype=torch.float32, device="cuda")
    acc = torch.zeros(npx * 4, dtype=torch.float32, device="cuda")

    def frame(seed):
        f = spk.FrameDesc()
        f.width, f.height, f.frame_seed, f.max_bounce = 0, 0, seed, max_bounce
        f.debug_mode, f.sky_enabled, f.flags, f.output_limit = 0, 0, 0, 0.0
        return f

    if schedule is None:
        (blocks, srt, refill, tail), tune_ms = ctx.autotune_batch(frame(0), rgba)
    else:
        (blocks, srt, refill, tail), tune_ms = schedule, 0.0
        ctx.set_schedule(*schedule, batch=True)
        ctx.render_batch(frame(10 ** 6), rgba)  # graph capture
    torch.cuda.synchronize()
    rays = 0
    t0 = time.perf_counter()
    for k in range(spp):
        rays += ctx.render_batch(frame(k), rgba).rays_traced
        ctx.accumulate(rgba, acc, npx, k, 1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    img = acc.view(-1, 4)
    finite = bool(torch.isfinite(img).all().item())
    tris = [sc.num_triangles for sc, _, _, _ in items]
    out = {
        "scenes": count,
        "triangles_total": int(sum(tris)),
        "triangles_min_max": [int(min(tris)), int(max(tris))],
        "pixels_per_frame": npx,
        "spp": spp,
        "max_bounce": max_bounce,
        "seconds": round(dt, 3),
        "samples_per_s": round(npx * spp / dt, 1),
        "mrays_per_s": round(rays / dt / 1e6, 1),
        "schedule": {"blocks_per_cu": blocks, "material_sort": bool(srt), "refill": refill, "tail_bounce": tail,
                     "tuned_frame_ms": round(tune_ms, 3)},
        "build_s": round(t_build, 2),
        "finite": finite,
    }
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays", type=int, default=8 * 1024 * 1024)
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-coherent", action="store_true")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-buffer (spk_trace_closest_host) measurement")
    ap.add_argument("--config", type=str, default=None,
                    help="traversal variant width,stack,block[,refill[,ray_sort]]")
    ap.add_argument("--leaf", type=int, default=3, help="triangles per BVH leaf slot (1..3)")
    ap.add_argument("--relmse", type=str, default="1920x1080,64,4096",
                    help="equal-time relMSE experiment WxH,spp,reference-spp ('off' to skip)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batch measurement")
    ap.add_argument("--trace-only", action="store_true",
                    help="only the closest-hit measurement (for counter runs): no cache, path-trace, batch or relMSE")
    ap.add_argument("--render-only", type=int, default=0,
                    help="only run the full path trace of the many-instance scene for N spp (one JSON line)")
    ap.add_argument("--render-mode", choices=sorted(RENDER_MODES), default="plain",
                    help="estimator of --render-only")
    ap.add_argument("--render-schedule", type=str, default=None,
                    help="blocks_per_cu,material_sort,refill,tail_bounce for --render-only and --batch-only "
                         "(default: autotuned)")
    ap.add_argument("--two-level", action="store_true",
                    help="only the flattened vs two-level comparison on the instanced scene (one JSON line)")
    ap.add_argument("--relmse-only", action="store_true",
                    help="only the equal-time relMSE experiment (one JSON line)")
    ap.add_argument("--cache-kernels", action="store_true",
                    help="only time the learned cache's inference and training kernels (one JSON line)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the 64-scene batch measurement (one JSON line)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run N device copies of 256 MiB (counter calibration for the traffic figure)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(args.copy_calibration):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel(), "copy_bytes_written": b.numel(),
                          "copies": args.copy_calibration}))
        return
    if args.render_only:
        import torch
        import sparkle_amd as spk
        from sparkle_amd import scenes

        stream = torch.cuda.Stream()
        torch.cuda.set_stream(stream)
        fixed = tuple(int(x) for x in args.render_schedule.split(",")) if args.render_schedule else None
        print(json.dumps({"full_path": render_experiment(spk, scenes, stream, spp=args.render_only,
                                                         mode=args.render_mode, schedule=fixed)}), flush=True)
        return
    if args.two_level:
        import torch
        import sparkle_amd as spk
        from sparkle_amd import scenes

        stream = torch.cuda.Stream()
        torch.cuda.set_stream(stream)
        print(json.dumps({"two_level": two_level_experiment(spk, scenes, stream)}), flush=True)
        return
    if args.relmse_only:
        import torch
        import sparkle_amd as spk
        from sparkle_amd import scenes

        stream = torch.cuda.Stream()
        torch.cuda.set_stream(stream)
        res, spp, ref_spp = args.relmse.split(",")
        w_, h_ = (int(x) for x in res.split("x"))
        print(json.dumps({"relmse": relmse_experiment(spk, scenes, stream, w_, h_, int(spp), int(ref_spp))}),
              flush=True)
        return
    if args.cache_kernels:
        import torch
        import sparkle_amd as spk
        from sparkle_amd import scenes

        stream = torch.cuda.Stream()
        torch.cuda.set_stream(stream)
        print(json.dumps({"cache_kernels": cache_kernel_bench(spk, scenes, stream)}), flush=True)
        return
    if args.batch_only:
        import torch
        import sparkle_amd as spk
        from sparkle_amd import scenes

        stream = torch.cuda.Stream()
        torch.cuda.set_stream(stream)
        fixed = tuple(int(x) for x in args.render_schedule.split(",")) if args.render_schedule else None
        print(json.dumps({"batch": batch_experiment(spk, scenes, stream, schedule=fixed)}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards by image tiles or scenes, "
              f"see the design notes", file=sys.stderr)
        sys.exit(2)

    import torch
    import sparkle_amd as spk
    from sparkle_amd import scenes

    log("generating heightfield and rays")
    sc = scenes.heightfield_scene(cells=args.cells)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=args.rays, seed=0)
    n = rays.shape[0]

    ctx = spk.Context(0)
    # a dedicated (default-priority) torch stream, so torch events time the kernels launched on it
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx.set_stream(stream.cuda_stream)
    assert ctx.stream() == stream.cuda_stream
    log("building BVH and uploading")
    st = ctx.upload_scene(sc, max_leaf_tris=args.leaf)
    d_rays = torch.from_numpy(rays.view(np.uint8).reshape(-1)).cuda()
    d_hits = torch.empty(n * 8, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()

    tune = None
    if args.config:
        parts = [int(x) for x in args.config.split(",")]
        w_, s_, b_ = parts[:3]
        r_ = parts[3] if len(parts) > 3 else 0
        o_ = parts[4] if len(parts) > 4 else 0
        ctx.set_trace_config(w_, s_, b_, r_, o_)
        st = ctx.stats()
        tune = {"bvh_width": w_, "stack": s_, "block": b_, "refill": r_, "ray_sort": o_, "fixed": True}
    elif not args.no_autotune:
        log("autotuning traversal variants")
        # the whole batch: the coherence order pays off only with enough rays per bin
        cfg, ms = ctx.autotune(d_rays, n, tune_width=True)
        tune = {"bvh_width": cfg[0], "stack": cfg[1], "block": cfg[2], "refill": cfg[3], "ray_sort": cfg[4]}
        st = ctx.stats()

    # traversal counts for the algorithmic-bytes model
    cnt_nodes = torch.empty(n, dtype=torch.int32, device="cuda")
    cnt_tris = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.trace_count(d_rays, cnt_nodes, cnt_tris, n)
    mean_nodes = cnt_nodes.double().mean().item()
    mean_tris = cnt_tris.double().mean().item()
    bytes_per_ray = 32 + 8 + st.node_bytes * mean_nodes + st.triangle_bytes * mean_tris
    sort_bytes_per_ray = 0
    if ctx.trace_config()[4]:
        # the traversal kernel reads its ray's index in the coherence order; the two sort passes before
        # it read each ray once more and write its key (36 B), then move the key and index by
        # direction bin (4 in, 8 out), count the keys again (4) and move the index by full key (8 in, 4 out)
        bytes_per_ray += 4
        sort_bytes_per_ray = 36 + 12 + 4 + 12
    bytes_per_launch = bytes_per_ray * n

    for _ in range(args.warmup):
        ctx.trace(d_rays, d_hits, n, sync=False)
    ctx.synchronize()
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    # events on the context stream around the traversal kernel of each timed call (a coherence sort,
    # when the configuration has one, runs before it and stays outside)
    ctx.trace_kernel_timing(args.steps)
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        rc = ctx.trace(d_rays, d_hits, n, sync=False)
        if rc != spk.SPK_OK:
            raise RuntimeError(f"trace failed {rc}")
        ends[k].record(stream)
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ctx.trace_kernel_times()
    assert len(kernel_ms) == args.steps
    ctx.trace_kernel_timing(0)
    avg_ms = float(np.mean(kernel_ms))
    region_ms = starts[0].elapsed_time(ends[-1])
    value = n * args.steps / (region_ms / 1e3) / 1e6

    hits = d_hits.cpu().numpy().view(spk.HIT_DTYPE)
    hit_frac = float((hits["prim"] != spk.SPK_MISS).mean())
    log("checking the timed call's hits against the oracle")
    parity = oracle_check(sc, rays, hits)

    # coherent 1080p primary rays on the same scene (reported beside the metric)
    coherent_mrays = None
    if not args.no_coherent:
        rays_a = scenes.coherent_rays()
        d_rays_a = torch.from_numpy(rays_a.view(np.uint8).reshape(-1)).cuda()
        d_hits_a = torch.empty(rays_a.shape[0] * 8, dtype=torch.uint8, device="cuda")
        for _ in range(2):
            ctx.trace(d_rays_a, d_hits_a, rays_a.shape[0], sync=False)
        torch.cuda.synchronize()
        ea, eb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ea.record(stream)
        for _ in range(10):
            ctx.trace(d_rays_a, d_hits_a, rays_a.shape[0], sync=False)
        eb.record(stream)
        torch.cuda.synchronize()
        coherent_mrays = round(rays_a.shape[0] * 10 / (ea.elapsed_time(eb) / 1e3) / 1e6, 1)

    # the same rays handed over in pageable host memory (spk_trace_closest_host: 256 MiB of rays in
    # over PCIe through the pinned staging ring, the traversal, 64 MiB of hits out into a hit array
    # the caller reuses); reported beside the metric, never as it. The first call sets the ring up.
    # (--no-host skips it: its chunked traversal calls would mix with the full-size launches in a
    # profiler's per-kernel averages)
    host_ms = []
    host_mrays = None
    if not args.no_host:
        hits_h = np.zeros(n, dtype=spk.HIT_DTYPE)
        for _ in range(5):
            t_h = time.perf_counter()
            rc_h, hits_h = ctx.trace_host(rays, hits_h)
            host_ms.append((time.perf_counter() - t_h) * 1e3)
            if rc_h != spk.SPK_OK:
                raise RuntimeError(f"host-buffer trace failed {rc_h}")
        host_mrays = round(n / (float(np.median(host_ms)) / 1e3) / 1e6, 1)
        if not np.array_equal(hits_h.view(np.uint8).reshape(-1), d_hits.cpu().numpy()):
            raise RuntimeError("host-buffer hits differ from the device-buffer hits")

    achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9
    cur = ctx.trace_config()[:4] if not ctx.trace_config()[4] else ctx.trace_config()
    traffic, l2_hit, traffic_src = load_traffic(",".join(str(x) for x in cur))
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(region_ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"incoherent extend rays: {n} rays (origins uniform in the scene box, directions uniform "
                        f"on the sphere, seed 0) vs a {args.cells}x{args.cells}-cell value-noise heightfield "
                        f"({st.num_triangles} triangles), closest hit with id and t",
            "bvh": {"width": st.bvh_width, "nodes": st.num_nodes, "depth": st.depth, "node_bytes": st.node_bytes,
                    "triangle_bytes": st.triangle_bytes, "sah": round(st.sah_cost, 3),
                    "build_s": round(st.build_seconds, 3)},
            "autotuned": tune,
            "hit_fraction": round(hit_frac, 4),
            "coherent_1080p_mrays": coherent_mrays,
            "host_buffers_mrays": host_mrays,
            "host_buffers_ms": round(float(np.median(host_ms)), 3) if host_ms else None,
            "mean_nodes_per_ray": round(mean_nodes, 3),
            "mean_tris_per_ray": round(mean_tris, 3),
            "wall_s": round(wall, 4),
            "parity": parity,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_ray": round(bytes_per_ray, 1),
            "compulsory_bytes_per_ray": 40,
            "sort_bytes_per_ray": sort_bytes_per_ray,
            "kernel_avg_ms": round(avg_ms, 4),
            "kernel": f"k_trace<{ctx.trace_config()[0]}, {ctx.trace_config()[1]}, {ctx.trace_config()[2]}, false, "
                      f"false>" if not ctx.trace_config()[3] else "k_trace_persistent",
            "traffic_source": traffic_src,
            "l2_hit_rate": l2_hit,
        },
    }
    if not args.no_cpu:
        log("timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(sc, rays)
    ctx.close()
    if args.trace_only:
        print(json.dumps(out), flush=True)
        return
    out["cache_kernels"] = cache_kernel_bench(spk, scenes, stream)
    out["full_path"] = render_experiment(spk, scenes, stream)
    out["c1"] = c1_experiment(spk, scenes, stream)
    out["two_level"] = two_level_experiment(spk, scenes, stream)
    if not args.no_batch:
        out["batch"] = batch_experiment(spk, scenes, stream)
    if args.relmse != "off":
        res, spp, ref_spp = args.relmse.split(",")
        w_, h_ = (int(x) for x in res.split("x"))
        out["relmse"] = relmse_experiment(spk, scenes, stream, w_, h_, int(spp), int(ref_spp))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
