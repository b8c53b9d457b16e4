"""The benchmark's path-tracing configurations at their own scene sizes (MI355X).

* The many-instance scene (400 instances flattened to 5,533,044 triangles: icospheres, tori and
  displaced boxes, 60% Lambertian, 30% GGX metal, 10% IOR-1.5 glass, three area lights), 8 bounces
  with Russian roulette, at a reduced 240x136 image: per pixel within 3 sigma of the Monte Carlo
  error against the oracle, identical sample paths on most pixels, relMSE of the two means against
  its Monte Carlo floor.
* The 64-scene batch (triangle budgets log-uniform in [1k, 500k], images of 64^2-512^2, 6 bounces)
  in one launch under the autotuned batch schedule: range properties on every scene, every 8th scene
  bit-identical to the same scene rendered alone, and the smallest scene against the oracle.
* The learned cache at its benchmark size (16 levels of 2^19 entries, 2^16 training samples per
  frame) trained online on the many-instance scene: forward and gradients of the trained network
  against the FP64 oracle on a sample of queries in the scene.
"""
import math

import numpy as np
import pytest

import sparkle_amd as spk
from sparkle_amd import scenes
from oracle import refcpu
from tests.conftest import print_case
from tests.test_render_gpu import Renderer, frame

pytestmark = pytest.mark.gpu


def _paired(frames_g, frames_c):
    d = np.stack(frames_g) - np.stack(frames_c)
    mg, mc = np.stack(frames_g).mean(0), np.stack(frames_c).mean(0)
    se = d.std(0, ddof=1) / math.sqrt(d.shape[0])
    within = (np.abs(d.mean(0)) <= 3 * se + 1e-6).mean()
    same = (np.abs(d).max(-1) < 1e-3).mean()
    rel_mse = float(np.mean((mg - mc) ** 2 / (mc ** 2 + 1e-2)))
    floor = float(np.mean(d.var(0, ddof=1) / d.shape[0] / (mc ** 2 + 1e-2)))
    return within, same, rel_mse, floor, mg, mc


def test_many_instance_scene_matches_oracle(gpu_ctx):
    w, h = 240, 136
    sc, cam = scenes.many_instance_scene(seed=0, aspect=w / h)
    assert sc.num_triangles == 5_533_044
    print_case("many-instance upload")
    r = Renderer(gpu_ctx, sc, cam, w, h)
    o = refcpu.OracleScene(sc)
    frames = 16
    g, c = [], []
    for k in range(frames):
        print_case("many-instance frame", k)
        g.append(r.render(k, bounce=8)[0][..., :3])
        c.append(o.render_frame(cam, w, h, frame_seed=k, max_bounce=8)[0][..., :3])
    within, same, rel_mse, floor, mg, mc = _paired(g, c)
    print(f"many-instance 240x136: within 3 sigma {within:.4f}, identical paths {same:.4f}, relMSE {rel_mse:.3g}, "
          f"Monte Carlo floor {floor:.3g}, flux {mg.mean():.5f} vs {mc.mean():.5f}")
    assert within >= 0.99
    assert same > 0.8
    assert rel_mse < 2.0 * floor + 1e-5
    # glass and metal are in view: first-hit albedo equals the oracle's, with glass pixels present
    ga, _, _ = r.render(0, debug=spk.DEBUG_ALBEDO)
    ca, _ = o.render_frame(cam, w, h, frame_seed=0, debug_mode=spk.DEBUG_ALBEDO)
    assert (np.abs(ga[..., :3] - ca[..., :3]).max(-1) < 1e-4).mean() > 0.99
    gm, _, _ = r.render(0, debug=spk.DEBUG_METALLIC)
    assert (gm[..., 0] == 1.0).mean() > 0.05


def test_batch_of_64_scenes_at_size():
    import torch

    items = scenes.batch_scenes(seed=0)
    assert len(items) == 64
    ctx = spk.Context(0)
    try:
        print_case("batch upload 64 scenes")
        offs = ctx.upload_batch(items)
        rgba = torch.zeros(offs[-1] * 4, dtype=torch.float32, device="cuda")
        print_case("batch autotune")
        ctx.autotune_batch(frame(0, 0, 0, bounce=6), rgba)
        seeds = (0, 7)
        batch = {}
        for s in seeds:
            print_case("batch frame", s)
            ctx.render_batch(frame(0, 0, s, bounce=6), rgba)
            img = rgba.cpu().numpy()
            assert np.isfinite(img).all()
            px = img.reshape(-1, 4)
            assert np.isin(px[:, 3], (-1.0, 1.0)).all()
            assert (px[:, :3] >= 0).all() and (px[:, :3] <= 6.0).all()
            for k in range(len(items)):
                batch[k, s] = img[offs[k] * 4:offs[k + 1] * 4].copy()
                if s == seeds[0]:
                    assert (batch[k, s].reshape(-1, 4)[:, 3] == 1.0).mean() > 0.5, k
        alone = spk.Context(0)
        try:
            for k in range(0, 64, 8):
                sc, cam, w, h = items[k]
                alone.upload_scene(sc)
                alone.set_camera(cam)
                one = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
                for s in seeds:
                    print_case("batch scene alone", k, "seed", s)
                    alone.render_frame(frame(w, h, s, bounce=6), one)
                    assert np.array_equal(one.cpu().numpy(), batch[k, s]), (k, s)
        finally:
            alone.close()
        # the scene with the fewest triangles among the 128^2 ones against the oracle, under a sky so
        # that most samples carry light, 24 frames of the whole batch
        k = min((i for i in range(64) if items[i][2] == 128), key=lambda i: items[i][0].num_triangles)
        sc, cam, w, h = items[k]
        o = refcpu.OracleScene(sc)
        g, c = [], []
        sky = (0.6, 0.7, 0.9)
        for s in range(24):
            print_case("batch oracle frame", s, "scene", k)
            ctx.render_batch(frame(0, 0, 100 + s, bounce=6, sky=sky), rgba)
            g.append(rgba[offs[k] * 4:offs[k + 1] * 4].cpu().numpy().reshape(h, w, 4)[..., :3])
            c.append(o.render_frame(cam, w, h, frame_seed=100 + s, max_bounce=6, sky=sky)[0][..., :3])
        assert (np.stack(c) > 0).mean() > 0.5
        within, same, rel_mse, floor, _, _ = _paired(g, c)
        print(f"batch scene {k} ({sc.num_triangles} triangles, {w}x{h}): within 3 sigma {within:.4f}, "
              f"identical paths {same:.4f}, relMSE {rel_mse:.3g} (floor {floor:.3g})")
        assert within >= 0.99 and same > 0.8 and rel_mse < 2.0 * floor + 1e-5
    finally:
        ctx.close()


def test_cache_at_benchmark_size_matches_fp64(gpu_ctx):
    import torch

    from tests.test_cache_gpu import oracle_forward, random_attrs, to_dev

    w, h = 240, 136
    sc, cam = scenes.many_instance_scene(seed=0, aspect=w / h)
    print_case("cache config upload")
    gpu_ctx.upload_scene(sc)
    gpu_ctx.set_camera(cam)
    cfgc = spk.cache_config()
    assert cfgc.log2_table == 19 and cfgc.batch == 1 << 16
    gpu_ctx.cache_create(cfgc)
    rgba = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
    losses = []
    for k in range(24):
        print_case("cache config frame", k)
        f = frame(w, h, k, bounce=8, flags=spk.FRAME_CACHE | spk.FRAME_NEE)
        gpu_ctx.render_frame(f, rgba)
        losses.append(gpu_ctx.cache_info()["last_loss"])
    info = gpu_ctx.cache_info()
    assert info["steps"] >= 20 and np.isfinite(losses).all()
    assert np.isfinite(rgba.cpu().numpy()).all()
    wts, tab = gpu_ctx.cache_get_params()
    cfg = refcpu.cache_cfg(log2_table=19, inputs=spk.CACHE_INPUTS)
    rng = np.random.Generator(np.random.PCG64(11))
    n = 256
    pts = rng.random((n, 3)).astype(np.float32)  # the cache's own [0,1]^3 frame of the scene box
    at = random_attrs(rng, n)
    out = torch.zeros(n * 3, device="cuda")
    print_case("cache config forward")
    gpu_ctx.cache_infer(to_dev(pts.reshape(-1)), out, n, attrs_dev=to_dev(at.reshape(-1)))
    g = out.cpu().numpy().reshape(n, 3)
    targets = rng.random((n, 3)).astype(np.float32)
    ref_loss, ref, ref_dw, ref_dx, _ = oracle_forward(cfg, wts, tab, pts, targets, at)
    err = np.abs(g - ref)
    assert (err <= 0.03 * np.abs(ref) + 2e-3).mean() > 0.99, err.max()
    print_case("cache config gradients")
    loss, dw, dtab = gpu_ctx.cache_gradients(to_dev(pts.reshape(-1)), to_dev(targets.reshape(-1)), n, table=True,
                                             attrs_dev=to_dev(at.reshape(-1)))
    assert abs(loss - ref_loss) <= 0.03 * abs(ref_loss)
    assert np.linalg.norm(dw - ref_dw) / np.linalg.norm(ref_dw) < 0.05
    ref_tab = refcpu.table_backward(cfg, pts, ref_dx)
    assert np.linalg.norm(dtab - ref_tab) / np.linalg.norm(ref_tab) < 0.05
    gpu_ctx.cache_destroy()


def test_path_with_a_nan_direction_does_not_walk_the_tree(gpu_ctx):
    """Frame seed 1,000,559 of the 1080p many-instance render holds a path whose refraction past the
    critical angle met a Fresnel draw of exactly 1 (RandomUnit reaches 1.0), so its next direction is
    NaN, as in the reference. Such a ray hits nothing; it used to open every box of the tree (NaN slab
    distances) and cost ~10 s of one lane. The frame now takes as long as its neighbours."""
    import torch

    w, h = 1920, 1080
    sc, cam = scenes.many_instance_scene(seed=0, aspect=w / h)
    print_case("nan-direction frame upload")
    gpu_ctx.upload_scene(sc)
    gpu_ctx.set_camera(cam)
    gpu_ctx.set_schedule(8, 0, 0, 1)
    rgba = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
    times = []
    for seed in (999_997, 1_000_558, 1_000_559, 1_000_560):
        print_case("nan-direction frame", seed)
        st = gpu_ctx.render_frame(frame(w, h, seed, bounce=8, flags=spk.FRAME_SEED64, limit=1e30), rgba)
        times.append(st.milliseconds)
    print("frame ms", times)
    assert times[2] < 5 * max(times[1], times[3]) and times[2] < 1000.0
