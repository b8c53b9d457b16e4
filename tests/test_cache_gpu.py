"""Learned radiance cache on the MI355X against the FP64 oracle: BF16-MFMA forward and backward,
the hash-grid gradient scatter, Adam, and online fitting.

The network's input is the 32 hash-grid features of the position, then the query's 8 surface
attributes (shading normal, albedo, roughness, 1), then zeros up to 64.

Tolerances: the device computes the MLP on BF16 operands (8-bit mantissa) with FP32 accumulation,
so outputs are compared to the FP64 oracle (fed the same FP16 table and BF16-rounded weights) within
3% relative + 2e-3 absolute, and gradients by relative vector error: 5% over all weights, 8% per
layer, 2% for the table gradient of 4 samples and 8% for that of 300 (where the BF16 accumulation
of many contributions per coarse entry dominates).
"""
import numpy as np
import pytest

import sparkle_amd as spk
from sparkle_amd import scenes
from oracle import refcpu
from tests.conftest import print_case

pytestmark = pytest.mark.gpu

LOG2 = 14  # a small table keeps the FP64 oracle fast


def to_dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def setup_cache(ctx, seed=0, log2=LOG2, lr=1e-2):
    sc = scenes.cornell_scene()
    ctx.upload_scene(sc)
    ctx.cache_create(spk.cache_config(log2_table=log2, learning_rate=lr, seed=seed))
    cfg = refcpu.cache_cfg(log2_table=log2, inputs=spk.CACHE_INPUTS)
    rng = np.random.Generator(np.random.PCG64(seed + 100))
    info = ctx.cache_info()
    # larger weights and table values than the initialisation, so every layer is exercised
    w = rng.normal(0, 0.15, info["num_weights"]).astype(np.float32)
    t = rng.normal(0, 0.5, info["num_table_params"]).astype(np.float32)
    ctx.cache_set_params(w, t)
    return cfg, w, t


def random_attrs(rng, n):
    """Per-query attributes as the path tracer writes them: a unit normal, an albedo, a roughness, 1."""
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return np.concatenate([nrm, rng.uniform(0.05, 0.9, (n, 3)), rng.uniform(0, 1, (n, 1)), np.ones((n, 1))],
                          1).astype(np.float32)


def oracle_forward(cfg, w, t, pts, targets, attrs=None):
    """FP64 loss, outputs and gradients of the network on the device's BF16 operands: the FP16 table's
    features and the attributes rounded to BF16, zero padding to the input width."""
    t16 = t.astype(np.float16).astype(np.float32)
    feats = refcpu.hashgrid_encode(cfg, t16, pts)
    n = pts.shape[0]
    x = np.zeros((n, cfg.inputs or feats.shape[1]), np.float32)
    x[:, :feats.shape[1]] = feats
    if attrs is not None:
        x[:, feats.shape[1]:feats.shape[1] + attrs.shape[1]] = attrs
    feats_bf = refcpu.to_bf16(x).astype(np.float64)
    wb = refcpu.to_bf16(w).astype(np.float64)
    return refcpu.mlp_loss_grad(cfg, wb, feats_bf, targets) + (feats,)


def test_forward_matches_fp64(gpu_ctx):
    import torch

    cfg, w, t = setup_cache(gpu_ctx)
    rng = np.random.Generator(np.random.PCG64(5))
    n = 777  # not a multiple of the 64-query tile
    pts = rng.random((n, 3)).astype(np.float32)
    at = random_attrs(rng, n)
    out = torch.zeros(n * 3, device="cuda")
    print_case("cache forward", n)
    gpu_ctx.cache_infer(to_dev(pts.reshape(-1)), out, n, attrs_dev=to_dev(at.reshape(-1)))
    g = out.cpu().numpy().reshape(n, 3)
    _, ref, _, _, _ = oracle_forward(cfg, w, t, pts, np.zeros((n, 3)), at)
    err = np.abs(g - ref)
    assert (err <= 0.03 * np.abs(ref) + 2e-3).mean() > 0.995, err.max()
    assert np.abs(ref).mean() > 0.05  # the test is not comparing zeros


def test_gradients_match_fp64(gpu_ctx):
    cfg, w, t = setup_cache(gpu_ctx, seed=1)
    rng = np.random.Generator(np.random.PCG64(6))
    n = 300
    pts = rng.random((n, 3)).astype(np.float32)
    targets = rng.random((n, 3)).astype(np.float32)
    at = random_attrs(rng, n)
    print_case("cache gradients", n)
    loss, dw, dtab = gpu_ctx.cache_gradients(to_dev(pts.reshape(-1)), to_dev(targets.reshape(-1)), n, table=True,
                                             attrs_dev=to_dev(at.reshape(-1)))
    ref_loss, out, ref_dw, ref_dx, feats = oracle_forward(cfg, w, t, pts, targets, at)
    assert abs(loss - ref_loss) <= 0.03 * abs(ref_loss)
    rel = np.linalg.norm(dw - ref_dw) / np.linalg.norm(ref_dw)
    assert rel < 0.05, rel
    # per layer as well (the output layer has only 192 weights)
    offs = [0, 4096, 8192, 12288, 16384, 16576]
    for a, b in zip(offs[:-1], offs[1:]):
        r = np.linalg.norm(dw[a:b] - ref_dw[a:b]) / max(np.linalg.norm(ref_dw[a:b]), 1e-30)
        assert r < 0.08, (a, r)
    # At 300 samples the coarse levels' entries sum hundreds of mixed-sign contributions in BF16
    # (8-bit mantissa, by design: Adam normalises by the gradient's RMS), which dominates the table
    # gradient's error (4.7% here; element-wise the median is 0.7%).
    ref_tab = refcpu.table_backward(cfg, pts, ref_dx)
    rel_t = np.linalg.norm(dtab - ref_tab) / np.linalg.norm(ref_tab)
    assert rel_t < 0.08, rel_t


def test_table_gradient_of_few_samples_matches_fp64(gpu_ctx):
    """With 4 samples almost every table entry takes one contribution, so the BF16 accumulation does
    not enter: the table gradient pins d(loss)/d(features) of the backward pass, the trilinear
    weights and the corner indices of the scatter (0.45% measured; every touched entry must be the
    oracle's)."""
    cfg, w, t = setup_cache(gpu_ctx, seed=1)
    rng = np.random.Generator(np.random.PCG64(6))
    n = 4
    pts = rng.random((n, 3)).astype(np.float32)
    targets = rng.random((n, 3)).astype(np.float32)
    at = random_attrs(rng, n)
    print_case("cache table gradient", n)
    _, _, dtab = gpu_ctx.cache_gradients(to_dev(pts.reshape(-1)), to_dev(targets.reshape(-1)), n, table=True,
                                         attrs_dev=to_dev(at.reshape(-1)))
    _, _, _, ref_dx, _ = oracle_forward(cfg, w, t, pts, targets, at)
    ref_tab = refcpu.table_backward(cfg, pts, ref_dx)
    np.testing.assert_array_equal(dtab != 0, ref_tab != 0)
    assert np.linalg.norm(dtab - ref_tab) / np.linalg.norm(ref_tab) < 0.02


def test_adam_first_step_is_sign_of_gradient(gpu_ctx):
    lr = 1e-3
    cfg, w, t = setup_cache(gpu_ctx, seed=2, lr=lr)
    rng = np.random.Generator(np.random.PCG64(7))
    n = 512
    pts = to_dev(rng.random((n, 3)).reshape(-1))
    targets = to_dev(rng.random((n, 3)).reshape(-1))
    _, dw, dtab = gpu_ctx.cache_gradients(pts, targets, n, table=True)
    w0, t0 = gpu_ctx.cache_get_params()
    gpu_ctx.cache_train(pts, targets, n)
    w1, t1 = gpu_ctx.cache_get_params()
    # first Adam step: m_hat = g, v_hat = g^2 -> delta = -lr * g / (|g| + eps)
    big = np.abs(dw) > 1e-5
    np.testing.assert_allclose(w1[big] - w0[big], -lr * np.sign(dw[big]), rtol=1e-3, atol=1e-7)
    bigt = np.abs(dtab) > 1e-5
    np.testing.assert_allclose(t1[bigt] - t0[bigt], -lr * np.sign(dtab[bigt]), rtol=1e-3, atol=1e-7)
    assert np.array_equal(t1[dtab == 0], t0[dtab == 0])


def test_table_adam_is_lazy_and_consumes_its_gradient(gpu_ctx):
    """The table's Adam skips entries (feature pairs) whose gradient is zero (torch.optim.SparseAdam's
    rule): after a step on points A and a step on disjoint points B, an entry A touched and B did not
    keeps its value, although its first moment is non-zero (plain Adam would move it again). The step also
    clears the gradient it read: a gradient pass on points C afterwards is non-zero only where C's
    own gradient is, with no trace of B's."""
    print_case("cache lazy table Adam")
    setup_cache(gpu_ctx, seed=5, lr=1e-3)
    rng = np.random.Generator(np.random.PCG64(11))
    n = 512
    pa = rng.uniform(0.05, 0.45, (n, 3))
    pb = rng.uniform(0.55, 0.95, (n, 3))
    pc = rng.uniform(0.05, 0.95, (64, 3))
    targets = to_dev(rng.random((n, 3)).reshape(-1))
    _, _, ga = gpu_ctx.cache_gradients(to_dev(pa.reshape(-1)), targets, n, table=True)
    _, _, gb = gpu_ctx.cache_gradients(to_dev(pb.reshape(-1)), targets, n, table=True)
    _, _, gc_fresh = gpu_ctx.cache_gradients(to_dev(pc.reshape(-1)), targets, 64, table=True)
    gpu_ctx.cache_train(to_dev(pa.reshape(-1)), targets, n)
    _, t1 = gpu_ctx.cache_get_params()
    gpu_ctx.cache_train(to_dev(pb.reshape(-1)), targets, n)
    _, t2 = gpu_ctx.cache_get_params()
    # an entry is its two features (one BF16 gradient pair); it is skipped when both are zero
    ga2, gb2 = ga.reshape(-1, 2), gb.reshape(-1, 2)
    only_a = (ga2 != 0).any(1) & (gb2 == 0).all(1)
    assert only_a.sum() > 10000
    np.testing.assert_array_equal(t2.reshape(-1, 2)[only_a], t1.reshape(-1, 2)[only_a])
    moved = np.abs(gb) > 1e-5
    assert (t2[moved] != t1[moved]).mean() > 0.99
    _, _, gc = gpu_ctx.cache_gradients(to_dev(pc.reshape(-1)), targets, 64, table=True)
    assert (gc != 0).sum() > 0
    assert not ((gc != 0) & (gc_fresh == 0)).any()


def test_online_training_fits_an_analytic_field(gpu_ctx):
    """Fitting f(x) = a smooth positive RGB field: the relative L2 loss falls tenfold and the
    trained network beats the untrained one on held-out points."""
    import torch

    sc = scenes.cornell_scene()
    gpu_ctx.upload_scene(sc)
    gpu_ctx.cache_create(spk.cache_config(log2_table=16, learning_rate=1e-2, seed=3))

    def field(p):
        return np.stack([1.0 + np.sin(6 * p[:, 0]) * np.cos(4 * p[:, 1]),
                         0.5 + 0.4 * p[:, 2] ** 2,
                         0.8 + 0.5 * np.sin(9 * p[:, 0] * p[:, 1])], 1).astype(np.float32)

    rng = np.random.Generator(np.random.PCG64(8))
    held = rng.random((4096, 3)).astype(np.float32)
    held_t = field(held)

    def held_error():
        out = torch.zeros(held.shape[0] * 3, device="cuda")
        gpu_ctx.cache_infer(to_dev(held.reshape(-1)), out, held.shape[0])
        y = out.cpu().numpy().reshape(-1, 3)
        return float(np.mean((y - held_t) ** 2 / (held_t ** 2 + 1e-2)))

    e0 = held_error()
    losses = []
    for step in range(300):
        p = rng.random((1 << 14, 3)).astype(np.float32)
        losses.append(gpu_ctx.cache_train(to_dev(p.reshape(-1)), to_dev(field(p).reshape(-1)), p.shape[0]))
    e1 = held_error()
    print("loss first/last", losses[0], losses[-1], "held-out relL2", e0, e1)
    assert losses[-1] < 0.1 * losses[0]
    assert e1 < 0.1 * e0 and e1 < 0.02
    assert gpu_ctx.cache_info()["steps"] == 300
    gpu_ctx.cache_destroy()


def test_cache_render_trains_online_and_stays_close_to_plain(gpu_ctx):
    """The path tracer terminates paths into the cache and trains it on its own samples: the loss
    drops from its untrained value, the converged image keeps the plain tracer's flux, and the
    per-pixel variance is below the plain tracer's at equal samples (what the trained network gains
    over tracing the tail), more so when paths terminate after the first bounce."""
    import torch

    w = h = 64
    sc = scenes.cornell_scene()
    gpu_ctx.upload_scene(sc)
    gpu_ctx.set_camera(scenes.cornell_camera(w, h))
    rgba = torch.zeros(w * h * 4, device="cuda")

    def frame(seed, flags):
        f = spk.FrameDesc()
        f.width, f.height, f.frame_seed, f.max_bounce, f.flags = w, h, seed, 6, flags
        return f

    plain = []
    for k in range(200):
        gpu_ctx.render_frame(frame(1000 + k, 0), rgba)
        plain.append(rgba.cpu().numpy().reshape(h, w, 4)[..., :3].copy())
    plain = np.array(plain)
    for bounce in (2, 1):
        print_case("cache render, terminate bounce", bounce)
        gpu_ctx.cache_create(spk.cache_config(log2_table=16, batch=1 << 12, seed=4, terminate_bounce=bounce))
        losses = []
        cache_imgs = []
        for k in range(300):
            gpu_ctx.render_frame(frame(k, spk.FRAME_CACHE), rgba)
            losses.append(gpu_ctx.cache_info()["last_loss"])
            if k >= 100:
                cache_imgs.append(rgba.cpu().numpy().reshape(h, w, 4)[..., :3].copy())
        cache_imgs = np.array(cache_imgs)
        mc, mp = cache_imgs.mean(0), plain.mean(0)
        vc, vp = cache_imgs.var(0).mean(), plain.var(0).mean()
        print("bounce", bounce, "loss first/median late", losses[1], np.median(losses[-100:]),
              "flux cache/plain", mc.mean(), mp.mean(), "variance cache/plain", vc, vp)
        # each frame trains on the previous frame's samples, so the first frame takes no step
        assert losses[1] > 0.0
        assert losses[1] > 3 * np.median(losses[-100:])
        assert abs(mc.mean() - mp.mean()) / mp.mean() < 0.1
        assert gpu_ctx.cache_info()["steps"] >= 290
        # terminating earlier removes more of the tail's variance; what remains is the direct hits
        # of the small light at the first vertices, which the cache does not replace
        if bounce == 1:
            assert vc < 0.9 * vp
        else:
            assert vc < vp
        # reset: back to the initial parameters, same buffers
        gpu_ctx.cache_reset()
        assert gpu_ctx.cache_info()["steps"] == 0
        gpu_ctx.cache_destroy()


def test_training_paths_end_in_the_cache_after_their_suffix(gpu_ctx):
    """Self-training: a training path ends in the cache train_suffix hit indices after
    terminate_bounce (one in 16 excepted), so a frame with a suffix traces fewer rays than one whose
    training paths run to their end, and fewer with a short suffix than with a long one (the same
    paths otherwise: same seeds)."""
    import torch

    w = h = 64
    sc = scenes.cornell_scene()
    gpu_ctx.upload_scene(sc)
    gpu_ctx.set_camera(scenes.cornell_camera(w, h))
    rgba = torch.zeros(w * h * 4, device="cuda")
    f = spk.FrameDesc()
    f.width, f.height, f.frame_seed, f.max_bounce, f.flags = w, h, 7, 8, spk.FRAME_CACHE
    rays = {}
    for suffix in (0, 1, 3):
        print_case("cache train suffix", suffix)
        # every path a training path, so the suffix decides how far most paths run
        gpu_ctx.cache_create(spk.cache_config(log2_table=14, batch=1 << 14, seed=4, terminate_bounce=1,
                                              train_fraction=1.0, train_suffix=suffix))
        rays[suffix] = gpu_ctx.render_frame(f, rgba).rays_traced
        assert np.isfinite(rgba.cpu().numpy()).all()
        gpu_ctx.cache_destroy()
    print("rays by suffix", rays)
    assert rays[1] < rays[3] < rays[0]
