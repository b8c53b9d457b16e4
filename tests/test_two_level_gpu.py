"""Two-level scenes on the MI355X: a TLAS over instances and one object-space BLAS per mesh
(spk_build_options.two_level), and instance moves without BLAS rebuilds (spk_scene_update_instances).

Bars: hit ids and t bit-identical to the oracle's two-level watertight twin (object-space tests
through the float inverse transform) on every ray, for BVH widths 8 and 4 and with a sphere; ids equal
to the FP64 oracle on every triangle hit it does not flag (it has no spheres: a sphere hit must lie in
front of its triangle hit); after moving instances, hits bit-identical to a
fresh two-level upload of the moved scene and to the twin of the moved scene; path-traced images
bit-identical across launch schedules, and equal to the flattened scene's where the two agree on
the geometry (the first-hit albedo on nearly every pixel, colour statistically).
"""
import numpy as np
import pytest

import sparkle_amd as spk
from oracle import refcpu
from sparkle_amd import scenes
from tests.conftest import print_case
from tests.helpers import gpu_trace, instanced_scene, rotation
from tests.test_render_gpu import frame

pytestmark = pytest.mark.gpu


def _rays(sc, n, seed):
    lo, hi = scenes.scene_aabb(sc)
    return scenes.incoherent_rays(lo, hi, n=n, seed=seed)


def _check_twin(ctx, sc, rays, label):
    print_case("two-level parity", label, "rays", rays.shape[0])
    rc, h = gpu_trace(ctx, rays)
    assert rc == spk.SPK_OK, ctx.lib.spk_last_error(ctx.h)
    o = refcpu.OracleScene(sc)
    tw = o.trace(rays, refcpu.TRACE_WT_TWO)
    np.testing.assert_array_equal(h["prim"], tw["prim"])
    np.testing.assert_array_equal(h["t"].view(np.uint32), tw["t"].view(np.uint32))
    h64, flags, t64 = o.trace(rays, refcpu.TRACE_FP64)
    clean = flags == 0
    # the FP64 oracle covers triangles only: on rays that hit a sphere its triangle hit lies behind
    sph = h["prim"] != spk.SPK_MISS
    sph &= h["prim"] >= sc.num_triangles
    np.testing.assert_array_equal(h["prim"][clean & ~sph], h64["prim"][clean & ~sph])
    behind = (h64["prim"][sph] == spk.SPK_MISS) | (t64[sph] >= h["t"][sph] * (1 - 1e-6))
    assert behind.all()
    return h


@pytest.mark.parametrize("width", [8, 4])
def test_two_level_hits_match_the_twin(gpu_ctx, width):
    sc, _ = instanced_scene(seed=11, count=30)
    sc.add_sphere((2.0, 2.0, 2.0), 0.35, 0)
    gpu_ctx.upload_scene(sc, bvh_width=width, two_level=True)
    st = gpu_ctx.stats()
    assert st.num_triangles == sc.num_triangles and st.bvh_width == width
    rays = _rays(sc, 60_000, 12)
    h = _check_twin(gpu_ctx, sc, rays, f"width {width}")
    assert 0.1 < (h["prim"] != spk.SPK_MISS).mean() < 0.9
    assert (h["prim"] == sc.num_triangles).any()  # the sphere
    cfg = gpu_ctx.trace_config()
    assert cfg[0] == width and cfg[1] >= 24 and cfg[3] == 0


def test_two_level_autotune_keeps_results(gpu_ctx):
    sc, _ = instanced_scene(seed=13, count=20)
    gpu_ctx.upload_scene(sc, two_level=True)
    rays = _rays(sc, 20_000, 14)
    _, before = gpu_trace(gpu_ctx, rays)
    import torch

    from tests.helpers import to_dev

    print_case("two-level autotune")
    cfg, ms = gpu_ctx.autotune(to_dev(rays), rays.shape[0], tune_width=True)
    assert ms > 0 and cfg[3] == 0
    _, after = gpu_trace(gpu_ctx, rays)
    if cfg[0] == gpu_ctx.stats().bvh_width:
        np.testing.assert_array_equal(after["prim"], before["prim"])
        np.testing.assert_array_equal(after["t"].view(np.uint32), before["t"].view(np.uint32))
    torch.cuda.synchronize()


def _moved(xfs, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = xfs.copy()
    for k in range(0, xfs.shape[0] - 1, 3):  # every third instance (not the light)
        R = rotation(rng.normal(size=3), rng.uniform(0, 1.0)).astype(np.float64)
        M = xfs[k].astype(np.float64)
        M[:, :3] = R @ M[:, :3]
        M[:, 3] += rng.uniform(-0.5, 0.5, 3)
        out[k] = M.astype(np.float32)
    return out


def test_update_instances_equals_a_fresh_upload(gpu_ctx):
    sc, xfs = instanced_scene(seed=17, count=33)
    rays = _rays(sc, 40_000, 18)
    gpu_ctx.upload_scene(sc, two_level=True)
    for step, seed in enumerate((1, 2)):
        xfs = _moved(xfs, seed)
        print_case("update instances step", step)
        gpu_ctx.update_instances(xfs)
        for inst, M in zip(sc.instances, xfs):
            inst["M"] = M
        _, moved = gpu_trace(gpu_ctx, rays)
        h = _check_twin(gpu_ctx, sc, rays, f"after update {step}")
        np.testing.assert_array_equal(moved["prim"], h["prim"])
    fresh = spk.Context(0)
    try:
        fresh.upload_scene(sc, two_level=True)
        _, f = gpu_trace(fresh, rays)
    finally:
        fresh.close()
    np.testing.assert_array_equal(moved["prim"], f["prim"])
    np.testing.assert_array_equal(moved["t"].view(np.uint32), f["t"].view(np.uint32))


def test_update_instances_errors(gpu_ctx):
    sc, xfs = instanced_scene(seed=19, count=6)
    gpu_ctx.upload_scene(sc)
    with pytest.raises(spk.SpkError) as e:
        gpu_ctx.update_instances(xfs)
    assert e.value.code == spk.SPK_ERR_UNSUPPORTED
    gpu_ctx.upload_scene(sc, two_level=True)
    with pytest.raises(spk.SpkError) as e:
        gpu_ctx.update_instances(xfs[:-1])
    assert e.value.code == spk.SPK_ERR_INVALID_ARGUMENT
    bad = xfs.copy()
    bad[2, :, :3] = 0.0
    with pytest.raises(spk.SpkError) as e:
        gpu_ctx.update_instances(bad)
    assert e.value.code == spk.SPK_ERR_INVALID_SCENE
    bad = xfs.copy()
    bad[1, 0, 3] = np.nan
    with pytest.raises(spk.SpkError) as e:
        gpu_ctx.update_instances(bad)
    assert e.value.code == spk.SPK_ERR_INVALID_SCENE
    # the scene is unchanged after rejected updates
    rays = _rays(sc, 5000, 20)
    _check_twin(gpu_ctx, sc, rays, "after rejected updates")


def _camera(w, h):
    return spk.camera_desc(position=(2.0, -3.5, 2.5), front=(0.0, 1.0, -0.35), vertical_fov=np.radians(60.0),
                           aspect=w / h, focus_distance=4.0)


def test_two_level_path_tracing(gpu_ctx):
    import torch

    sc, _ = instanced_scene(seed=23, count=30)
    w, h = 64, 48
    rgba = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
    imgs = {}
    for two in (False, True):
        gpu_ctx.upload_scene(sc, two_level=two)
        gpu_ctx.set_camera(_camera(w, h))
        print_case("two-level render", two)
        gpu_ctx.render_frame(frame(w, h, 0, bounce=1, debug=spk.DEBUG_ALBEDO), rgba)
        imgs[two, "albedo"] = rgba.cpu().numpy().copy()
        acc = np.zeros(w * h * 4)
        for s in range(64):
            gpu_ctx.render_frame(frame(w, h, s, bounce=6, flags=spk.FRAME_NEE, sky=(0.2, 0.3, 0.5)), rgba)
            acc += rgba.cpu().numpy()
        imgs[two, "mean"] = acc / 64
        if two:
            # the image does not depend on the launch schedule (per-path tail, sort, grid)
            f = frame(w, h, 7, bounce=6, flags=spk.FRAME_NEE, sky=(0.2, 0.3, 0.5))
            gpu_ctx.set_schedule(8, 1, 16, 0)
            gpu_ctx.render_frame(f, rgba)
            ref = rgba.cpu().numpy().copy()
            for sched in ((8, 0, 64, 1), (4, 1, 16, 3)):
                gpu_ctx.set_schedule(*sched)
                gpu_ctx.render_frame(f, rgba)
                np.testing.assert_array_equal(rgba.cpu().numpy(), ref)
            gpu_ctx.set_schedule(8, 1, 16)
    a, b = imgs[False, "albedo"], imgs[True, "albedo"]
    same = np.all(a.reshape(-1, 4) == b.reshape(-1, 4), axis=1)
    assert same.mean() > 0.995, same.mean()
    ma, mb = imgs[False, "mean"].reshape(-1, 4)[:, :3], imgs[True, "mean"].reshape(-1, 4)[:, :3]
    assert np.isfinite(mb).all() and mb.mean() > 0
    # 64-frame means of the two scenes: same expectation, nearly the same samples
    assert abs(ma.mean() - mb.mean()) / ma.mean() < 0.01
    assert np.median(np.abs(ma - mb)) < 1e-3


def test_update_instances_back_and_forth(gpu_ctx):
    """Moving instances, a rejected update, then moving them back to where they were: the hits are
    those of the original placement (the host copy the later updates compare against stays true to
    what the device holds)."""
    sc, xfs = instanced_scene(seed=23, count=12)
    rays = _rays(sc, 8000, 24)
    gpu_ctx.upload_scene(sc, two_level=True)
    moved = _moved(xfs, 7)
    print_case("update instances away")
    gpu_ctx.update_instances(moved)
    bad = moved.copy()
    bad[0, :, :3] = 0.0
    with pytest.raises(spk.SpkError):
        gpu_ctx.update_instances(bad)
    print_case("update instances back")
    gpu_ctx.update_instances(xfs)
    _check_twin(gpu_ctx, sc, rays, "moved away and back")
