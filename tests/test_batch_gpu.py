"""Batches of independent scenes of uneven size in one launch (MI355X).

Every scene of a batch must render bit-identically to the same scene rendered alone with
spk_render_frame (same per-pixel seeds, same camera, same closest hits), whatever the schedule
(persistent grid size, material sort on or off).
"""
import numpy as np
import pytest

import sparkle_amd as spk
from sparkle_amd import scenes
from tests.conftest import print_case
from tests.test_render_gpu import frame

pytestmark = pytest.mark.gpu


def small_batch():
    items = scenes.batch_scenes(seed=3, count=5, min_tris=1_000, max_tris=30_000, sizes=(16, 24, 32))
    items.append((scenes.cornell_scene(0), scenes.cornell_camera(40, 24), 40, 24))
    return items


def render_alone(ctx, items, seeds, **kw):
    import torch

    out = {}
    for k, (sc, cam, w, h) in enumerate(items):
        ctx.upload_scene(sc)
        ctx.set_camera(cam)
        rgba = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
        for s in seeds:
            print_case("alone", k, "seed", s, kw)
            ctx.render_frame(frame(w, h, s, **kw), rgba)
            out[k, s] = rgba.cpu().numpy().copy()
    return out


def test_batch_matches_each_scene_rendered_alone():
    import torch

    items = small_batch()
    seeds = (0, 5)
    ctx = spk.Context(0)
    try:
        offs = ctx.upload_batch(items)
        assert offs[0] == 0 and offs[-1] == sum(w * h for _, _, w, h in items)
        rgba = torch.zeros(offs[-1] * 4, dtype=torch.float32, device="cuda")
        batch = {}
        rays = 0
        for s in seeds:
            print_case("batch seed", s)
            st = ctx.render_batch(frame(0, 0, s, bounce=6, sky=(0.3, 0.5, 0.9)), rgba)
            rays += st.rays_traced
            img = rgba.cpu().numpy()
            for k in range(len(items)):
                batch[k, s] = img[offs[k] * 4:offs[k + 1] * 4].copy()
        assert rays >= 2 * offs[-1]
        alone_ctx = spk.Context(0)
        try:
            alone = render_alone(alone_ctx, items, seeds, bounce=6, sky=(0.3, 0.5, 0.9))
        finally:
            alone_ctx.close()
        for key, ref in alone.items():
            assert np.array_equal(batch[key], ref), key
            assert np.isfinite(ref).all() and ref[3::4].min() != 0.0
    finally:
        ctx.close()


def test_batch_depth_mode_uses_each_scenes_camera():
    import torch

    items = small_batch()[-3:]
    ctx = spk.Context(0)
    try:
        offs = ctx.upload_batch(items)
        rgba = torch.zeros(offs[-1] * 4, dtype=torch.float32, device="cuda")
        print_case("batch depth mode")
        ctx.render_batch(frame(0, 0, 2, bounce=2, debug=spk.DEBUG_DEPTH), rgba)
        img = rgba.cpu().numpy()
        alone = render_alone(ctx, items, (2,), bounce=2, debug=spk.DEBUG_DEPTH)
        for k in range(len(items)):
            assert np.array_equal(img[offs[k] * 4:offs[k + 1] * 4], alone[k, 2]), k
    finally:
        ctx.close()


def test_batch_schedule_does_not_change_results():
    import torch

    items = small_batch()
    ctx = spk.Context(0)
    try:
        offs = ctx.upload_batch(items)
        rgba = torch.zeros(offs[-1] * 4, dtype=torch.float32, device="cuda")
        f = frame(0, 0, 1, bounce=6)
        print_case("batch default schedule")
        ctx.render_batch(f, rgba)
        ref = rgba.cpu().numpy().copy()
        print_case("batch autotune")
        (blocks, srt, refill, tail), ms = ctx.autotune_batch(f, rgba)
        assert blocks in (2, 4, 8, 16) and srt in (0, 1) and refill in (0, 16, 64) and tail in (0, 1, 2, 3, 4, 5) and ms > 0
        assert ctx.schedule(batch=True) == (blocks, srt, refill, tail)
        ctx.render_batch(f, rgba)
        assert np.array_equal(rgba.cpu().numpy(), ref)
        for tail in (1, 3):
            print_case("batch per-path tail from", tail)
            ctx.set_schedule(8, 1, 16, tail, batch=True)
            ctx.render_batch(f, rgba)
            assert np.array_equal(rgba.cpu().numpy(), ref)
        print_case("batch direct extends")
        ctx.set_schedule(4, 1, 0, 0, batch=True)
        ctx.render_batch(f, rgba)
        assert np.array_equal(rgba.cpu().numpy(), ref)
        print_case("batch unsorted")
        ctx.render_batch(frame(0, 0, 1, bounce=6, flags=spk.FRAME_NO_SORT), rgba)
        assert np.array_equal(rgba.cpu().numpy(), ref)
    finally:
        ctx.close()


def test_batch_rejects_single_scene_features_and_bad_scenes():
    import torch

    items = small_batch()[:2]
    ctx = spk.Context(0)
    try:
        with pytest.raises(spk.SpkError) as e:
            ctx.render_batch(frame(0, 0, 0), torch.zeros(4, device="cuda"))
        assert e.value.code == spk.SPK_ERR_NO_SCENE
        offs = ctx.upload_batch(items)
        rgba = torch.zeros(offs[-1] * 4, dtype=torch.float32, device="cuda")
        with pytest.raises(spk.SpkError) as e:
            ctx.render_batch(frame(0, 0, 0, flags=spk.FRAME_NEE), rgba)
        assert e.value.code == spk.SPK_ERR_UNSUPPORTED
        bad = scenes.cornell_scene(0)
        bad.instances[0]["material"] = 999
        with pytest.raises(spk.SpkError) as e:
            ctx.upload_batch(items + [(bad, scenes.cornell_camera(8, 8), 8, 8)])
        assert e.value.code == spk.SPK_ERR_INVALID_SCENE and "scene 2" in str(e.value)
        with pytest.raises(spk.SpkError) as e:
            ctx.upload_batch([(items[0][0], items[0][1], 1, 8)])
        assert e.value.code == spk.SPK_ERR_INVALID_ARGUMENT
    finally:
        ctx.close()
