"""Library work on a caller's stream (spk_set_stream), as a host that runs torch on a stream of its
own would set it up (MI355X): results equal those on the context's own stream, and work the caller
queued on its stream before a call is ordered before the call's kernels."""
import numpy as np
import pytest

import sparkle_amd as spk
from sparkle_amd import scenes
from tests.conftest import print_case
from tests.helpers import gpu_trace
from tests.test_render_gpu import Renderer, frame

pytestmark = pytest.mark.gpu


def test_closest_hit_on_a_caller_stream(gpu_ctx):
    """Closest hit through device arrays and through host arrays (the pinned staging pipeline) on a
    caller's torch stream gives the hits of the context's own stream, in input and coherence order.
    The hit array is filled with 0xFF on the caller's stream right before the call, so a call that
    ran ahead of the caller's queued work would leave that fill in place."""
    import torch

    sc = scenes.heightfield_scene(cells=128)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=600_001, seed=43)
    gpu_ctx.upload_scene(sc)
    w, stack, block = gpu_ctx.trace_config()[:3]
    stream = torch.cuda.Stream()
    d_rays = torch.from_numpy(rays.view(np.uint8).reshape(-1)).cuda()
    d_hits = torch.empty(rays.shape[0] * 8, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    try:
        for ray_sort in (0, 1):
            gpu_ctx.set_stream(None)
            gpu_ctx.set_trace_config(w, stack, block, 0, ray_sort)
            print_case("own stream", ray_sort)
            rc0, h0 = gpu_trace(gpu_ctx, rays)
            assert rc0 == spk.SPK_OK
            gpu_ctx.set_stream(stream.cuda_stream)
            print_case("caller stream, device arrays", ray_sort)
            with torch.cuda.stream(stream):
                d_hits.fill_(0xFF)
            assert gpu_ctx.trace(d_rays, d_hits, rays.shape[0]) == spk.SPK_OK
            h1 = d_hits.cpu().numpy().view(spk.HIT_DTYPE)
            print_case("caller stream, host arrays", ray_sort)
            rc2, h2 = gpu_ctx.trace_host(rays)
            assert rc2 == spk.SPK_OK
            for h in (h1, h2):
                np.testing.assert_array_equal(h["prim"], h0["prim"])
                np.testing.assert_array_equal(h["t"].view(np.uint32), h0["t"].view(np.uint32))
    finally:
        gpu_ctx.set_stream(None)
        gpu_ctx.set_trace_config(w, stack, block, 0, 0)
    torch.cuda.synchronize()


def test_frame_on_a_caller_stream(gpu_ctx):
    """spk_render_frame on a caller's stream (the frame graph captured there, then replayed) renders
    the image of the context's own stream: the reference's estimator, NEE, and the learned cache's
    first frame after a reset (its training branch forks from and joins the caller's stream)."""
    import torch

    w, h = 48, 40
    sc = scenes.textured_cornell_scene(seed=2)
    r = Renderer(gpu_ctx, sc, scenes.cornell_camera(w, h), w, h)
    stream = torch.cuda.Stream()
    try:
        for flags in (0, spk.FRAME_NEE):
            f = frame(w, h, 3, bounce=8, flags=flags)
            gpu_ctx.set_stream(None)
            gpu_ctx.render_frame(f, r.rgba)
            ref = r.rgba.cpu().numpy().copy()
            gpu_ctx.set_stream(stream.cuda_stream)
            for k in range(2):  # the capture on the caller's stream, then a replay
                print_case("caller stream frame", flags, k)
                with torch.cuda.stream(stream):
                    r.rgba.fill_(-1.0)
                gpu_ctx.render_frame(f, r.rgba)
                np.testing.assert_array_equal(r.rgba.cpu().numpy(), ref)
        gpu_ctx.cache_create(spk.cache_config(terminate_bounce=2))
        f = frame(w, h, 5, bounce=8, flags=spk.FRAME_CACHE | spk.FRAME_NEE)
        gpu_ctx.set_stream(None)
        gpu_ctx.render_frame(f, r.rgba)  # capture
        gpu_ctx.cache_reset()
        gpu_ctx.render_frame(f, r.rgba)
        ref = r.rgba.cpu().numpy().copy()
        assert np.isfinite(ref).all()
        gpu_ctx.set_stream(stream.cuda_stream)
        for k in range(2):
            print_case("caller stream cache frame", k)
            gpu_ctx.cache_reset()
            gpu_ctx.render_frame(f, r.rgba)
            np.testing.assert_array_equal(r.rgba.cpu().numpy(), ref)
        gpu_ctx.set_stream(None)
        gpu_ctx.cache_destroy()
    finally:
        gpu_ctx.set_stream(None)
    torch.cuda.synchronize()
