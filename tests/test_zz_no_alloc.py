"""spk_trace_closest_noalloc (the closest hit with the library's fixed scratch, no allocation during
the call) against spk_trace_closest on the same rays (MI355X).

Two calls in a row, the second with fewer rays, in the coherence-sorted configuration (the one whose
kernels use the scratch) and in input order; hits bit-identical to the allocating entry point's,
which the other GPU tests check against the oracle. The hit arrays are filled with a sentinel
first, so a ray the sort dropped would show. Calls above the maximum are rejected on the host.
"""
import numpy as np
import pytest

import sparkle_amd as spk
from sparkle_amd import scenes
from tests.conftest import print_case
from tests.helpers import empty_dev, hits_from_dev, to_dev

pytestmark = pytest.mark.gpu


def _trace(ctx, fn, rays):
    import torch

    n = rays.shape[0]
    d_rays = to_dev(rays)
    d_hits = empty_dev(n * 8)
    d_hits.fill_(0xAB)
    torch.cuda.synchronize()
    rc = fn(d_rays, d_hits, n)
    return rc, hits_from_dev(d_hits, n)


def test_noalloc_equals_allocating_entry_point(gpu_ctx):
    sc = scenes.heightfield_scene(cells=256)
    lo, hi = scenes.scene_aabb(sc)
    big = scenes.incoherent_rays(lo, hi, n=spk.NOALLOC_MAX_RAYS - 5, seed=31)
    small = scenes.incoherent_rays(lo, hi, n=100_003, seed=32)
    small["dir"][17] = 0.0  # one invalid ray: reported the same way by both entry points
    gpu_ctx.upload_scene(sc)
    for cfg in ((8, 12, 64, 0, 1), (8, 12, 64, 0, 0)):
        gpu_ctx.set_trace_config(*cfg)
        for name, rays in (("big", big), ("small", small)):
            print_case("noalloc", cfg, name, rays.shape[0])
            rc_a, h_a = _trace(gpu_ctx, gpu_ctx.trace, rays)
            rc_b, h_b = _trace(gpu_ctx, gpu_ctx.trace_noalloc, rays)
            assert rc_a == rc_b == (spk.SPK_ERR_INVALID_RAY if name == "small" else spk.SPK_OK)
            assert not (h_b["prim"] == 0xABABABAB).any()
            np.testing.assert_array_equal(h_b["prim"], h_a["prim"])
            np.testing.assert_array_equal(h_b["t"].view(np.uint32), h_a["t"].view(np.uint32))
    # the sorted configuration once more after the unsorted calls: the scratch keeps the last call's
    # contents, which must not leak into this one
    gpu_ctx.set_trace_config(8, 12, 64, 0, 1)
    print_case("noalloc sorted after unsorted")
    rc_a, h_a = _trace(gpu_ctx, gpu_ctx.trace, small[:5000])
    rc_b, h_b = _trace(gpu_ctx, gpu_ctx.trace_noalloc, small[:5000])
    assert rc_a == rc_b == spk.SPK_ERR_INVALID_RAY
    np.testing.assert_array_equal(h_b.view(np.uint8), h_a.view(np.uint8))
    assert gpu_ctx.trace_noalloc(to_dev(small[:1]), empty_dev(8), 0) == spk.SPK_OK
    # above the maximum: rejected before any device work
    assert gpu_ctx.trace_noalloc(to_dev(small[:1]), empty_dev(8), spk.NOALLOC_MAX_RAYS + 1) == \
        spk.SPK_ERR_INVALID_ARGUMENT
    assert "at most" in gpu_ctx.lib.spk_last_error(gpu_ctx.h).decode()
