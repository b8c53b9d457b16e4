"""Closest-hit parity of the HIP traversal against the CPU oracle (runs on the MI355X).

Bar: hit ids and t bit-identical to the oracle's watertight FP32 twin on every ray; against the
FP64 oracle, ids equal on every ray it does not flag (edge within 1e-4 barycentric, t-tie within
1e-6, grazing below 1e-3) and t within 1e-6 relative.
"""
import numpy as np
import pytest

import sparkle_amd as spk
from sparkle_amd import scenes
from oracle import refcpu
from tests.conftest import print_case
from tests.helpers import empty_dev, gpu_trace, hits_from_dev, multi_mesh_scene, soup_scene, to_dev

pytestmark = pytest.mark.gpu

T_REL_TOL = 1e-6


def check_against_oracle(ctx, sc, rays, width=8, label=""):
    print_case("trace parity", label, "width", width, "rays", rays.shape[0])
    ctx.upload_scene(sc, bvh_width=width)
    rc, h = gpu_trace(ctx, rays)
    assert rc == spk.SPK_OK, ctx.lib.spk_last_error(ctx.h)
    o = refcpu.OracleScene(sc)
    wt = o.trace(rays, refcpu.TRACE_WT)
    np.testing.assert_array_equal(h["prim"], wt["prim"])
    np.testing.assert_array_equal(h["t"].view(np.uint32), wt["t"].view(np.uint32))
    h64, flags, t64 = o.trace(rays, refcpu.TRACE_FP64)
    clean = flags == 0
    np.testing.assert_array_equal(h["prim"][clean], h64["prim"][clean])
    hit = clean & (h64["prim"] != spk.SPK_MISS)
    rel = np.abs(h["t"][hit].astype(np.float64) - t64[hit]) / t64[hit]
    assert rel.size == 0 or rel.max() <= T_REL_TOL, rel.max()
    return h, flags


@pytest.mark.parametrize("width", [8, 4])
def test_heightfield_incoherent(gpu_ctx, width):
    sc = scenes.heightfield_scene(cells=128)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=200_000, seed=3)
    h, _ = check_against_oracle(gpu_ctx, sc, rays, width, "heightfield128 incoherent")
    assert 0.3 < (h["prim"] != spk.SPK_MISS).mean() < 0.7


def test_heightfield_coherent(gpu_ctx):
    sc = scenes.heightfield_scene(cells=256)
    rays = scenes.coherent_rays(320, 180)
    h, _ = check_against_oracle(gpu_ctx, sc, rays, 8, "heightfield256 coherent")
    assert (h["prim"] != spk.SPK_MISS).mean() > 0.1


@pytest.mark.parametrize("seed", [0, 1])
def test_triangle_soup(gpu_ctx, seed):
    sc = soup_scene(5000, seed=seed)
    rng = np.random.Generator(np.random.PCG64(seed + 10))
    n = 100_000
    rays = np.zeros(n, dtype=spk.RAY_DTYPE)
    rays["org"] = rng.random((n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    rays["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    check_against_oracle(gpu_ctx, sc, rays, 8, f"soup seed {seed}")


def test_multi_mesh_ids(gpu_ctx):
    sc = multi_mesh_scene()
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=50_000, seed=5)
    h, _ = check_against_oracle(gpu_ctx, sc, rays, 8, "multi mesh")
    assert h["prim"][h["prim"] != spk.SPK_MISS].max() < sc.num_triangles


def test_shared_edge_ties_go_to_smaller_id(gpu_ctx):
    # rays through the shared diagonal of a unit quad: both triangles report the same t
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], dtype=np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)
    sc = spk.Scene()
    sc.add_instance(sc.add_mesh(pos, idx), sc.add_material())
    s = np.linspace(0.05, 0.95, 64, dtype=np.float32)
    rays = np.zeros(64, dtype=spk.RAY_DTYPE)
    rays["org"] = np.stack([s, s, np.ones_like(s)], 1)
    rays["dir"] = np.array([0, 0, -1], dtype=np.float32)
    rays["tmax"] = np.inf
    gpu_ctx.upload_scene(sc)
    rc, h = gpu_trace(gpu_ctx, rays)
    assert rc == spk.SPK_OK
    assert (h["prim"] == 0).all() and (h["t"] == 1.0).all()


def test_tmin_tmax_and_empty(gpu_ctx):
    sc = scenes.heightfield_scene(cells=64)
    gpu_ctx.upload_scene(sc)
    rays = np.zeros(3, dtype=spk.RAY_DTYPE)
    rays["org"] = [[0.5, 0.5, 1.0]] * 3
    rays["dir"] = [[0, 0, -1.0]] * 3
    rays["tmin"] = [0.0, 0.0, 2.0]
    rays["tmax"] = [np.inf, 0.5, np.inf]
    rc, h = gpu_trace(gpu_ctx, rays)
    assert rc == spk.SPK_OK
    assert h["prim"][0] != spk.SPK_MISS and 0.5 < h["t"][0] < 1.0
    assert h["prim"][1] == spk.SPK_MISS and h["prim"][2] == spk.SPK_MISS
    # zero rays is a no-op
    assert gpu_ctx.trace(empty_dev(32), empty_dev(8), 0) == spk.SPK_OK


def test_invalid_rays_reported_as_misses(gpu_ctx):
    sc = scenes.heightfield_scene(cells=64)
    gpu_ctx.upload_scene(sc)
    rays = np.zeros(4, dtype=spk.RAY_DTYPE)
    rays["org"] = [[0.5, 0.5, 1.0], [np.nan, 0.5, 1.0], [0.5, 0.5, 1.0], [0.2, 0.3, 1.0]]
    rays["dir"] = [[0, 0, -1.0], [0, 0, -1.0], [0, 0, 0], [0, 0, -1.0]]
    rays["tmax"] = np.inf
    rc, h = gpu_trace(gpu_ctx, rays)
    assert rc == spk.SPK_ERR_INVALID_RAY
    assert h["prim"][1] == spk.SPK_MISS and h["prim"][2] == spk.SPK_MISS
    assert h["prim"][0] != spk.SPK_MISS and h["prim"][3] != spk.SPK_MISS
    # the error is reported once; the next call is clean
    rc, _ = gpu_trace(gpu_ctx, rays[[0, 3]])
    assert rc == spk.SPK_OK


def test_axis_parallel_rays(gpu_ctx):
    # d.x == 0 rays: the reference's InverseAt gives NaN and drops them; the device test does not
    sc = scenes.heightfield_scene(cells=64)
    n = 4096
    rng = np.random.Generator(np.random.PCG64(9))
    rays = np.zeros(n, dtype=spk.RAY_DTYPE)
    rays["org"] = np.concatenate([rng.random((n, 2)), np.full((n, 1), 0.5)], 1).astype(np.float32)
    d = np.zeros((n, 3))
    d[:, 1] = rng.normal(size=n)
    d[:, 2] = -np.abs(rng.normal(size=n))
    rays["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmax"] = np.inf
    h, _ = check_against_oracle(gpu_ctx, sc, rays, 8, "axis-parallel")
    assert (h["prim"] != spk.SPK_MISS).mean() > 0.3


def test_all_kernel_variants_agree(gpu_ctx):
    """Every compiled (width, stack, block) variant gives identical results."""
    sc = scenes.heightfield_scene(cells=128)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=50_000, seed=11)
    ref = None
    seen = 0
    for width in (8, 4):
        gpu_ctx.upload_scene(sc, bvh_width=width)
        for cfg in gpu_ctx.list_trace_configs():
            print_case("variant", cfg)
            gpu_ctx.set_trace_config(*cfg)
            assert gpu_ctx.trace_config() == cfg
            rc, h = gpu_trace(gpu_ctx, rays)
            assert rc == spk.SPK_OK
            if ref is None:
                ref = h
            np.testing.assert_array_equal(h["prim"], ref["prim"])
            np.testing.assert_array_equal(h["t"].view(np.uint32), ref["t"].view(np.uint32))
            seen += 1
    assert seen >= 10


def test_persistent_kernels_match_counts_and_invalid_rays(gpu_ctx):
    """Persistent lanes with refill give the same hits and traversal counts as one ray per thread,
    report invalid rays the same way, and handle a ray count that is not a multiple of the wave."""
    import torch

    sc = scenes.heightfield_scene(cells=128)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=200_003, seed=21)
    rays["org"][[5, 70_000]] = np.nan
    rays["dir"][199_999] = 0.0
    n = rays.shape[0]
    for width in (4, 8):
        gpu_ctx.upload_scene(sc, bvh_width=width)
        gpu_ctx.set_trace_config(width, 16, 64, 0)
        rc, ref = gpu_trace(gpu_ctx, rays)
        assert rc == spk.SPK_ERR_INVALID_RAY
        d = to_dev(rays)
        cn = torch.zeros(n, dtype=torch.int32, device="cuda")
        ct = torch.zeros(n, dtype=torch.int32, device="cuda")
        rc0 = gpu_ctx.lib.spk_trace_count(gpu_ctx.h, d.data_ptr(), n, cn.data_ptr(), ct.data_ptr())
        assert rc0 == spk.SPK_ERR_INVALID_RAY
        ref_counts = (cn.cpu().numpy().copy(), ct.cpu().numpy().copy())
        for refill in (1, 16, 32):
            print_case("persistent", width, refill)
            gpu_ctx.set_trace_config(width, 16, 64, refill)
            assert gpu_ctx.trace_config() == (width, 16, 64, refill, 0)
            rc, h = gpu_trace(gpu_ctx, rays)
            assert rc == spk.SPK_ERR_INVALID_RAY
            np.testing.assert_array_equal(h["prim"], ref["prim"])
            np.testing.assert_array_equal(h["t"].view(np.uint32), ref["t"].view(np.uint32))
            cn.fill_(-1)
            ct.fill_(-1)
            rc1 = gpu_ctx.lib.spk_trace_count(gpu_ctx.h, d.data_ptr(), n, cn.data_ptr(), ct.data_ptr())
            assert rc1 == spk.SPK_ERR_INVALID_RAY
            np.testing.assert_array_equal(cn.cpu().numpy(), ref_counts[0])
            np.testing.assert_array_equal(ct.cpu().numpy(), ref_counts[1])
        # coherence order: seven sorting chunks, the last one ragged, invalid rays in bin 0
        print_case("ray sort", width)
        gpu_ctx.set_trace_config(width, 16, 64, 0, 1)
        assert gpu_ctx.trace_config() == (width, 16, 64, 0, 1)
        rc, h = gpu_trace(gpu_ctx, rays)
        assert rc == spk.SPK_ERR_INVALID_RAY
        np.testing.assert_array_equal(h["prim"], ref["prim"])
        np.testing.assert_array_equal(h["t"].view(np.uint32), ref["t"].view(np.uint32))
        cn.fill_(-1)
        ct.fill_(-1)
        rc1 = gpu_ctx.lib.spk_trace_count(gpu_ctx.h, d.data_ptr(), n, cn.data_ptr(), ct.data_ptr())
        assert rc1 == spk.SPK_ERR_INVALID_RAY
        np.testing.assert_array_equal(cn.cpu().numpy(), ref_counts[0])
        np.testing.assert_array_equal(ct.cpu().numpy(), ref_counts[1])
        gpu_ctx.set_trace_config(width, 16, 64, 0, 0)
        rc, _ = gpu_trace(gpu_ctx, rays[100:1100])  # no invalid ray in this slice
        assert rc == spk.SPK_OK


def test_autotune_keeps_results(gpu_ctx):
    sc = scenes.heightfield_scene(cells=256)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=100_000, seed=12)
    gpu_ctx.upload_scene(sc)
    rc, before = gpu_trace(gpu_ctx, rays)
    d = to_dev(rays)
    print_case("autotune")
    cfg, ms = gpu_ctx.autotune(d, tune_width=True)
    assert ms > 0 and cfg[0] in (4, 8)
    assert gpu_ctx.trace_config() == cfg
    rc, after = gpu_trace(gpu_ctx, rays)
    np.testing.assert_array_equal(after["prim"], before["prim"])
    np.testing.assert_array_equal(after["t"].view(np.uint32), before["t"].view(np.uint32))


def test_full_heightfield_2m_subsampled(gpu_ctx):
    """1024^2-cell heightfield (2,097,152 triangles), 8,388,608 incoherent rays traced on the GPU in
    the post-upload configuration, in the configuration the benchmark runs (BVH8, stack 12, 64
    threads, coherence sort: 256 pass-2 sorting chunks with a permuted write-back) and in the one the
    autotuner picks on this batch; the hit array is filled with a sentinel before each call, so a ray
    the sort dropped would show. Every 64th ray is checked bit for bit against the oracle's
    watertight twin and against FP64 on unflagged rays, every ray for range properties, and all
    8,388,608 hits of the three configurations against each other."""
    import torch

    sc = scenes.heightfield_scene(cells=1024)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=8 * 1024 * 1024, seed=0)
    n = rays.shape[0]
    print_case("full heightfield upload")
    st = gpu_ctx.upload_scene(sc)
    assert st.num_triangles == 2 * 1024 * 1024
    sub = np.arange(0, n, 64)
    o = refcpu.OracleScene(sc)
    wt = o.trace(rays[sub], refcpu.TRACE_WT)
    h64, flags, t64 = o.trace(rays[sub], refcpu.TRACE_FP64)
    clean = flags == 0
    m = clean & (h64["prim"] != spk.SPK_MISS)
    d_rays = to_dev(rays)
    d_hits = empty_dev(n * 8)
    default_cfg = gpu_ctx.trace_config()
    print_case("full heightfield autotune 8M")
    tuned_cfg, _ = gpu_ctx.autotune(d_rays, tune_width=False)
    first = None
    for cfg in (default_cfg, (8, 12, 64, 0, 1), tuned_cfg):
        print_case("full heightfield trace 8M", cfg)
        gpu_ctx.set_trace_config(*cfg)
        d_hits.fill_(0xAB)
        torch.cuda.synchronize()
        assert gpu_ctx.trace(d_rays, d_hits, n) == spk.SPK_OK
        h = hits_from_dev(d_hits, n)
        assert not (h["prim"].view(np.uint32) == 0xABABABAB).any()
        hit = h["prim"] != spk.SPK_MISS
        assert (h["prim"][hit] < 2 * 1024 * 1024).all()
        assert (h["t"][hit] > 0).all() and np.isfinite(h["t"][hit]).all()
        assert np.isinf(h["t"][~hit]).all()
        np.testing.assert_array_equal(h["prim"][sub], wt["prim"])
        np.testing.assert_array_equal(h["t"][sub].view(np.uint32), wt["t"].view(np.uint32))
        np.testing.assert_array_equal(h["prim"][sub][clean], h64["prim"][clean])
        rel = np.abs(h["t"][sub][m].astype(np.float64) - t64[m]) / t64[m]
        assert rel.max() <= T_REL_TOL
        if first is None:
            first = h
        else:
            np.testing.assert_array_equal(h["prim"], first["prim"])
            np.testing.assert_array_equal(h["t"].view(np.uint32), first["t"].view(np.uint32))
    gpu_ctx.set_trace_config(*default_cfg)
    # the coherent 1080p set on the same scene, every 16th ray
    rays_a = scenes.coherent_rays()
    print_case("full heightfield trace coherent 1080p")
    rc, ha = gpu_trace(gpu_ctx, rays_a)
    assert rc == spk.SPK_OK
    sub = np.arange(0, rays_a.shape[0], 16)
    wa = o.trace(rays_a[sub], refcpu.TRACE_WT)
    np.testing.assert_array_equal(ha["prim"][sub], wa["prim"])
    np.testing.assert_array_equal(ha["t"][sub].view(np.uint32), wa["t"].view(np.uint32))


def test_analytic_spheres_match_oracle(gpu_ctx):
    """Spheres beside a triangle soup: ids (num_tris + k) and t bit-identical to the oracle's
    watertight twin for rays from outside and inside the spheres, for every traversal form."""
    sc = soup_scene(3000, seed=4)
    mat = 0
    rng = np.random.Generator(np.random.PCG64(17))
    for k in range(12):
        sc.add_sphere(tuple(rng.random(3)), float(rng.uniform(0.02, 0.2)), mat)
    o = refcpu.OracleScene(sc)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(np.minimum(lo, 0.0), np.maximum(hi, 1.0), n=60_000, seed=5)
    want = o.trace(rays, refcpu.TRACE_WT)
    assert (want["prim"] >= 3000).sum() > 1000 and (want["prim"] < 3000).sum() > 1000
    gpu_ctx.upload_scene(sc, bvh_width=4)
    for cfg in ((4, 16, 64, 0), (4, 16, 64, 16)):
        print_case("spheres", cfg)
        gpu_ctx.set_trace_config(*cfg)
        rc, h = gpu_trace(gpu_ctx, rays)
        assert rc == spk.SPK_OK
        np.testing.assert_array_equal(h["prim"], want["prim"])
        np.testing.assert_array_equal(h["t"].view(np.uint32), want["t"].view(np.uint32))


def test_kernel_timing_records_each_call(gpu_ctx):
    """spk_trace_kernel_timing times the traversal kernel of each of the next `capacity` calls, with
    and without the coherence sort before it; later calls and capacity 0 record nothing."""
    import torch

    sc = scenes.heightfield_scene(cells=128)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=100_000, seed=31)
    gpu_ctx.upload_scene(sc)
    d = to_dev(rays)
    hits = torch.empty(rays.shape[0] * 8, dtype=torch.uint8, device="cuda")
    w, stack, block = gpu_ctx.trace_config()[:3]
    for sort in (0, 1):
        print_case("kernel timing", sort)
        gpu_ctx.set_trace_config(w, stack, block, 0, sort)
        gpu_ctx.trace_kernel_timing(3)
        for _ in range(5):
            assert gpu_ctx.trace(d, hits, sync=False) == spk.SPK_OK
        t = gpu_ctx.trace_kernel_times()
        assert len(t) == 3 and all(0.0 < x < 1000.0 for x in t), t
    gpu_ctx.trace_kernel_timing(0)
    assert gpu_ctx.trace(d, hits) == spk.SPK_OK
    assert gpu_ctx.trace_kernel_times() == []
    with pytest.raises(spk.SpkError) as e:
        gpu_ctx.trace_kernel_timing(1 << 17)  # at most 65536 timed calls
    assert e.value.code == spk.SPK_ERR_INVALID_ARGUMENT
    gpu_ctx.set_trace_config(w, stack, block, 0, 0)


def test_host_buffer_trace_matches_device_buffers(gpu_ctx):
    """spk_trace_closest_host (rays and hits in host memory, staged through a ring of three pinned
    buffers in chunks of 524,288 rays at these sizes) gives the device-buffer call's hits bit for bit,
    in input and coherence order, reports invalid rays the same way (one in the last, ragged chunk),
    and treats zero rays as a no-op. 2,100,003 rays make five chunks, so ring slots are reused; 600,001
    make two; 70,001 one partial chunk. A caller's hit array is written in place."""
    sc = scenes.heightfield_scene(cells=128)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=2_100_003, seed=41)
    rays["dir"][2_099_000] = 0.0
    gpu_ctx.upload_scene(sc)
    w, stack, block = gpu_ctx.trace_config()[:3]
    for ray_sort in (0, 1):
        gpu_ctx.set_trace_config(w, stack, block, 0, ray_sort)
        for n in (70_001, 2_100_003, 600_001):  # the first call sets the ring up, the second grows it
            print_case("host buffers", ray_sort, n)
            sub = rays[:n]
            rc_d, h_d = gpu_trace(gpu_ctx, sub)
            out = np.full(n, 0xFF, dtype=np.uint8).repeat(8).view(spk.HIT_DTYPE) if n == 600_001 else None
            rc_h, h_h = gpu_ctx.trace_host(sub, out)
            if out is not None:
                assert h_h is out
            bad = spk.SPK_ERR_INVALID_RAY if n > 2_099_000 else spk.SPK_OK
            assert rc_d == rc_h == bad, (rc_d, rc_h)
            np.testing.assert_array_equal(h_h["prim"], h_d["prim"])
            np.testing.assert_array_equal(h_h["t"].view(np.uint32), h_d["t"].view(np.uint32))
    rc, h = gpu_ctx.trace_host(rays[:0])
    assert rc == spk.SPK_OK and h.shape[0] == 0
    gpu_ctx.set_trace_config(w, stack, block, 0, 0)
