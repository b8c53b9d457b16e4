"""Pins the CPU oracle (no GPU): RNG against an independent implementation and published values,
triangle tests against known answers and FP64, BVH traversal against brute force."""
import math

import numpy as np
import pytest

import sparkle_amd as spk
from sparkle_amd import scenes
from oracle import refcpu
from tests.helpers import soup_scene

M64 = (1 << 64) - 1
M32 = (1 << 32) - 1


def splitmix64(seed):
    x = seed & M64
    while True:
        x = (x + 0x9E3779B97F4A7C15) & M64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        yield z ^ (z >> 31)


def xoshiro128p(seed, n):
    sm = splitmix64(seed)
    s = [next(sm) & M32 for _ in range(4)]
    out = []
    for _ in range(n):
        out.append((s[0] + s[3]) & M32)
        t = (s[1] << 9) & M32
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = ((s[3] << 11) | (s[3] >> 21)) & M32
    return np.array(out, dtype=np.uint32)


def test_splitmix64_published_first_output():
    # SplitMix64 (Steele, Lea, Flood 2014) seeded with 0 first returns 0xe220a8397b1dcdaf
    assert next(splitmix64(0)) == 0xE220A8397B1DCDAF


@pytest.mark.parametrize("seed", [0, 1, 12345, (1 << 32) + 5, 2**63 + 17])
def test_xoshiro_matches_independent_implementation(seed):
    np.testing.assert_array_equal(refcpu.xoshiro_stream(seed, 1000), xoshiro128p(seed, 1000))


def test_random_unit_is_float_of_u32_over_2pow32():
    u = refcpu.xoshiro_stream(77, 5000)
    f = refcpu.random_unit_stream(77, 5000)
    expect = u.astype(np.float32) / np.float32(4294967296.0)
    np.testing.assert_array_equal(f, expect)
    assert f.min() >= 0.0 and f.max() <= 1.0
    # values within 128 of 2^32 round to exactly 1.0, as float(UINT32_MAX) = 2^32
    assert np.float32(np.uint32(M32)) / np.float32(4294967296.0) == 1.0


def test_random_unit_moments():
    f = refcpu.random_unit_stream(3, 200_000).astype(np.float64)
    assert abs(f.mean() - 0.5) < 4 * math.sqrt(1 / 12 / f.size)
    assert abs(f.var() - 1 / 12) < 0.002


# ---- triangle tests -----------------------------------------------------------------------
V0, V1, V2 = [0, 0, 0], [1, 0, 0], [0, 1, 0]


def test_axis_aligned_triangle_known_answers():
    hit, t, u, v = refcpu.tri_mt([0.25, 0.5, 2.0], [0, 0, -1], V0, V1, V2)
    assert hit and t == 2.0 and u == 0.25 and v == 0.5
    hit, t, u, v = refcpu.tri_wt([0.25, 0.5, 2.0], [0, 0, -1], V0, V1, V2)
    assert hit and t == 2.0 and abs(u - 0.25) < 1e-7 and abs(v - 0.5) < 1e-7
    # from below, both tests are two-sided
    assert refcpu.tri_mt([0.25, 0.25, -1.0], [0, 0, 1], V0, V1, V2)[0]
    assert refcpu.tri_wt([0.25, 0.25, -1.0], [0, 0, 1], V0, V1, V2)[0]


def test_miss_outside_and_behind_and_parallel():
    for f in (refcpu.tri_mt, refcpu.tri_wt):
        assert not f([0.75, 0.75, 1.0], [0, 0, -1], V0, V1, V2)[0]  # outside the hypotenuse
        assert not f([0.25, 0.25, 1.0], [0, 0, 1], V0, V1, V2)[0]  # triangle behind the origin
        assert not f([0.25, 0.25, 1.0], [1, 0, 0], V0, V1, V2)[0]  # parallel to the plane


def test_edge_and_vertex_hits_are_inclusive_in_watertight_test():
    assert refcpu.tri_wt([0.5, 0.0, 1.0], [0, 0, -1], V0, V1, V2)[0]  # on edge v0-v1
    assert refcpu.tri_wt([0.5, 0.5, 1.0], [0, 0, -1], V0, V1, V2)[0]  # on the hypotenuse
    assert refcpu.tri_wt([0.0, 0.0, 1.0], [0, 0, -1], V0, V1, V2)[0]  # on a vertex


def test_reference_mt_rejects_a_flat_heightfield_cell():
    # |det| of a 1/1024 cell seen head-on is (1/1024)^2 = 9.5e-7 < 1e-6 (MeshRenderProxy.cpp:143-150)
    h = 1.0 / 1024
    o, d = [0.3 * h, 0.2 * h, 1.0], [0, 0, -1]
    a, b, c = [0, 0, 0], [h, 0, 0], [h, h, 0]
    assert not refcpu.tri_mt(o, d, a, b, c)[0]
    hit, t, _, _ = refcpu.tri_wt(o, d, a, b, c)
    assert hit and t == 1.0


def test_refined_t_axis_plane_closed_form():
    rng = np.random.Generator(np.random.PCG64(4))
    for _ in range(200):
        z = float(np.float32(rng.uniform(-1, 1)))
        o = rng.uniform(-1, 1, 3).astype(np.float32)
        d = rng.normal(size=3)
        d[2] = -abs(d[2]) - 0.05 if o[2] > z else abs(d[2]) + 0.05
        d = (d / np.linalg.norm(d)).astype(np.float32)
        t_exact = (z - float(o[2])) / float(d[2])
        tri = [[-10, -10, z], [10, -10, z], [0, 10, z]]
        t = refcpu.refine_t(o, d, *tri)
        assert abs(t - t_exact) <= 2 * np.spacing(np.float32(t_exact))


# ---- scene traversal ----------------------------------------------------------------------
def random_rays(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    rays = np.zeros(n, dtype=spk.RAY_DTYPE)
    rays["org"] = rng.random((n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    rays["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmax"] = np.inf
    return rays


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_watertight_bvh_equals_brute_force(seed):
    sc = soup_scene(800, seed=seed)
    rays = random_rays(20_000, seed + 100)
    o = refcpu.OracleScene(sc)
    a = o.trace(rays, refcpu.TRACE_WT)
    b = o.trace(rays, refcpu.TRACE_WT_BRUTE)
    np.testing.assert_array_equal(a["prim"], b["prim"])
    np.testing.assert_array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))


def test_fp64_agrees_with_watertight_on_unflagged_rays():
    sc = scenes.heightfield_scene(cells=96)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=50_000, seed=2)
    o = refcpu.OracleScene(sc)
    wt = o.trace(rays, refcpu.TRACE_WT)
    h64, flags, t64 = o.trace(rays, refcpu.TRACE_FP64)
    clean = flags == 0
    assert clean.mean() > 0.99
    np.testing.assert_array_equal(wt["prim"][clean], h64["prim"][clean])
    m = clean & (h64["prim"] != spk.SPK_MISS)
    rel = np.abs(wt["t"][m] - t64[m]) / t64[m]
    assert rel.max() < 1e-6


def test_fp64_flags_shared_edges():
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], dtype=np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)
    sc = spk.Scene()
    sc.add_instance(sc.add_mesh(pos, idx), sc.add_material())
    rays = np.zeros(2, dtype=spk.RAY_DTYPE)
    rays["org"] = [[0.5, 0.5, 1.0], [0.7, 0.2, 1.0]]
    rays["dir"] = [[0, 0, -1], [0, 0, -1]]
    rays["tmax"] = np.inf
    h64, flags, _ = refcpu.OracleScene(sc).trace(rays, refcpu.TRACE_FP64)
    assert flags[0] & refcpu.FLAG_EDGE and flags[0] & refcpu.FLAG_TIE
    assert flags[1] == 0 and h64["prim"][1] == 0


def test_reference_traversal_semantics():
    """REF_MT: closest hit over meshes with a shared candidate; d.x == 0 rays never hit a mesh
    (world_t = (p - o).x / d.x is NaN, MeshRenderProxy.cpp:179-182)."""
    sc = scenes.heightfield_scene(cells=32)
    o = refcpu.OracleScene(sc)
    rays = np.zeros(2, dtype=spk.RAY_DTYPE)
    rays["org"] = [[0.5013, 0.4987, 1.0], [0.5013, 0.4987, 1.0]]
    rays["dir"] = [[0.0, 0.0, -1.0], [0.01, 0.0, -0.99995]]
    rays["tmax"] = np.inf
    mt = o.trace(rays, refcpu.TRACE_REF_MT)
    wt = o.trace(rays, refcpu.TRACE_WT)
    assert mt["prim"][0] == spk.SPK_MISS and wt["prim"][0] != spk.SPK_MISS
    assert mt["prim"][1] == wt["prim"][1]
    assert abs(mt["t"][1] - wt["t"][1]) < 1e-5


def test_reference_mt_agrees_with_fp64_on_a_coarse_mesh():
    # triangles large enough for the |det| >= 1e-6 cutoff: the reference test finds the same hits
    sc = scenes.heightfield_scene(cells=32)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=20_000, seed=8)
    rays["dir"][:, 0] = np.where(rays["dir"][:, 0] == 0, np.float32(1e-3), rays["dir"][:, 0])
    o = refcpu.OracleScene(sc)
    mt = o.trace(rays, refcpu.TRACE_REF_MT)
    h64, flags, t64 = o.trace(rays, refcpu.TRACE_FP64)
    clean = flags == 0
    agree = (mt["prim"][clean] == h64["prim"][clean]).mean()
    assert agree > 0.999


def test_instance_transforms_in_the_reference_traversal():
    # a translated instance must be hit where its world-space copy is
    p, i, n, t, uv = scenes.icosphere(2, radius=0.3)
    a = spk.Scene()
    ma = a.add_material()
    a.add_instance(a.add_mesh(p, i, n, t, uv), ma, np.array([[1, 0, 0, 0.5], [0, 1, 0, 0.5], [0, 0, 1, 0.5]]))
    b = spk.Scene()
    b.add_instance(b.add_mesh(p + 0.5, i, n, t, uv), b.add_material())
    rays = random_rays(5000, 3)
    ha = refcpu.OracleScene(a).trace(rays, refcpu.TRACE_REF_MT)
    hb = refcpu.OracleScene(b).trace(rays, refcpu.TRACE_REF_MT)
    same = (ha["prim"] == hb["prim"]).mean()
    assert same > 0.999
    m = (ha["prim"] == hb["prim"]) & (ha["prim"] != spk.SPK_MISS)
    assert np.abs(ha["t"][m] - hb["t"][m]).max() < 1e-5


def test_edge_on_false_hits_are_rejected():
    """Regression: for these rays the FP32 edge functions of distant, edge-on heightfield cells
    round to exact zeros; the certified test re-evaluates them and finds no hit (as FP64 does)."""
    sc = scenes.heightfield_scene(cells=1024)
    lo, hi = scenes.scene_aabb(sc)
    rays = scenes.incoherent_rays(lo, hi, n=8 * 1024 * 1024, seed=0)
    ids = np.array([3764800, 3784576, 4979136, 6322624, 6998848])
    o = refcpu.OracleScene(sc)
    br = o.trace(rays[ids], refcpu.TRACE_WT_BRUTE)
    bv = o.trace(rays[ids], refcpu.TRACE_WT)
    h64, flags, _ = o.trace(rays[ids], refcpu.TRACE_FP64)
    np.testing.assert_array_equal(br["prim"], h64["prim"])
    np.testing.assert_array_equal(bv["prim"], h64["prim"])
    # the first one: both plain FP32 edge values are exactly zero for a cell 0.87 away
    P = sc.meshes[0]["positions"]
    I = sc.meshes[0]["indices"]
    assert not refcpu.tri_wt(rays["org"][ids[0]], rays["dir"][ids[0]], *P[I[466407]])[0]


def test_analytic_sphere_closed_forms():
    """SphereRenderProxy: near root from outside, far root from inside, misses, id num_tris + k,
    and ordering against triangles (smaller t wins; the triangle wins an exact tie)."""
    sc = spk.Scene()
    m = sc.add_mesh(np.array([[-1, -1, 3], [1, -1, 3], [0, 1, 3]], np.float32), np.array([[0, 1, 2]], np.uint32),
                    np.array([[0, 0, -1]] * 3, np.float32))
    mat = sc.add_material()
    sc.add_instance(m, mat)
    sc.add_sphere((0.0, 0.0, 0.0), 1.0, mat)
    sc.add_sphere((0.0, 0.0, 6.0), 0.5, mat)
    o = refcpu.OracleScene(sc)
    rays = np.zeros(5, dtype=spk.RAY_DTYPE)
    rays["org"] = [[0, 0, -5], [0, 0, 0], [0, 2, -5], [0, 0, 4], [0.25, 0.25, -5]]
    rays["dir"] = [[0, 0, 1], [0, 0, 1], [0, 0, 1], [0, 0, 1], [0, 0, 1]]
    rays["tmax"] = np.inf
    for mode in (refcpu.TRACE_REF_MT, refcpu.TRACE_WT):
        h = o.trace(rays, mode)
        assert h["prim"][0] == 1 and h["t"][0] == 4.0          # outside: near root
        assert h["prim"][1] == 1 and h["t"][1] == 1.0          # inside: far root
        assert h["prim"][2] == spk.SPK_MISS                    # passes beside both
        assert h["prim"][3] == 2 and h["t"][3] == 1.5          # second sphere, id num_tris + 1
        # through the unit sphere: its far side at z = +sqrt(1 - 0.125) is nearer than the triangle
        assert h["prim"][4] == 1
    far = rays[[0]].copy()
    far["org"] = [[0.0, 0.0, 2.0]]
    h = o.trace(far, refcpu.TRACE_WT)
    assert h["prim"][0] == 0 and h["t"][0] == 1.0              # the triangle at z = 3 before the sphere at 5.5
