"""The oracle's twin of two-level scenes (TLAS over instances, one BLAS per mesh, object-space tests
through the float inverse transform), pinned on the CPU.

Pins: with identity transforms the object-space ray is the world ray bit for bit, so the twin equals
the flattened watertight twin exactly; with rotations, non-uniform and mirrored scales the ids agree
with the FP64 oracle on every ray it does not flag, and the hit distance agrees to 1e-6 of the
coordinate scale (|t - t64| <= 1e-6 (t64 + |o|)): an object-space test (the reference's own method
for instances, MeshRenderProxy::IntersectInternal) sees the surface an ulp of the world coordinates
away from the baked world triangles, which a relative bound on t cannot absorb for hits near the
origin; a mutation that drops the translation of the inverse transform is caught.
"""

T_POS_TOL = 1e-6
import numpy as np

import sparkle_amd as spk
from oracle import refcpu
from sparkle_amd import scenes
from tests.helpers import instanced_scene


def _rays(sc, n, seed):
    lo, hi = scenes.scene_aabb(sc)
    return scenes.incoherent_rays(lo, hi, n=n, seed=seed)


def test_identity_instances_match_the_flattened_twin():
    sc, _ = instanced_scene(seed=1, count=9)
    for inst in sc.instances:
        inst["M"] = np.eye(3, 4, dtype=np.float32)
    rays = _rays(sc, 4000, 2)
    o = refcpu.OracleScene(sc)
    a = o.trace(rays, refcpu.TRACE_WT_TWO)
    b = o.trace(rays, refcpu.TRACE_WT_BRUTE)
    np.testing.assert_array_equal(a["prim"], b["prim"])
    np.testing.assert_array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    assert (a["prim"] != spk.SPK_MISS).sum() > 100  # the meshes overlap at the origin, the light far above


def test_transformed_instances_agree_with_fp64():
    sc, _ = instanced_scene(seed=3, count=24)
    rays = _rays(sc, 6000, 4)
    o = refcpu.OracleScene(sc)
    h = o.trace(rays, refcpu.TRACE_WT_TWO)
    h64, flags, t64 = o.trace(rays, refcpu.TRACE_FP64)
    clean = flags == 0
    assert clean.mean() > 0.95
    np.testing.assert_array_equal(h["prim"][clean], h64["prim"][clean])
    hit = clean & (h64["prim"] != spk.SPK_MISS)
    assert hit.mean() > 0.1
    err = np.abs(h["t"][hit].astype(np.float64) - t64[hit])
    scale = t64[hit] + np.linalg.norm(rays["org"][hit].astype(np.float64), axis=1)
    assert (err / scale).max() <= T_POS_TOL, (err / scale).max()
    # the flattened twin (world-space triangles baked from the same transforms) agrees on clean rays
    w = o.trace(rays, refcpu.TRACE_WT_BRUTE)
    np.testing.assert_array_equal(w["prim"][clean], h["prim"][clean])


def test_a_wrong_inverse_is_caught():
    """Dropping the instances' translations on the object-space side only (the mistake of using the
    inverse 3x3 without -inv * t) changes which triangles are hit: the FP64 comparison fails."""
    sc, _ = instanced_scene(seed=5, count=12)
    rays = _rays(sc, 3000, 6)
    o_ok = refcpu.OracleScene(sc)
    h_ok = o_ok.trace(rays, refcpu.TRACE_WT_TWO)
    h64, flags, _ = o_ok.trace(rays, refcpu.TRACE_FP64)
    clean = flags == 0
    assert (h_ok["prim"][clean] == h64["prim"][clean]).all()
    broken = spk.Scene()
    broken.meshes, broken.materials = sc.meshes, sc.materials
    for inst in sc.instances:
        M = inst["M"].copy()
        M[:, 3] = 0.0
        broken.add_instance(inst["mesh"], inst["material"], M)
    h_bad = refcpu.OracleScene(broken).trace(rays, refcpu.TRACE_WT_TWO)
    assert (h_bad["prim"][clean] != h64["prim"][clean]).mean() > 0.05
