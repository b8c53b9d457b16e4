"""Host BVH builder (no GPU): invariants of the quantised wide tree, and input validation."""
import numpy as np
import pytest

import sparkle_amd as spk
from sparkle_amd import scenes
from tests.helpers import multi_mesh_scene, soup_scene


@pytest.mark.parametrize("width", [8, 4])
@pytest.mark.parametrize("leaf", [1, 2, 3])
def test_heightfield_invariants(width, leaf):
    sc = scenes.heightfield_scene(cells=128)
    rc, st, msg = spk.build_check(sc, bvh_width=width, max_leaf_tris=leaf)
    assert rc == spk.SPK_OK, msg
    assert st.num_triangles == 2 * 128 * 128
    assert st.bvh_width == width and 1 <= st.depth <= 64
    assert st.node_bytes == (80 if width == 8 else 64)
    assert st.degenerate_tris == 0


@pytest.mark.parametrize("width", [8, 4])
def test_soups_and_multi_mesh(width):
    for sc in (soup_scene(3000, seed=4), soup_scene(7, seed=5), soup_scene(1, seed=6), multi_mesh_scene()):
        rc, st, msg = spk.build_check(sc, bvh_width=width)
        assert rc == spk.SPK_OK, msg
        assert st.num_triangles == sc.num_triangles


def test_far_from_origin_and_tiny_scenes():
    # quantisation must stay conservative for small triangles far from the origin
    sc = soup_scene(2000, seed=7, scale=1.0, spread=1e-4)
    sc.meshes[0]["positions"] += np.float32(3000.0)
    rc, st, msg = spk.build_check(sc)
    assert rc == spk.SPK_OK, msg
    sc = soup_scene(500, seed=8, scale=1e-3, spread=1e-6)
    rc, st, msg = spk.build_check(sc)
    assert rc == spk.SPK_OK, msg


def test_degenerate_triangles_are_counted_not_rejected():
    pos = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 1, 0]], dtype=np.float32)
    idx = np.array([[0, 1, 2], [0, 1, 3], [1, 1, 1]], dtype=np.uint32)
    sc = spk.Scene()
    sc.add_instance(sc.add_mesh(pos, idx), sc.add_material())
    rc, st, msg = spk.build_check(sc)
    assert rc == spk.SPK_OK, msg
    assert st.degenerate_tris == 2


def test_invalid_scenes_are_rejected_with_a_reason():
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float32)
    sc = spk.Scene()
    sc.add_instance(sc.add_mesh(pos, np.array([[0, 1, 3]], dtype=np.uint32)), sc.add_material())
    rc, _, msg = spk.build_check(sc)
    assert rc == spk.SPK_ERR_INVALID_SCENE and "out of range" in msg
    sc = spk.Scene()
    sc.add_mesh(pos, np.array([[0, 1, 2]], dtype=np.uint32))
    sc.add_instance(5, 0)
    sc.add_material()
    rc, _, msg = spk.build_check(sc)
    assert rc == spk.SPK_ERR_INVALID_SCENE and "mesh" in msg
    bad = pos.copy()
    bad[1, 0] = np.nan
    sc = spk.Scene()
    sc.add_instance(sc.add_mesh(bad, np.array([[0, 1, 2]], dtype=np.uint32)), sc.add_material())
    rc, _, msg = spk.build_check(sc)
    assert rc == spk.SPK_ERR_INVALID_SCENE and "non-finite" in msg
    rc, _, msg = spk.build_check(spk.Scene())
    assert rc == spk.SPK_ERR_INVALID_ARGUMENT


def test_full_size_heightfield_builds():
    sc = scenes.heightfield_scene(cells=1024)
    rc, st, msg = spk.build_check(sc)
    assert rc == spk.SPK_OK, msg
    assert st.num_triangles == 2 * 1024 * 1024
    assert st.depth <= 24
    assert st.build_seconds < 30


def test_batch_scene_generator_spans_the_requested_sizes():
    from sparkle_amd import scenes

    items = scenes.batch_scenes(seed=0, count=64)
    tris = np.array([sc.num_triangles for sc, _, _, _ in items])
    sides = {w for _, _, w, h in items}
    assert all(w == h for _, _, w, h in items)
    assert sides <= {64, 128, 256, 512} and len(sides) == 4
    # log-uniform budgets in [1k, 500k]: small and large scenes both present, within a factor of the range
    assert tris.min() < 5_000 and tris.max() > 200_000
    assert tris.min() >= 1_000 and tris.max() < 1_500_000
    # the many-instance scene itself is unchanged by the per-instance budget cap
    sc, _ = scenes.many_instance_scene(seed=0)
    assert sc.num_triangles == 5_533_044


def test_instanced_many_instance_scene_places_the_same_geometry():
    """The shared form of the many-instance generator (prototype meshes placed by transforms) gives
    the flattened form's world triangles up to float rounding, instance by instance."""
    from sparkle_amd import scenes

    flat, _ = scenes.many_instance_scene(seed=2, count=40, target_tris=200_000)
    shared, _ = scenes.many_instance_scene(seed=2, count=40, target_tris=200_000, shared=True)
    assert len(shared.instances) == len(flat.instances) and shared.num_triangles == flat.num_triangles
    assert len(shared.meshes) < len(flat.meshes)
    for a, b in zip(flat.instances, shared.instances):
        ma, mb = flat.meshes[a["mesh"]], shared.meshes[b["mesh"]]
        np.testing.assert_array_equal(ma["indices"], mb["indices"])
        pa = ma["positions"].astype(np.float64) @ a["M"][:, :3].T.astype(np.float64) + a["M"][:, 3]
        pb = mb["positions"].astype(np.float64) @ b["M"][:, :3].T.astype(np.float64) + b["M"][:, 3]
        np.testing.assert_allclose(pa, pb, atol=2e-5)
        assert a["material"] == b["material"]
