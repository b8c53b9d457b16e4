"""Pins the oracle's integrator (no GPU): closed forms of the camera, sky and emitters, the clamp,
the running mean, thread-count invariance and seed handling."""
import math

import numpy as np

import sparkle_amd as spk
from sparkle_amd import scenes
from oracle import refcpu


def test_camera_center_ray_is_front_and_corners_span_the_focus_plane():
    cam = spk.camera_desc(position=(1, 2, 3), front=(0, 1, 0), vertical_fov=math.radians(60), focus_distance=2.0,
                          aspect=2.0)
    ll = np.array(cam.lower_left)
    mu, mv = np.array(cam.max_u), np.array(cam.max_v)
    center = ll + 0.5 * mu + 0.5 * mv
    np.testing.assert_allclose(center, [1, 4, 3], atol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(mv), 2 * math.tan(math.radians(30)) * 2.0, rtol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(mu), 2 * np.linalg.norm(mv), rtol=1e-6)
    np.testing.assert_allclose(cam.right, [1, 0, 0], atol=1e-7)
    np.testing.assert_allclose(cam.up, [0, 0, 1], atol=1e-7)


def _room_with_light_only():
    sc = spk.Scene()
    pos = np.array([[-1, 1, -1], [1, 1, -1], [1, 1, 1], [-1, 1, 1]], dtype=np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)
    nrm = np.tile([0, -1, 0], (4, 1)).astype(np.float32)
    tan = np.tile([1, 0, 0, 1], (4, 1)).astype(np.float32)
    m = sc.add_mesh(pos, idx, nrm, tan)
    light = sc.add_material(emissive=(15.0, 2.0, 0.5), metallic=0.0)
    sc.add_instance(m, light)
    return sc


def test_directly_viewed_emitter_is_clamped_to_six():
    sc = _room_with_light_only()
    cam = spk.camera_desc(position=(0, -1, 0), front=(0, 1, 0), vertical_fov=math.radians(20), focus_distance=2.0)
    rgba, nrm = refcpu.OracleScene(sc).render_frame(cam, 16, 16, frame_seed=0)
    # every pixel sees the quad: emission (15, 2, 0.5) clamped per channel to 6
    np.testing.assert_array_equal(rgba[..., 0], 6.0)
    np.testing.assert_allclose(rgba[..., 1], 2.0)
    np.testing.assert_allclose(rgba[..., 2], 0.5)
    assert (rgba[..., 3] == 1.0).all()
    np.testing.assert_allclose(nrm[..., 1], -1.0)


def test_sky_gradient_closed_form():
    sc = _room_with_light_only()
    # look away from the quad: every ray misses and returns lerp(1, sky, 0.5 d.z + 0.5)
    cam = spk.camera_desc(position=(0, -1, 0), front=(0, -1, 0.3), vertical_fov=math.radians(40),
                          focus_distance=1.0, aperture_radius=0.0)
    sky = np.array([0.2, 0.5, 0.9], dtype=np.float32)
    w = h = 8
    rgba, _ = refcpu.OracleScene(sc).render_frame(cam, w, h, frame_seed=1, sky=tuple(sky))
    assert (rgba[..., :3] > 0).all()
    # recover each pixel's direction z from the colour and check it is a unit-length direction
    t = (rgba[..., 0] - 1.0) / (sky[0] - 1.0)
    dz = 2 * t - 1
    t2 = (rgba[..., 1] - 1.0) / (sky[1] - 1.0)
    np.testing.assert_allclose(t, t2, atol=2e-6)
    assert (np.abs(dz) <= 1 + 1e-5).all()
    # rows further up the image look higher
    assert (np.diff(dz.mean(1)) > 0).all()


def test_thread_count_invariance():
    sc = scenes.cornell_scene()
    cam = scenes.cornell_camera(32, 32)
    o = refcpu.OracleScene(sc)
    a, an = o.render_frame(cam, 32, 32, frame_seed=4, threads=1)
    b, bn = o.render_frame(cam, 32, 32, frame_seed=4, threads=8)
    assert np.array_equal(a, b) and np.array_equal(an, bn)


def test_seed_changes_per_frame_and_wraps_in_uint32():
    sc = scenes.cornell_scene()
    cam = scenes.cornell_camera(16, 16)
    o = refcpu.OracleScene(sc)
    a, _ = o.render_frame(cam, 16, 16, frame_seed=1)
    b, _ = o.render_frame(cam, 16, 16, frame_seed=2)
    assert not np.array_equal(a, b)
    # frame f and f + 2^32/(W*H) wrap to the same uint32 pixel seeds (CPURenderer.cpp:442-443)
    wrap = (1 << 32) // (16 * 16)
    c, _ = o.render_frame(cam, 16, 16, frame_seed=1 + wrap)
    assert np.array_equal(a, c)
    d, _ = o.render_frame(cam, 16, 16, frame_seed=1 + wrap, seed64=True)
    assert not np.array_equal(a, d)


def test_running_mean_is_the_arithmetic_mean():
    rng = np.random.Generator(np.random.PCG64(0))
    acc = np.zeros((4, 4, 4), dtype=np.float32)
    xs = []
    for n in range(10):
        x = rng.random((4, 4, 4)).astype(np.float32)
        xs.append(x)
        refcpu.accumulate(x, acc, n, 1)
    np.testing.assert_allclose(acc, np.mean(xs, 0), rtol=1e-5)


def test_samples_are_clamped_and_valid_flags_are_signed():
    sc = scenes.cornell_scene()
    cam = scenes.cornell_camera(32, 32)
    rgba, _ = refcpu.OracleScene(sc).render_frame(cam, 32, 32, frame_seed=0)
    assert rgba[..., :3].max() <= 6.0 and rgba[..., :3].min() >= 0.0
    assert set(np.unique(rgba[..., 3])).issubset({-1.0, 1.0})
    # the light is visible: some samples reach the clamp
    assert (rgba[..., :3] == 6.0).any()
