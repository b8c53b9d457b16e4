"""Camera records the device consumes (SetupViewRay's inputs), pinned on the CPU.

The orbit camera (OrbitCameraComponent::Setup / UpdateTransform) is checked against the reference's
own inverse, OrbitCameraComponent::SetupFromTransform (OrbitCameraComponent.cpp:33-52: yaw =
atan2(f.x, f.y), pitch = -asin(f.z), centre = position + f * radius), which derives the angles from
the camera's forward vector by a different route than the rotation that builds it; the default
camera every loaded scene starts with (SceneManager.cpp:142-147) against its closed form.
"""
import math

import numpy as np
import pytest

import sparkle_amd as spk


def test_default_scene_camera_closed_form():
    c = spk.orbit_camera(aspect=1.0)
    np.testing.assert_allclose(c.position[:], [0.0, -10.0, 0.0], atol=1e-6)
    np.testing.assert_allclose(c.front[:], [0.0, 1.0, 0.0], atol=1e-7)
    np.testing.assert_allclose(c.right[:], [1.0, 0.0, 0.0], atol=1e-7)
    np.testing.assert_allclose(c.up[:], [0.0, 0.0, 1.0], atol=1e-7)
    # 35 mm lens on a 24 mm sensor: 37.85 degrees; f/22: lens radius 7.95e-4; focus on the centre
    h = 2.0 * 10.0 * (0.024 / (2 * 0.035))
    np.testing.assert_allclose(c.max_v[:], [0.0, 0.0, h], rtol=1e-6)
    np.testing.assert_allclose(c.max_u[:], [h, 0.0, 0.0], rtol=1e-6)
    np.testing.assert_allclose(c.lower_left[:], [-h / 2, 0.0, -h / 2], rtol=1e-6, atol=1e-6)
    assert math.degrees(2 * math.atan(0.024 / 0.07)) == pytest.approx(37.85, abs=0.01)
    assert c.lens_radius == pytest.approx(7.95e-4, rel=1e-3)


@pytest.mark.parametrize("pitch,yaw", [(30.0, 0.0), (0.0, 90.0), (-20.0, 135.0), (45.0, -60.0), (10.0, 170.0)])
def test_orbit_camera_inverts_through_setup_from_transform(pitch, yaw):
    center = np.array([0.3, -1.2, 0.7])
    radius = 4.5
    pos, right, front, up = spk.orbit_basis(center, radius, pitch, yaw)
    # SetupFromTransform's recovery of the angles and the orbit centre
    assert math.degrees(math.atan2(front[0], front[1])) == pytest.approx(yaw, abs=1e-9)
    assert math.degrees(-math.asin(front[2])) == pytest.approx(pitch, abs=1e-9)
    np.testing.assert_allclose(pos + front * radius, center, atol=1e-12)
    # a proper rotation: orthonormal rows, right x front = up, and no roll (right stays horizontal)
    basis = np.stack([right, front, up])
    np.testing.assert_allclose(basis @ basis.T, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(np.cross(right, front), up, atol=1e-12)
    assert abs(right[2]) < 1e-12
    # positive pitch looks down, toward -z
    assert (front[2] < 0) == (pitch > 0)


def test_orbit_camera_record_matches_the_focus_plane():
    c = spk.orbit_camera(center=(1.0, 2.0, 0.5), radius=3.0, pitch=25.0, yaw=40.0, aspect=16 / 9)
    P, F = np.array(c.position[:]), np.array(c.front[:])
    # the focus plane centre lies on the orbit centre, at the focus distance = radius
    centre = np.array(c.lower_left[:]) + 0.5 * np.array(c.max_u[:]) + 0.5 * np.array(c.max_v[:])
    np.testing.assert_allclose(centre, [1.0, 2.0, 0.5], atol=2e-6)
    np.testing.assert_allclose(P + 3.0 * F, [1.0, 2.0, 0.5], atol=2e-6)
    assert np.linalg.norm(c.max_u[:]) / np.linalg.norm(c.max_v[:]) == pytest.approx(16 / 9, rel=1e-6)
