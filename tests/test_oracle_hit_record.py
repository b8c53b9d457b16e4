"""The oracle's hit record (MeshRenderProxy::GetIntersection, MeshRenderProxy.cpp:340-361; Mesh.h
attribute interpolation), pinned on the CPU through the first-hit normal G-buffer.

Pins: barycentric interpolation reproduces affine fields exactly, so vertex normals that are an affine
function of position, (a x + b, c y + d, 1), must come back as that field at the hit point, normalised:
each pixel's normal decodes to a point inside its own footprint on the quad (a swapped or dropped
barycentric term puts it elsewhere); a flat mesh (equal vertex normals) gives that normal on every hit.
"""
import numpy as np

import sparkle_amd as spk
from oracle import refcpu

A, B, C, D = 0.3, 0.1, -0.2, 0.05


def _quad(normals):
    pos = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    sc = spk.Scene()
    sc.add_instance(sc.add_mesh(pos, idx, normals=np.asarray(normals, np.float32)), sc.add_material())
    return sc


def _render(sc, w=48):
    cam = spk.camera_desc(position=(0.0, 0.0, 3.0), front=(0.0, 0.0, -1.0), up_hint=(0.0, 1.0, 0.0),
                          vertical_fov=np.radians(30.0), aspect=1.0, focus_distance=3.0)
    _, nrm = refcpu.OracleScene(sc).render_frame(cam, w, w, frame_seed=2, max_bounce=1)
    return cam, nrm


def test_interpolated_normals_reproduce_an_affine_field():
    pos = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float64)
    nrm = np.stack([A * pos[:, 0] + B, C * pos[:, 1] + D, np.ones(4)], 1)
    w = 48
    cam, g = _render(_quad(nrm), w)
    hit = np.abs(g).sum(-1) > 0
    assert hit.mean() > 0.5
    n = g[hit].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)
    # decode the hit point from the normal field and compare with the pixel's footprint on z = 0
    x = (n[:, 0] / n[:, 2] - B) / A
    y = (n[:, 1] / n[:, 2] - D) / C
    jj, ii = np.nonzero(hit)
    P = np.array(cam.position[:]); LL = np.array(cam.lower_left[:])
    MU = np.array(cam.max_u[:]); MV = np.array(cam.max_v[:])
    u = (ii + 0.5) / (w - 1)
    v = (jj + 0.5) / (w - 1)
    focus = LL[None] + MU[None] * u[:, None] + MV[None] * v[:, None]
    d = focus - P[None]
    t = -P[2] / d[:, 2]
    ph = P[None] + d * t[:, None]
    pixel = np.linalg.norm(MU) / (w - 1)  # one pixel on the focus plane, which is the quad's plane
    assert np.abs(x - ph[:, 0]).max() < pixel and np.abs(y - ph[:, 1]).max() < pixel
    assert (np.abs(x) <= 1 + 1e-5).all() and (np.abs(y) <= 1 + 1e-5).all()


def test_flat_mesh_normal_is_constant():
    n0 = np.array([0.2, -0.1, 1.0]) / np.linalg.norm([0.2, -0.1, 1.0])
    _, g = _render(_quad(np.tile(n0, (4, 1))))
    hit = np.abs(g).sum(-1) > 0
    assert hit.mean() > 0.5
    np.testing.assert_allclose(g[hit], np.tile(n0, (hit.sum(), 1)), atol=1e-6)
