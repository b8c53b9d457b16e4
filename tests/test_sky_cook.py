"""The textured sky light's cook (SkyLightCookJob::Execute, SkyLight.cpp:83-135) and its loading from
USD (LoadSkyLight, USDLoader.cpp:474-497), no GPU.

The library's cook (spk_sky_cook) is compared bit for bit with the oracle's restatement; the oracle
is pinned by what the mathematics fixes: FP16 rounding against numpy's binary16 conversion, a
constant map cooking to constant faces, a map that is a linear ramp in the equirectangular
coordinates cooking to atan2/asin of each texel centre's direction (which fixes the face
orientations, the texel-centre convention and CartesianToEquirectangular), and the cooked cube
sampled back through the renderer's Image2DCube lookup returning the direction it was asked for.
"""
import math
import os

import numpy as np
import pytest

import sparkle_amd as spk
from sparkle_amd import usd
from oracle import refcpu


def test_half_rounding_matches_binary16():
    rng = np.random.Generator(np.random.PCG64(1))
    vals = np.concatenate([
        rng.normal(0, 1, 2000), rng.normal(0, 1000, 2000), rng.uniform(-7e4, 7e4, 500),
        np.ldexp(rng.uniform(0.5, 1, 500), rng.integers(-30, -10, 500)),  # subnormal halves and below
        1.0 + np.arange(0, 64) * 2.0 ** -11,  # exact ties between neighbouring halves
        np.array([65504.0, 65519.0, 65520.0, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, 0.0, -0.0])]).astype(np.float32)
    with np.errstate(over="ignore"):  # values past 65520 round to infinity, as binary16 does
        ref = vals.astype(np.float16).astype(np.float32)
    got = np.array([refcpu.half_round(v) for v in vals], np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_library_cook_equals_oracle_bit_for_bit():
    rng = np.random.Generator(np.random.PCG64(2))
    for shape, face in (((32, 64, 4), 48), ((17, 41, 3), 33), ((8, 16, 4), 5)):
        img = (rng.random(shape) * rng.choice([1.0, 50.0, 7e4])).astype(np.float32)
        src = spk.sky_source_texels(img)
        assert (src[..., :3] <= 65504.0).all()
        a = np.stack(spk.sky_cook(img, face))
        b = refcpu.sky_cook(src, face)
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    ldr = rng.integers(0, 256, (16, 32, 4), dtype=np.uint8)
    src = spk.sky_source_texels(ldr)
    np.testing.assert_array_equal(src, ldr.astype(np.float32) / np.float32(255.0))
    np.testing.assert_array_equal(np.stack(spk.sky_cook(ldr, 24)), refcpu.sky_cook(src, 24))


def test_constant_map_cooks_to_constant_faces():
    img = np.zeros((9, 20, 4), np.float32)
    img[...] = (0.3, 1.7, 1234.567, 1.0)
    faces = refcpu.sky_cook(img, 16)
    expect = np.array([0.3, 1.7, 1234.567], np.float32).astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(faces[..., :3], np.broadcast_to(expect, faces[..., :3].shape))
    assert (faces[..., 3] == 1.0).all()


def _texel_directions(n):
    """Directions of every cube texel centre in float64, face by face as Image2DCube lays them out:
    +X (1, -v, -u), -X (-1, -v, u), +Y (u, 1, v), -Y (u, -1, -v), +Z (u, -v, 1), -Z (-u, -v, -1),
    u from the column and v from the row, texel centres at (k + 0.5) / n * 2 - 1."""
    c = (np.arange(n) + 0.5) / n * 2 - 1
    v, u = np.meshgrid(c, c, indexing="ij")  # [row j, column i]
    one = np.ones_like(u)
    d = np.stack([np.stack(x, -1) for x in ((one, -v, -u), (-one, -v, u), (u, one, v), (u, -one, -v),
                                                (u, -v, one), (-u, -v, -one))])
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def test_linear_ramp_map_cooks_to_the_equirectangular_coordinates():
    """Texel (i, j) of a w x h map holds (i / (w - 1), j / (h - 1)): bilinear sampling at (eu, ev)
    returns (eu, ev) itself away from the wrap seam, so every cube texel must hold
    eu = atan2(y, x) / 2pi + 1/2 and ev = asin(-z) / pi + 1/2 of its direction (to FP16)."""
    w, h, n = 257, 129, 40
    jj, ii = np.mgrid[0:h, 0:w]
    img = np.stack([ii / (w - 1), jj / (h - 1), np.full((h, w), 0.5), np.ones((h, w))], -1).astype(np.float32)
    faces = refcpu.sky_cook(img, n)
    d = _texel_directions(n)
    eu = np.arctan2(d[..., 1], d[..., 0]) / (2 * np.pi) + 0.5
    ev = np.arcsin(-d[..., 2]) / np.pi + 0.5
    away = (eu > 2.0 / (w - 1)) & (eu < 1 - 2.0 / (w - 1))
    assert away.mean() > 0.98
    tol = 2e-3 * np.maximum(eu, 1e-3)  # binary16 keeps 11 bits; the float32 path adds a few ulps
    assert (np.abs(faces[..., 0] - eu)[away] <= tol[away] + 1e-6).all()
    assert (np.abs(faces[..., 1] - ev) <= 2e-3 * ev + 1e-6).all()
    assert (faces[..., 2] == 0.5).all()
    # a face permuted or mirrored is caught: the +X face sees eu near 1/2 (x > 0), the -Z face ev > 1/2
    assert abs(faces[0, ..., 0].mean() - 0.5) < 0.02 and faces[5, ..., 1].min() > 0.5


def test_cooked_cube_samples_back_to_the_ray_direction():
    """The oracle renders a pinhole at the origin whose rays all miss: each pixel is the cube sample
    of its ray direction (Image2DCube::DirectionToTextureCoordinate + Sample). With the ramp map
    the colour encodes (eu, ev), which decodes to a direction within the pixel's footprint."""
    w, h, n = 257, 129, 256
    jj, ii = np.mgrid[0:h, 0:w]
    img = np.stack([ii / (w - 1), jj / (h - 1), np.zeros((h, w)), np.ones((h, w))], -1).astype(np.float32)
    sc = spk.Scene()
    far = sc.add_mesh(np.array([[500, 0, 0], [501, 0, 0], [500, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32))
    sc.add_instance(far, sc.add_material(base_color=(0.5, 0.5, 0.5), metallic=0.0))
    sc.set_sky_map(img, n)
    W, H = 48, 32
    for front in ((1.0, 0.3, 0.2), (-0.4, -1.0, 0.5), (0.1, 0.2, -1.0)):
        cam = spk.camera_desc(position=(0.0, 0.0, 0.0), front=front, vertical_fov=math.radians(60.0),
                              focus_distance=1.0, aperture_radius=0.0, aspect=W / H)
        rgba, _ = refcpu.OracleScene(sc).render_frame(cam, W, H, frame_seed=0, max_bounce=1, sky=(1, 1, 1))
        phi = (rgba[..., 0].astype(np.float64) - 0.5) * 2 * np.pi
        z = -np.sin((rgba[..., 1].astype(np.float64) - 0.5) * np.pi)
        r = np.sqrt(np.maximum(0.0, 1 - z * z))
        d = np.stack([r * np.cos(phi), r * np.sin(phi), z], -1)
        # the nominal direction of each pixel centre
        ll, mu, mv = (np.array(x, np.float64) for x in (cam.lower_left, cam.max_u, cam.max_v))
        u = (np.arange(W) + 0.5) / (W - 1)
        v = (np.arange(H) + 0.5) / (H - 1)
        p = ll + u[None, :, None] * mu + v[:, None, None] * mv
        p /= np.linalg.norm(p, axis=-1, keepdims=True)
        ang = np.arccos(np.clip((d * p).sum(-1), -1, 1))
        # cube texels on either side of the map's wrap seam (the half plane y = 0, x < 0) hold eu near 0
        # and near 1, so a lookup between them decodes to no direction: leave a band around it out
        seam = np.abs(np.arctan2(p[..., 1], p[..., 0])) > math.pi - math.radians(6.0)
        assert (~seam).mean() > 0.7
        assert (ang[~seam] < math.radians(1.5)).all(), np.degrees(ang[~seam].max())


def test_usd_textured_dome_is_cooked_into_the_sky(tmp_path):
    from tests.test_usd import _glass_cornell

    rng = np.random.Generator(np.random.PCG64(4))
    sky_map = (rng.random((16, 32, 3)) * 4).astype(np.float32)
    path = str(tmp_path / "dome" / "scene.usda")
    os.makedirs(os.path.dirname(path))
    usd.write_usda(path, _glass_cornell(), sky=(0.2, 0.4, 0.6), sky_map=sky_map)
    back = usd.read_usda(path, sky_face_size=32)
    assert back.sky == tuple(float(np.float32(x)) for x in (0.2, 0.4, 0.6))
    assert back.sky_map_file.endswith(os.path.join("textures", "sky.hdr"))
    faces = back.scene.sky_faces
    assert faces is not None and len(faces) == 6
    with open(back.sky_map_file, "rb") as fh:
        decoded = usd.hdr_decode(fh.read())
    # RGBE keeps 8 bits relative to each pixel's largest channel (a shared exponent)
    assert (np.abs(decoded[..., :3] - sky_map) <= 2 ** -7 * sky_map.max(-1, keepdims=True)).all()
    cooked = np.stack(spk.sky_cook(decoded, 32))
    for k, t in enumerate(faces):
        px, fmt = back.scene.textures[t]
        assert fmt == spk.TEXTURE_RGBA32F
        np.testing.assert_array_equal(px, cooked[k])
    assert not any("dome" in s for s in back.skipped)
