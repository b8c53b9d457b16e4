"""USD (ASCII) reader and writer: the reference loader's mapping rules on hand-written stages, and
write -> read round trips that must give back the same scene and the same image.

The hand-written stage follows the layout of the reference's loader test
(tests/usd/UsdLoaderSemanticsTest.cpp: a distant light of colour (0.25, 0.5, 0.75), intensity 3,
exposure 2 -> (3, 6, 9); a dome light (0.1, 0.2, 0.3) x 2 -> (0.2, 0.4, 0.6); a triangle whose st
equals its xy, textured from an srgb_texture and a lin_srgb image; an untextured triangle without
uvs that still gets tangents) plus transforms, polygons, face-varying normals, a camera and a
transmissive material, whose expected values follow from USD's and the loader's definitions.
"""
import math
import os

import numpy as np
import pytest

import sparkle_amd as spk
from oracle import refcpu
from sparkle_amd import scenes, usd
from sparkle_amd.gltf import _png_rgba

STAGE = '''#usda 1.0
(
    defaultPrim = "Root"
    upAxis = "Z"
)

def Xform "Root"
{
    double3 xformOp:translate = (1, 2, 3)
    float xformOp:rotateZ = 90
    uniform token[] xformOpOrder = ["xformOp:translate", "xformOp:rotateZ"]

    def DistantLight "Sun"
    {
        color3f inputs:color = (0.25, 0.5, 0.75)
        float inputs:exposure = 2
        float inputs:intensity = 3
    }

    def DomeLight "Sky"
    {
        color3f inputs:color = (0.1, 0.2, 0.3)
        float inputs:intensity = 2
    }

    def Mesh "TexturedTriangle" (
        prepend apiSchemas = ["MaterialBindingAPI"]
    )
    {
        int[] faceVertexCounts = [3]
        int[] faceVertexIndices = [0, 1, 2]
        rel material:binding = </Root/Materials/TestMaterial>
        normal3f[] normals = [(0, 0, 1), (0, 0, 1), (0, 0, 1)] (
            interpolation = "vertex"
        )
        point3f[] points = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]
        texCoord2f[] primvars:st = [(0, 0), (1, 0), (0, 1)] (
            interpolation = "vertex"
        )
        uniform token subdivisionScheme = "none"
    }

    def Mesh "UntexturedTriangle" (
        prepend apiSchemas = ["MaterialBindingAPI"]
    )
    {
        int[] faceVertexCounts = [3]
        int[] faceVertexIndices = [0, 1, 2]
        rel material:binding = </Root/Materials/PlainMaterial>
        normal3f[] normals = [(0, 0, 1), (0, 0, 1), (0, 0, 1)] (
            interpolation = "vertex"
        )
        point3f[] points = [(2, 0, 0), (3, 0, 0), (2, 1, 0)]
        uniform token subdivisionScheme = "none"
    }

    def Xform "Child"
    {
        matrix4d xformOp:transform = ( (2, 0, 0, 0), (0, 2, 0, 0), (0, 0, 2, 0), (10, 0, 0, 1) )
        uniform token[] xformOpOrder = ["xformOp:transform"]

        def Mesh "Quad" (
            prepend apiSchemas = ["MaterialBindingAPI"]
        )
        {
            int[] faceVertexCounts = [4]
            int[] faceVertexIndices = [0, 1, 2, 3]
            rel material:binding = </Root/Materials/Glass>
            normal3f[] normals = [(0, 0, 1), (0, 0, 1), (0, 0, 1), (0, 0, -1)] (
                interpolation = "faceVarying"
            )
            point3f[] points = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]
        }

        def Mesh "Unbound"
        {
            int[] faceVertexCounts = [3]
            int[] faceVertexIndices = [0, 1, 2]
            point3f[] points = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]
        }
    }

    def Xform "CameraNode"
    {
        double3 xformOp:translate = (0, -5, 1)
        uniform token[] xformOpOrder = ["xformOp:translate"]

        def Camera "Cam"
        {
            float focalLength = 50
            float verticalAperture = 24
            float fStop = 2.8
            float focusDistance = 5
            token projection = "perspective"
        }
    }

    def Scope "Materials"
    {
        def Material "TestMaterial"
        {
            token outputs:surface.connect = </Root/Materials/TestMaterial/Surface.outputs:surface>

            def Shader "StReader"
            {
                uniform token info:id = "UsdPrimvarReader_float2"
                string inputs:varname = "st"
                float2 outputs:result
            }

            def Shader "SrgbTexture"
            {
                uniform token info:id = "UsdUVTexture"
                asset inputs:file = @srgb.png@ (
                    colorSpace = "srgb_texture"
                )
                float2 inputs:st.connect = </Root/Materials/TestMaterial/StReader.outputs:result>
                float3 outputs:rgb
            }

            def Shader "LinearTexture"
            {
                uniform token info:id = "UsdUVTexture"
                asset inputs:file = @linear.png@ (
                    colorSpace = "lin_srgb"
                )
                float2 inputs:st.connect = </Root/Materials/TestMaterial/StReader.outputs:result>
                float3 outputs:rgb
            }

            def Shader "Surface"
            {
                uniform token info:id = "UsdPreviewSurface"
                color3f inputs:diffuseColor.connect = </Root/Materials/TestMaterial/SrgbTexture.outputs:rgb>
                color3f inputs:emissiveColor.connect = </Root/Materials/TestMaterial/LinearTexture.outputs:rgb>
                float inputs:metallic = 0
                float inputs:roughness = 1
                token outputs:surface
            }
        }

        def Material "PlainMaterial"
        {
            token outputs:surface.connect = </Root/Materials/PlainMaterial/Surface.outputs:surface>

            def Shader "Surface"
            {
                uniform token info:id = "UsdPreviewSurface"
                color3f inputs:diffuseColor = (0.5, 0.5, 0.5)
                float inputs:metallic = 0
                float inputs:roughness = 1
                token outputs:surface
            }
        }

        def Material "Glass"
        {
            token outputs:surface.connect = </Root/Materials/Glass/Surface.outputs:surface>

            def Shader "Surface"
            {
                uniform token info:id = "UsdPreviewSurface"
                float inputs:opacity = 0
                float inputs:ior = 1.33
                token outputs:surface
            }
        }
    }
}
'''


@pytest.fixture
def stage(tmp_path):
    px = np.array([[[64, 128, 191, 255]]], np.uint8)
    (tmp_path / "srgb.png").write_bytes(_png_rgba(px))
    (tmp_path / "linear.png").write_bytes(_png_rgba(px))
    p = tmp_path / "scene.usda"
    p.write_text(STAGE)
    return usd.read_usda(str(p))


def _world_of(u, name):
    k = u.mesh_names.index(name)
    inst = u.scene.instances[k]
    return u.scene.meshes[inst["mesh"]], inst["M"], u.scene.materials[inst["material"]]


def test_lights_apply_intensity_and_exposure(stage):
    np.testing.assert_allclose(stage.directional, (3.0, 6.0, 9.0), rtol=1e-6)
    np.testing.assert_allclose(stage.sky, (0.2, 0.4, 0.6), rtol=1e-6)


def test_textures_uvs_and_tangents_follow_the_loader(stage):
    mesh, _, mat = _world_of(stage, "TexturedTriangle")
    # st is the primary uv set, unflipped: uv_i == (x_i, y_i) of the points
    np.testing.assert_array_equal(mesh["uvs"], mesh["positions"][:, :2])
    base = stage.scene.textures[mat["base_color_texture"]]
    emis = stage.scene.textures[mat["emissive_texture"]]
    assert base[1] == spk.TEXTURE_RGBA8_SRGB and emis[1] == spk.TEXTURE_RGBA8_UNORM
    assert base[0].shape == (1, 1, 4) and emis[0].shape == (1, 1, 4)
    assert mat["emissive"] == (1.0, 1.0, 1.0) and mat["metallic"] == 0.0 and mat["roughness"] == 1.0
    # fallback tangents (n x a) x n with a = GetPossibleMajorAxis(n): n = +Z has |x| < 1/sqrt 3 -> a = Right
    np.testing.assert_allclose(mesh["tangents"], np.tile([[1.0, 0.0, 0.0, 1.0]], (3, 1)), atol=1e-7)
    plain, _, pmat = _world_of(stage, "UntexturedTriangle")
    assert plain["uvs"] is None and plain["tangents"].shape == (3, 4)
    assert pmat["base_color"] == (0.5, 0.5, 0.5) and "base_color_texture" not in pmat


def test_transforms_compose_first_op_outermost(stage):
    # Root: translate (1, 2, 3) then rotateZ 90 (points rotate first): x -> (1, 2, 3) + (0, 1, 0)
    _, M, _ = _world_of(stage, "TexturedTriangle")
    p = M @ np.array([1.0, 0.0, 0.0, 1.0])
    np.testing.assert_allclose(p, (1.0, 3.0, 3.0), atol=1e-6)
    # Child: row-vector matrix (scale 2, translate (10, 0, 0)) under Root
    _, Mc, _ = _world_of(stage, "Quad")
    q = Mc @ np.array([1.0, 0.0, 0.0, 1.0])
    np.testing.assert_allclose(q, Mc[:, :3] @ [1, 0, 0] + Mc[:, 3], atol=1e-6)
    np.testing.assert_allclose(Mc[:, 3], (1.0, 2.0, 3.0) + np.array([0.0, 10.0, 0.0]), atol=1e-5)
    np.testing.assert_allclose(np.linalg.norm(Mc[:, :3], axis=0), (2.0, 2.0, 2.0), atol=1e-6)


def test_polygons_fan_and_face_varying_normals_split(stage):
    quad, _, glass = _world_of(stage, "Quad")
    assert quad["positions"].shape == (4, 3) and quad["indices"].tolist() == [[0, 1, 2], [0, 2, 3]]
    assert quad["normals"][:, 2].tolist() == [1.0, 1.0, 1.0, -1.0]
    assert glass["type"] == spk.MATERIAL_DIELECTRIC and math.isclose(glass["eta"], 1.33, rel_tol=1e-6)
    assert glass["base_color"] == tuple(float(np.float32(0.18)) for _ in range(3))  # UsdPreviewSurface default
    assert "Unbound" not in stage.mesh_names and any("Unbound" in s for s in stage.skipped)


def test_camera_attributes_and_basis(stage):
    c = stage.camera
    assert math.isclose(c["focal_length"], 0.05) and math.isclose(c["sensor_height"], 0.024)
    assert math.isclose(c["aperture"], 2.8, rel_tol=1e-6) and math.isclose(c["focus_distance"], 5.0)
    # CameraNode's translate composes with Root's translate + rotateZ 90
    np.testing.assert_allclose(c["position"], (1.0 + 5.0, 2.0, 4.0), atol=1e-6)
    cd = stage.camera_desc(aspect=1.5)
    fov = 2 * math.atan(0.024 / (2 * 0.05))
    height = 2 * math.tan(fov / 2) * 5.0
    assert math.isclose(np.linalg.norm(cd.max_v), height, rel_tol=1e-5)
    assert math.isclose(np.linalg.norm(cd.max_u), 1.5 * height, rel_tol=1e-5)
    assert math.isclose(cd.lens_radius, 0.05 / 2.8 * 0.5, rel_tol=1e-5)


def test_parser_rejects_bad_input(tmp_path):
    px = np.array([[[64, 128, 191, 255]]], np.uint8)
    (tmp_path / "srgb.png").write_bytes(_png_rgba(px))
    (tmp_path / "linear.png").write_bytes(_png_rgba(px))
    p = tmp_path / "bad.usda"
    p.write_text("not usd")
    with pytest.raises(usd.UsdError):
        usd.read_usda(str(p))
    p.write_text(STAGE.replace("</Root/Materials/PlainMaterial>", "</Root/Materials/Missing>"))
    with pytest.raises(usd.UsdError):
        usd.read_usda(str(p))


def _glass_cornell():
    """The textured Cornell room with a glass box. A textured channel carries no factor through USD
    (the exporter writes only the connection and the loader keeps MaterialResource's default), so the
    factors of textured channels are set to those defaults here: white base colour, metallic and
    roughness 1, emissive factor 1."""
    sc = scenes.textured_cornell_scene(seed=1)
    glass = sc.add_material(base_color=(1, 1, 1), eta=1.5, type=spk.MATERIAL_DIELECTRIC)
    sc.instances[-1]["material"] = glass
    for m in sc.materials:
        if m.get("base_color_texture", -1) >= 0:
            m["base_color"] = (1.0, 1.0, 1.0)
        if m.get("metallic_roughness_texture", -1) >= 0:
            m["metallic"] = m["roughness"] = 1.0
        if m.get("emissive_texture", -1) >= 0:
            m["emissive"] = (1.0, 1.0, 1.0)
    return sc


def _same_scene(a, b):
    assert len(a.instances) == len(b.instances)
    for ia, ib in zip(a.instances, b.instances):
        np.testing.assert_array_equal(ia["M"], ib["M"])
        ma, mb = a.meshes[ia["mesh"]], b.meshes[ib["mesh"]]
        for key in ("positions", "indices", "normals", "tangents", "uvs"):
            if ma[key] is None:
                assert mb[key] is None or key == "tangents"
            else:
                np.testing.assert_array_equal(ma[key], mb[key])
        pa, pb = a.materials[ia["material"]], b.materials[ib["material"]]
        for key in ("type", "base_color", "emissive", "metallic", "roughness", "eta"):
            assert np.array_equal(np.float32(pa[key]), np.float32(pb[key])), key
        for key in ("base_color_texture", "metallic_roughness_texture", "emissive_texture", "normal_texture"):
            ka, kb = pa.get(key, -1), pb.get(key, -1)
            assert (ka < 0) == (kb < 0), key
            if ka >= 0:
                assert a.textures[ka][1] == b.textures[kb][1]
                want = a.textures[ka][0]
                if a.textures[ka][1] == spk.TEXTURE_RGBA32F:  # .hdr: RGBE precision, alpha 1
                    want = usd.hdr_decode(usd.hdr_encode(want))
                np.testing.assert_array_equal(want, b.textures[kb][0])


def test_write_read_round_trip_is_exact(tmp_path):
    sc = _glass_cornell()
    cam = dict(position=np.array([0.1, -2.9, 1.0]), right=np.array([1.0, 0.0, 0.0]), front=np.array([0.0, 1.0, 0.0]),
               up=np.array([0.0, 0.0, 1.0]), focal_length=0.035, sensor_height=0.024, aperture=8.0, focus_distance=3.0)
    path = str(tmp_path / "out" / "scene.usda")
    os.makedirs(os.path.dirname(path))
    usd.write_usda(path, sc, camera=cam, sky=(0.3, 0.5, 0.9))
    back = usd.read_usda(path)
    _same_scene(sc, back.scene)
    assert back.sky == tuple(float(np.float32(x)) for x in (0.3, 0.5, 0.9))
    for k in ("position", "right", "front", "up"):
        np.testing.assert_allclose(back.camera[k], cam[k], atol=1e-7)
    for k in ("focal_length", "sensor_height", "aperture", "focus_distance"):
        assert math.isclose(back.camera[k], cam[k], rel_tol=1e-6), k


def test_grey_png_expands_to_rgba():
    import struct
    import zlib

    from sparkle_amd.gltf import _png_decode

    grey = np.array([[0, 128, 255]], np.uint8)
    raw = b"".join(b"\x00" + row.tobytes() for row in grey)

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    png = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", 3, 1, 8, 0, 0, 0, 0)) +
           chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))
    out = _png_decode(png)
    assert out.shape == (1, 3, 4)
    np.testing.assert_array_equal(out[0, :, :3], np.repeat(grey.T, 3, axis=1))
    assert (out[..., 3] == 255).all()


def test_rgbe_codec():
    """Radiance RGBE: values that are exact in 8 mantissa bits survive; others land within one
    step of the shared exponent; zero and alpha follow the format; RLE scanlines decode too."""
    v = np.array([[[1.0, 0.5, 0.25, 0.3], [3.0, 0.0, 1.5, 1.0], [0.0, 0.0, 0.0, 0.0]]], np.float32)
    out = usd.hdr_decode(usd.hdr_encode(v))
    np.testing.assert_array_equal(out[..., :3], v[..., :3])
    assert (out[..., 3] == 1.0).all()
    rng = np.random.default_rng(0)
    r = rng.uniform(0, 50, (5, 9, 4)).astype(np.float32)
    back = usd.hdr_decode(usd.hdr_encode(r))
    step = np.ldexp(1.0, np.frexp(r[..., :3].max(-1))[1] - 8)[..., None]
    assert (np.abs(back[..., :3] - r[..., :3]) <= step).all()
    # a run-length scanline: width 8, one run of 8 per channel
    head = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 8\n"
    body = bytes([2, 2, 0, 8]) + bytes([128 + 8, 128, 128 + 8, 64, 128 + 8, 0, 128 + 8, 129])  # R x 2^(129-136)
    rle = usd.hdr_decode(head + body)
    np.testing.assert_array_equal(rle[0, :, :3], np.tile([[1.0, 0.5, 0.0]], (8, 1)))


def test_round_trip_renders_the_same_image_on_the_oracle(tmp_path):
    sc = _glass_cornell()
    # RGBE-representable float texels, so the file round trip is exact
    for k, (px, fmt) in enumerate(sc.textures):
        if fmt == spk.TEXTURE_RGBA32F:
            sc.textures[k] = (usd.hdr_decode(usd.hdr_encode(px)), fmt)
    path = str(tmp_path / "scene.usda")
    usd.write_usda(path, sc)
    back = usd.read_usda(path).scene
    w = h = 24
    cam = scenes.cornell_camera(w, h)
    imgs = []
    for s in (sc, back):
        o = refcpu.OracleScene(s)
        rgba, _ = o.render_frame(cam, w, h, frame_seed=3, max_bounce=4)
        imgs.append(rgba)
    np.testing.assert_array_equal(imgs[0], imgs[1])
    assert np.isfinite(imgs[0]).all() and imgs[0][..., :3].sum() > 0
