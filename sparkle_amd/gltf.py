"""glTF 2.0 (ASCII, embedded buffers) writer and reader for `sparkle_amd.Scene`.

The reader applies the mapping rules of Sparkle's loader (libraries/source/io/scene/GLTFLoader.cpp):
  * every node is rotated +90 degrees about X after its own rotation (y-up files, z-up renderer,
    `:536-538`), translation kept, node scale baked into the vertex positions (`:190-212, :541-545`),
    children composed with their parent's transform (`:564-571`);
  * only the TRS fields of a node are read (no `matrix`); cameras, skins and animations are ignored;
  * material factors are read only when the matching texture exists (`:485-517`): an untextured
    material is white, fully metallic, roughness 1, not emissive (the `MaterialResource` defaults,
    io/Material.h:9-23); base = texture x factor, metallic = texture.b x factor, roughness =
    texture.g x factor, emissive = texture x factor (MaterialRenderProxy.h:117-162);
  * every material is a PBR material (the loader creates no dielectric);
  * base-colour and emissive textures are colour data (sRGB decode after filtering), metallic-
    roughness and normal textures are linear (TextureCompression profiles Color / Data / Normal,
    `:485-517`).
Images are embedded PNGs (8-bit RGB or RGBA, no interlace) decoded by this module.

The writer stores each instance as its own mesh in world space, rotated by -90 degrees about X so
the reader's +90 degrees restores it. A material channel without a texture gets a 1x1 white
texture so that the loader applies its factor; RGBA8 textures of the scene are written as PNG.
"""
import base64
import json
import math
import struct
import zlib

import numpy as np

from . import MATERIAL_PBR, TEXTURE_RGBA8_SRGB, TEXTURE_RGBA8_UNORM, Scene

_FLOAT, _UINT = 5126, 5125
_ARRAY_BUFFER, _ELEMENT_ARRAY_BUFFER = 34962, 34963


def _png_rgba(pixels):
    """Minimal PNG encoder (8-bit RGBA, no interlace) for small constant textures."""
    h, w = pixels.shape[:2]
    raw = b"".join(b"\x00" + pixels[j].astype(np.uint8).tobytes() for j in range(h))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def _png_decode(data):
    """Decodes an 8-bit grey, RGB or RGBA PNG without interlace into RGBA: grey replicates into RGB
    and a missing alpha is 255 (the ExpandToRgba rule of USDLoader.cpp:126-163)."""
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG image")
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        if tag == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, ctype, _, _, interlace = hdr
    if depth != 8 or ctype not in (0, 2, 6) or interlace:
        raise NotImplementedError("only 8-bit grey/RGB/RGBA PNG without interlace")
    ch = {0: 1, 2: 3, 6: 4}[ctype]
    raw = zlib.decompress(idat)
    out = np.zeros((h, w, ch), np.uint8)
    prev = np.zeros(w * ch, np.int32)
    stride = w * ch
    for j in range(h):
        f = raw[j * (stride + 1)]
        line = np.frombuffer(raw, np.uint8, stride, j * (stride + 1) + 1).astype(np.int32)
        cur = np.zeros(stride, np.int32)
        for i in range(stride):
            a = cur[i - ch] if i >= ch else 0
            b = prev[i]
            c = prev[i - ch] if i >= ch else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) // 2
            else:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            cur[i] = (line[i] + p) & 0xFF
        out[j] = cur.reshape(w, ch)
        prev = cur
    if ch == 1:
        out = np.repeat(out, 3, axis=-1)
    if ch in (1, 3):
        out = np.concatenate([out, np.full((h, w, 1), 255, np.uint8)], -1)
    return out


_RX_NEG90 = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, -1.0, 0.0]])


def write_gltf(scene, path):
    """Writes `scene` (a sparkle_amd.Scene) as an ASCII .gltf with one embedded buffer."""
    blob = bytearray()
    views, accessors, meshes, nodes = [], [], [], []

    def add_view(data, target):
        while len(blob) % 4:
            blob.append(0)
        views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": len(data), "target": target})
        blob.extend(data)
        return len(views) - 1

    def add_accessor(arr, kind, comp, target, minmax=False):
        v = add_view(np.ascontiguousarray(arr).tobytes(), target)
        acc = {"bufferView": v, "componentType": comp, "count": int(arr.shape[0]), "type": kind}
        if minmax:
            acc["min"] = [float(x) for x in arr.min(0)]
            acc["max"] = [float(x) for x in arr.max(0)]
        accessors.append(acc)
        return len(accessors) - 1

    for inst in scene.instances:
        m = scene.meshes[inst["mesh"]]
        M = inst["M"].astype(np.float64)
        p = m["positions"].astype(np.float64) @ M[:, :3].T + M[:, 3]
        p = (p @ _RX_NEG90.T).astype(np.float32)
        attrs = {"POSITION": add_accessor(p, "VEC3", _FLOAT, _ARRAY_BUFFER, True)}
        if m["normals"] is not None:
            n = m["normals"].astype(np.float64) @ M[:, :3].T
            n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
            attrs["NORMAL"] = add_accessor((n @ _RX_NEG90.T).astype(np.float32), "VEC3", _FLOAT, _ARRAY_BUFFER)
        if m["tangents"] is not None:
            t = m["tangents"].astype(np.float64).copy()
            t[:, :3] = (t[:, :3] @ M[:, :3].T) @ _RX_NEG90.T
            attrs["TANGENT"] = add_accessor(t.astype(np.float32), "VEC4", _FLOAT, _ARRAY_BUFFER)
        if m["uvs"] is not None:
            attrs["TEXCOORD_0"] = add_accessor(m["uvs"].astype(np.float32), "VEC2", _FLOAT, _ARRAY_BUFFER)
        idx = add_accessor(m["indices"].astype(np.uint32).reshape(-1), "SCALAR", _UINT, _ELEMENT_ARRAY_BUFFER)
        meshes.append({"primitives": [{"attributes": attrs, "indices": idx, "material": int(inst["material"]),
                                       "mode": 4}]})
        nodes.append({"mesh": len(meshes) - 1})

    images = ["data:image/png;base64," + base64.b64encode(_png_rgba(np.full((1, 1, 4), 255, np.uint8))).decode()]
    image_of = {}

    def tex_ref(k):
        if k is None or k < 0:
            return {"index": 0}
        if k not in image_of:
            px, fmt = scene.textures[k]
            if px.dtype != np.uint8:
                raise NotImplementedError("only 8-bit textures are written")
            images.append("data:image/png;base64," + base64.b64encode(_png_rgba(px)).decode())
            image_of[k] = len(images) - 1
        return {"index": image_of[k]}

    materials = []
    for mat in scene.materials:
        m = {
            "pbrMetallicRoughness": {
                "baseColorFactor": [float(x) for x in mat["base_color"]] + [1.0],
                "baseColorTexture": tex_ref(mat.get("base_color_texture")),
                "metallicFactor": float(mat["metallic"]),
                "roughnessFactor": float(mat["roughness"]),
                "metallicRoughnessTexture": tex_ref(mat.get("metallic_roughness_texture")),
            },
            "emissiveFactor": [float(x) for x in mat["emissive"]],
            "emissiveTexture": tex_ref(mat.get("emissive_texture")),
        }
        if mat.get("normal_texture", -1) >= 0:
            m["normalTexture"] = tex_ref(mat["normal_texture"])
        materials.append(m)
    doc = {
        "asset": {"version": "2.0", "generator": "sparkle_amd.gltf"},
        "scene": 0,
        "scenes": [{"nodes": list(range(len(nodes)))}],
        "nodes": nodes,
        "meshes": meshes,
        "materials": materials,
        "textures": [{"source": i} for i in range(len(images))],
        "images": [{"uri": u} for u in images],
        "buffers": [{"byteLength": len(blob),
                     "uri": "data:application/octet-stream;base64," + base64.b64encode(bytes(blob)).decode()}],
        "bufferViews": views,
        "accessors": accessors,
    }
    with open(path, "w") as f:
        json.dump(doc, f)


_COMP = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4}


def _data_uri(uri):
    if not uri.startswith("data:"):
        raise NotImplementedError("only embedded (data URI) buffers and images are read")
    return base64.b64decode(uri.split(",", 1)[1])


def _quat_matrix(q):
    x, y, z, w = (float(c) for c in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _rx90():
    a = float(np.float32(math.radians(90.0)))  # ToRadian(90.0f) in float
    c, s = math.cos(a), math.sin(a)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])


def read_gltf(path):
    """Reads an ASCII .gltf with embedded buffers into a Scene with the reference loader's rules."""
    doc = json.load(open(path))
    buffers = [_data_uri(b["uri"]) for b in doc.get("buffers", [])]

    def accessor(i):
        a = doc["accessors"][i]
        v = doc["bufferViews"][a["bufferView"]]
        dt = np.dtype(_COMP[a["componentType"]])
        nc = _NCOMP[a["type"]]
        stride = v.get("byteStride", 0) or dt.itemsize * nc
        base = v.get("byteOffset", 0) + a.get("byteOffset", 0)
        buf = buffers[v["buffer"]]
        if stride == dt.itemsize * nc:
            out = np.frombuffer(buf, dt, a["count"] * nc, base).reshape(a["count"], nc)
        else:
            out = np.stack([np.frombuffer(buf, dt, nc, base + k * stride) for k in range(a["count"])])
        return out.copy()

    images = [_png_decode(_data_uri(im["uri"])) for im in doc.get("images", [])]
    sc = Scene()
    made = {}

    def texture(ref, fmt):
        key = (doc["textures"][ref["index"]]["source"], fmt)
        if key not in made:
            made[key] = sc.add_texture(images[key[0]], fmt)
        return made[key]

    for m in doc.get("materials", []):
        # MaterialResource defaults (io/Material.h:9-23); factors only with their texture
        kw = dict(base_color=(1.0, 1.0, 1.0), emissive=(0.0, 0.0, 0.0), metallic=1.0, roughness=1.0)
        pbr = m.get("pbrMetallicRoughness", {})
        if "baseColorTexture" in pbr:
            kw["base_color_texture"] = texture(pbr["baseColorTexture"], TEXTURE_RGBA8_SRGB)
            kw["base_color"] = tuple(float(np.float32(x)) for x in pbr.get("baseColorFactor", [1, 1, 1, 1])[:3])
        if "normalTexture" in m:
            kw["normal_texture"] = texture(m["normalTexture"], TEXTURE_RGBA8_UNORM)
        if "metallicRoughnessTexture" in pbr:
            kw["metallic_roughness_texture"] = texture(pbr["metallicRoughnessTexture"], TEXTURE_RGBA8_UNORM)
            kw["metallic"] = float(np.float32(pbr.get("metallicFactor", 1.0)))
            kw["roughness"] = float(np.float32(pbr.get("roughnessFactor", 1.0)))
        if "emissiveTexture" in m:
            kw["emissive_texture"] = texture(m["emissiveTexture"], TEXTURE_RGBA8_SRGB)
            kw["emissive"] = tuple(float(np.float32(x)) for x in m.get("emissiveFactor", [0, 0, 0]))
        sc.add_material(type=MATERIAL_PBR, **kw)
    if not sc.materials:
        sc.add_material(metallic=1.0, roughness=1.0)

    R90 = _rx90()

    def visit(ni, parent):
        node = doc["nodes"][ni]
        R = R90 @ (_quat_matrix(node["rotation"]) if "rotation" in node else np.eye(3))
        T = np.asarray(node.get("translation", [0, 0, 0]), np.float64)
        S = np.asarray(node.get("scale", [1, 1, 1]), np.float64)
        local = np.eye(4)
        local[:3, :3] = R
        local[:3, 3] = T
        world = parent @ local
        if "mesh" in node:
            for prim in doc["meshes"][node["mesh"]]["primitives"]:
                if prim.get("mode", 4) != 4:
                    raise NotImplementedError("only triangle primitives")
                at = prim["attributes"]
                pos = accessor(at["POSITION"]).astype(np.float64) * S
                nrm = accessor(at["NORMAL"]) if "NORMAL" in at else None
                tan = accessor(at["TANGENT"]) if "TANGENT" in at else None
                uv = accessor(at["TEXCOORD_0"]) if "TEXCOORD_0" in at else None
                idx = (accessor(prim["indices"]).reshape(-1) if "indices" in prim
                       else np.arange(pos.shape[0], dtype=np.uint32))
                mi = sc.add_mesh(pos.astype(np.float32), idx.astype(np.uint32), nrm, tan, uv)
                sc.add_instance(mi, int(prim.get("material", 0)), world[:3, :].astype(np.float32))
        for c in node.get("children", []):
            visit(c, world)

    roots = doc["scenes"][doc.get("scene", 0)]["nodes"] if doc.get("scenes") else range(len(doc["nodes"]))
    for ni in roots:
        visit(ni, np.eye(4))
    return sc
