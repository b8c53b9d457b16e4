"""ctypes wrapper of liborcl.so (refcpu.cpp): the CPU restatement of the reference path.

Test infrastructure only (see refcpu.h for what each function restates and how it is pinned).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborcl.so")

TRACE_REF_MT = 0
TRACE_WT = 1
TRACE_WT_BRUTE = 2
TRACE_FP64 = 3
TRACE_WT_TWO = 4
FLAG_EDGE, FLAG_TIE, FLAG_GRAZE = 1, 2, 4

RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<u4"), ("t", "<f4")])


class OMesh(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_void_p), ("num_vertices", ctypes.c_uint32), ("indices", ctypes.c_void_p),
                ("num_triangles", ctypes.c_uint32), ("normals", ctypes.c_void_p), ("tangents", ctypes.c_void_p),
                ("uvs", ctypes.c_void_p)]


class OInstance(ctypes.Structure):
    _fields_ = [("mesh", ctypes.c_uint32), ("material", ctypes.c_uint32), ("world_from_local", ctypes.c_float * 12),
                ("local_from_world", ctypes.c_float * 12)]


class OMaterial(ctypes.Structure):
    _fields_ = [("type", ctypes.c_uint32), ("base_color", ctypes.c_float * 3), ("emissive", ctypes.c_float * 3),
                ("metallic", ctypes.c_float), ("roughness", ctypes.c_float), ("eta", ctypes.c_float),
                ("base_color_texture", ctypes.c_int32), ("metallic_roughness_texture", ctypes.c_int32),
                ("emissive_texture", ctypes.c_int32), ("normal_texture", ctypes.c_int32)]


class OSphere(ctypes.Structure):
    _fields_ = [("center", ctypes.c_float * 3), ("radius", ctypes.c_float), ("material", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 3)]


class OTexture(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("format", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("pixels", ctypes.c_void_p)]


class OCamera(ctypes.Structure):
    _fields_ = [("position", ctypes.c_float * 3), ("right", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("front", ctypes.c_float * 3), ("lower_left", ctypes.c_float * 3), ("max_u", ctypes.c_float * 3),
                ("max_v", ctypes.c_float * 3), ("lens_radius", ctypes.c_float), ("far_plane", ctypes.c_float)]


class OFrame(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("frame_seed", ctypes.c_uint32),
                ("max_bounce", ctypes.c_uint32), ("debug_mode", ctypes.c_uint32), ("sky_enabled", ctypes.c_uint32),
                ("sky_color", ctypes.c_float * 3), ("seed_hi", ctypes.c_uint32), ("nee", ctypes.c_uint32),
                ("output_limit", ctypes.c_float), ("rr_off", ctypes.c_uint32)]


class OCacheConfig(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_uint32), ("features", ctypes.c_uint32), ("log2_table", ctypes.c_uint32),
                ("base_resolution", ctypes.c_uint32), ("per_level_scale", ctypes.c_float),
                ("hidden", ctypes.c_uint32), ("hidden_layers", ctypes.c_uint32), ("outputs", ctypes.c_uint32),
                ("inputs", ctypes.c_uint32)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P, U32, U64, I, F, D = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float,
                                ctypes.c_double)
        sig = {
            "orcl_scene_create": (P, [P, U32, P, U32, P, U32]),
            "orcl_scene_set_textures": (None, [P, P, U32]),
            "orcl_scene_set_spheres": (None, [P, P, U32]),
            "orcl_scene_set_sky_faces": (None, [P, P]),
            "orcl_texture_sample": (None, [P, ctypes.c_int, ctypes.c_float, ctypes.c_float, P]),
            "orcl_scene_destroy": (None, [P]),
            "orcl_scene_num_triangles": (U64, [P]),
            "orcl_trace": (I, [P, P, U64, I, I, P, P, P]),
            "orcl_set_fp64_eps": (None, [D, D, D]),
            "orcl_tri_mt": (I, [P, P, P, P, P, P, P, P]),
            "orcl_tri_wt": (I, [P, P, F, P, P, P, P, P, P]),
            "orcl_refine_t": (F, [P, P, P, P, P]),
            "orcl_xoshiro_stream": (None, [U64, P, U32]),
            "orcl_xoshiro_advance": (None, [P, U64, P]),
            "orcl_sky_cook": (None, [P, U32, U32, U32, P]),
            "orcl_half_round": (F, [F]),
            "orcl_random_unit_stream": (None, [U64, P, U32]),
            "orcl_bxdf_sample": (None, [I, P, P, F, F, F, U64, U32, P, P, P]),
            "orcl_local_axes": (None, [P, P, P, P, P]),
            "orcl_render_frame": (I, [P, P, P, I, P, P]),
            "orcl_accumulate": (None, [P, P, U64, U32, U32]),
            "orcl_spatial_fill": (None, [P, P, P, U32, U32]),
            "orcl_last_render_rays": (U64, []),
            "orcl_camera_setup": (None, [P, P, P, P, F, F, F, F, P]),
            "orcl_hashgrid_encode": (None, [P, P, P, U32, P]),
            "orcl_hashgrid_index": (U32, [P, U32, U32, U32, U32]),
            "orcl_mlp_loss_grad": (D, [P, P, P, P, U32, P, P, P]),
            "orcl_adam": (None, [P, P, P, P, U64, D, D, D, D, U32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _inverse34(M):
    M = np.asarray(M, dtype=np.float64).reshape(3, 4)
    A = M[:, :3]
    Ai = np.linalg.inv(A)
    return np.concatenate([Ai, -(Ai @ M[:, 3])[:, None]], axis=1)


class OracleScene:
    """Two-level scene of the oracle built from a sparkle_amd.Scene-like description
    (dicts with numpy arrays; the arrays are copied by the library)."""

    def __init__(self, scene):
        L = lib()
        self._meshes = scene.meshes
        ms = (OMesh * len(scene.meshes))()
        for k, m in enumerate(scene.meshes):
            ms[k].positions = _p(m["positions"])
            ms[k].num_vertices = m["positions"].shape[0]
            ms[k].indices = _p(m["indices"])
            ms[k].num_triangles = m["indices"].shape[0]
            ms[k].normals = _p(m["normals"])
            ms[k].tangents = _p(m["tangents"])
            ms[k].uvs = _p(m["uvs"])
        ins = (OInstance * len(scene.instances))()
        for k, i in enumerate(scene.instances):
            ins[k].mesh = i["mesh"]
            ins[k].material = i["material"]
            ins[k].world_from_local[:] = [float(x) for x in np.asarray(i["M"], np.float32).reshape(-1)]
            ins[k].local_from_world[:] = [float(x) for x in _inverse34(i["M"]).astype(np.float32).reshape(-1)]
        mats = (OMaterial * max(1, len(scene.materials)))()
        for k, m in enumerate(scene.materials):
            mats[k].type = m["type"]
            mats[k].base_color[:] = m["base_color"]
            mats[k].emissive[:] = m["emissive"]
            mats[k].metallic = m["metallic"]
            mats[k].roughness = m["roughness"]
            mats[k].eta = m["eta"]
            mats[k].base_color_texture = m.get("base_color_texture", -1)
            mats[k].metallic_roughness_texture = m.get("metallic_roughness_texture", -1)
            mats[k].emissive_texture = m.get("emissive_texture", -1)
            mats[k].normal_texture = m.get("normal_texture", -1)
        self.h = L.orcl_scene_create(ms, len(scene.meshes), ins, len(scene.instances), mats, len(scene.materials))
        if not self.h:
            raise RuntimeError("orcl_scene_create failed")
        sph = getattr(scene, "spheres", [])
        if sph:
            arr = (OSphere * len(sph))()
            for k, (c, r, m) in enumerate(sph):
                arr[k].center[:] = list(c)
                arr[k].radius = r
                arr[k].material = m
            L.orcl_scene_set_spheres(self.h, arr, len(sph))
        texs = getattr(scene, "textures", [])
        if texs:
            arr = (OTexture * len(texs))()
            keep = []
            for k, (px, fmt) in enumerate(texs):
                px = np.ascontiguousarray(px, dtype=np.float32 if fmt == 2 else np.uint8)
                keep.append(px)
                arr[k].width, arr[k].height, arr[k].format = px.shape[1], px.shape[0], fmt
                arr[k].pixels = px.ctypes.data
            L.orcl_scene_set_textures(self.h, arr, len(texs))
        faces = getattr(scene, "sky_faces", None)
        if faces is not None:
            L.orcl_scene_set_sky_faces(self.h, (ctypes.c_int32 * 6)(*faces))

    def texture_sample(self, index, u, v):
        out = np.zeros(3, np.float32)
        lib().orcl_texture_sample(self.h, index, float(u), float(v), out.ctypes.data)
        return out

    def __del__(self):
        try:
            if self.h:
                lib().orcl_scene_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def trace(self, rays, mode=TRACE_WT, threads=None):
        rays = np.ascontiguousarray(rays)
        n = rays.shape[0]
        hits = np.zeros(n, dtype=HIT_DTYPE)
        flags = np.zeros(n, dtype=np.uint32)
        t64 = np.zeros(n, dtype=np.float64)
        threads = threads or os.cpu_count() or 1
        rc = lib().orcl_trace(self.h, rays.ctypes.data, n, mode, threads, hits.ctypes.data, flags.ctypes.data,
                              t64.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"orcl_trace failed: {rc}")
        if mode == TRACE_FP64:
            return hits, flags, t64
        return hits

    def render_frame(self, cam, width, height, frame_seed=0, max_bounce=4, debug_mode=0, sky=None, seed64=False,
                     threads=None, nee=False, output_limit=0.0, roulette=True):
        """One frame of BasePass. roulette=False is a checker switch with no reference counterpart: paths
        survive every roulette draw unweighted (the draw is still taken)."""
        fr = OFrame(width, height, frame_seed, max_bounce, debug_mode, 1 if sky is not None else 0,
                    (ctypes.c_float * 3)(*(sky if sky is not None else (0, 0, 0))), 1 if seed64 else 0,
                    1 if nee else 0, output_limit, 0 if roulette else 1)
        oc = OCamera()
        for f in ("position", "right", "up", "front", "lower_left", "max_u", "max_v"):
            getattr(oc, f)[:] = list(getattr(cam, f))
        oc.lens_radius = cam.lens_radius
        oc.far_plane = cam.far_plane
        rgba = np.zeros((height, width, 4), dtype=np.float32)
        nrm = np.zeros((height, width, 3), dtype=np.float32)
        rc = lib().orcl_render_frame(self.h, ctypes.byref(oc), ctypes.byref(fr), threads or os.cpu_count() or 1,
                                     rgba.ctypes.data, nrm.ctypes.data)
        if rc != 0:
            raise RuntimeError("orcl_render_frame failed")
        return rgba, nrm

    @staticmethod
    def last_render_rays():
        return int(lib().orcl_last_render_rays())


def accumulate(rgba, accum, cumulated, spp=1):
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    assert accum.flags["C_CONTIGUOUS"] and accum.dtype == np.float32
    lib().orcl_accumulate(rgba.ctypes.data, accum.ctypes.data, rgba.size // 4, cumulated, spp)


def spatial_fill(rgba, normal):
    """CPURenderer's 8-neighbour fill of invalid pixels (single-threaded order). rgba (H, W, 4),
    normal (H, W, 3), float32; returns the filled copy."""
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    normal = np.ascontiguousarray(normal, dtype=np.float32)
    h, w = rgba.shape[:2]
    out = np.empty_like(rgba)
    lib().orcl_spatial_fill(rgba.ctypes.data, normal.ctypes.data, out.ctypes.data, w, h)
    return out


def tri_mt(o, d, v0, v1, v2):
    f = lambda x: np.ascontiguousarray(x, dtype=np.float32)
    o, d, v0, v1, v2 = map(f, (o, d, v0, v1, v2))
    t, u, v = (np.zeros(1, np.float32) for _ in range(3))
    hit = lib().orcl_tri_mt(o.ctypes.data, d.ctypes.data, v0.ctypes.data, v1.ctypes.data, v2.ctypes.data,
                            t.ctypes.data, u.ctypes.data, v.ctypes.data)
    return bool(hit), float(t[0]), float(u[0]), float(v[0])


def tri_wt(o, d, v0, v1, v2, tmin=0.0):
    f = lambda x: np.ascontiguousarray(x, dtype=np.float32)
    o, d, v0, v1, v2 = map(f, (o, d, v0, v1, v2))
    t, u, v = (np.zeros(1, np.float32) for _ in range(3))
    hit = lib().orcl_tri_wt(o.ctypes.data, d.ctypes.data, float(tmin), v0.ctypes.data, v1.ctypes.data,
                            v2.ctypes.data, t.ctypes.data, u.ctypes.data, v.ctypes.data)
    return bool(hit), float(t[0]), float(u[0]), float(v[0])


def refine_t(o, d, v0, v1, v2):
    f = lambda x: np.ascontiguousarray(x, dtype=np.float32)
    o, d, v0, v1, v2 = map(f, (o, d, v0, v1, v2))
    return float(lib().orcl_refine_t(o.ctypes.data, d.ctypes.data, v0.ctypes.data, v1.ctypes.data, v2.ctypes.data))


def xoshiro_stream(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().orcl_xoshiro_stream(seed, out.ctypes.data, n)
    return out


def xoshiro_advance(state, n=1, outputs=False):
    """Advances a raw 4-word xoshiro128+ state n steps; returns the new state (and the outputs)."""
    st = np.ascontiguousarray(state, dtype=np.uint32).copy()
    out = np.zeros(n, dtype=np.uint32) if outputs else None
    lib().orcl_xoshiro_advance(st.ctypes.data, n, None if out is None else out.ctypes.data)
    return (st, out) if outputs else st


def sky_cook(equirect, face_size):
    """The oracle's textured-sky cook: (H, W, 4) float32 texels -> (6, face_size, face_size, 4)."""
    src = np.ascontiguousarray(equirect, dtype=np.float32)
    h, w = src.shape[:2]
    faces = np.zeros((6, face_size, face_size, 4), np.float32)
    lib().orcl_sky_cook(src.ctypes.data, w, h, face_size, faces.ctypes.data)
    return faces


def half_round(x):
    return float(lib().orcl_half_round(float(x)))


def random_unit_stream(seed, n):
    out = np.zeros(n, dtype=np.float32)
    lib().orcl_random_unit_stream(seed, out.ctypes.data, n)
    return out


BXDF_LAMBERT, BXDF_SPECULAR, BXDF_DIELECTRIC, BXDF_VNDF = 0, 1, 2, 3


def bxdf_sample(kind, wo, n, seed=1, base=(1.0, 1.0, 1.0), roughness=0.5, metallic=0.0, eta=1.5):
    """n samples of one lobe in the local frame: (wi [n, 3], throughput [n, 3], valid [n] bool)."""
    wo = np.ascontiguousarray(wo, dtype=np.float32)
    base = np.ascontiguousarray(base, dtype=np.float32)
    wi = np.zeros((n, 3), np.float32)
    thr = np.zeros((n, 3), np.float32)
    valid = np.zeros(n, np.int32)
    lib().orcl_bxdf_sample(kind, wo.ctypes.data, base.ctypes.data, roughness, metallic, eta, seed, n, wi.ctypes.data,
                           thr.ctypes.data, valid.ctypes.data)
    return wi, thr, valid.astype(bool)


def local_axes(normal, tangent):
    normal = np.ascontiguousarray(normal, dtype=np.float32)
    tangent = np.ascontiguousarray(tangent, dtype=np.float32)
    u, v, w = (np.zeros(3, np.float32) for _ in range(3))
    lib().orcl_local_axes(normal.ctypes.data, tangent.ctypes.data, u.ctypes.data, v.ctypes.data, w.ctypes.data)
    return u, v, w


# ---- learned radiance cache (FP64) ----------------------------------------------------------
def cache_cfg(log2_table=19, base_resolution=16, per_level_scale=1.3819, inputs=0):
    return OCacheConfig(16, 2, log2_table, base_resolution, per_level_scale, 64, 4, 3, inputs)


def level_resolutions(cfg):
    import math

    return [int(math.floor(cfg.base_resolution * cfg.per_level_scale ** l)) for l in range(cfg.levels)]


def hashgrid_index(cfg, level, x, y, z):
    return int(lib().orcl_hashgrid_index(ctypes.byref(cfg), level, x, y, z))


def hashgrid_encode(cfg, table, points):
    """table: float32 [levels * 2^log2_table * 2]; points [n, 3] in [0,1]; returns [n, 32] float64."""
    table = np.ascontiguousarray(table, dtype=np.float32)
    points = np.ascontiguousarray(points, dtype=np.float32)
    n = points.shape[0]
    out = np.zeros((n, cfg.levels * cfg.features), dtype=np.float64)
    lib().orcl_hashgrid_encode(ctypes.byref(cfg), table.ctypes.data, points.ctypes.data, n, out.ctypes.data)
    return out


def mlp_loss_grad(cfg, weights, features, targets):
    """Relative-L2 loss, outputs, dW and d(features) of the bias-free ReLU MLP in FP64."""
    w = np.ascontiguousarray(weights, dtype=np.float64)
    x = np.ascontiguousarray(features, dtype=np.float64)
    t = np.ascontiguousarray(targets, dtype=np.float64)
    n = x.shape[0]
    out = np.zeros((n, cfg.outputs), dtype=np.float64)
    dw = np.zeros_like(w)
    dx = np.zeros_like(x)
    loss = lib().orcl_mlp_loss_grad(ctypes.byref(cfg), w.ctypes.data, x.ctypes.data, t.ctypes.data, n,
                                    out.ctypes.data, dw.ctypes.data, dx.ctypes.data)
    return loss, out, dw, dx


def adam(param, grad, m, v, lr, beta1, beta2, eps, step):
    for a in (param, m, v):
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    g = np.ascontiguousarray(grad, dtype=np.float64)
    lib().orcl_adam(param.ctypes.data, g.ctypes.data, m.ctypes.data, v.ctypes.data, param.size, lr, beta1, beta2,
                    eps, step)


def table_backward(cfg, points, dfeatures):
    """Scatter d(features) into a dense FP64 table gradient (trilinear weights), the adjoint of
    hashgrid_encode. Pure numpy over the oracle's index function; for small n only."""
    import math

    T = 1 << cfg.log2_table
    g = np.zeros(cfg.levels * T * 2, dtype=np.float64)
    res = level_resolutions(cfg)
    for i in range(points.shape[0]):
        for l in range(cfg.levels):
            r = res[l]
            cell, frac = [], []
            for a in range(3):
                p = np.float32(min(max(points[i, a], 0.0), 1.0))
                pos = np.float32(p * np.float32(r))
                c = np.float32(math.floor(pos))
                c = min(max(c, np.float32(0)), np.float32(r - 1))
                cell.append(int(c))
                frac.append(float(np.float32(pos - c)))
            for k in range(8):
                cx, cy, cz = cell[0] + (k & 1), cell[1] + ((k >> 1) & 1), cell[2] + ((k >> 2) & 1)
                w = ((frac[0] if k & 1 else 1 - frac[0]) * (frac[1] if (k >> 1) & 1 else 1 - frac[1]) *
                     (frac[2] if (k >> 2) & 1 else 1 - frac[2]))
                idx = hashgrid_index(cfg, l, cx, cy, cz)
                g[(l * T + idx) * 2] += w * dfeatures[i, 2 * l]
                g[(l * T + idx) * 2 + 1] += w * dfeatures[i, 2 * l + 1]
    return g


def to_bf16(x):
    """Round float32 values to bfloat16 (nearest even) and back to float32."""
    a = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)
