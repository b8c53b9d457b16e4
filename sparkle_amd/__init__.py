"""sparkle_amd — MI355X-native hot path of Sparkle's `cpu` path tracer.

Python mirror of the C ABI in include/spk_pt.h (ctypes). The compute runs in the HIP library
sparkle_amd/lib/libspk_pt.so; there is no CPU fallback: importing the package on a machine without
the built library raises, and every call that needs a GPU fails with the library's error code.

Reference interfaces mirrored (paths relative to the Sparkle repository):
    Context.upload_scene   SceneRenderProxy::UpdateBVH (SceneRenderProxy.cpp:269-297)
    Context.trace          SceneRenderProxy::Intersect<false> (SceneRenderProxy.cpp:299-309)
    Context.render_frame   CPURenderer::BasePass (CPURenderer.cpp:419-450)
    Context.accumulate     CPURenderer::DenoisePass, temporal part (CPURenderer.cpp:506-552)
"""
import ctypes
import os

import math

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libspk_pt.so")

SPK_OK = 0
SPK_ERR_INVALID_ARGUMENT = -1
SPK_ERR_HIP = -2
SPK_ERR_OUT_OF_MEMORY = -3
SPK_ERR_NO_SCENE = -4
SPK_ERR_INVALID_SCENE = -5
SPK_ERR_STACK_DEPTH = -6
SPK_ERR_INVALID_RAY = -7
SPK_ERR_NO_DEVICE = -8
SPK_ERR_UNSUPPORTED = -9
SPK_MISS = 0xFFFFFFFF

MATERIAL_PBR = 0
MATERIAL_DIELECTRIC = 1

DEBUG_COLOR = 0
DEBUG_DEBUG = 1
DEBUG_RAY_DEPTH = 2
DEBUG_NORMAL = 3
DEBUG_RAY_DIRECTION = 4
DEBUG_INDIRECT_LIGHTING = 5
DEBUG_DIRECTIONAL_LIGHTING = 6
DEBUG_METALLIC = 7
DEBUG_ROUGHNESS = 8
DEBUG_ALBEDO = 9
DEBUG_EMISSIVE = 10
DEBUG_DEPTH = 11

NOALLOC_MAX_RAYS = 2097152  # SPK_NOALLOC_MAX_RAYS
SKY_FACE_SIZE = 1024  # the reference cooks 1024^2 cube faces (SkyLight.cpp:24)
FRAME_SEED64 = 1
FRAME_NEE = 2

RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<u4"), ("t", "<f4")])


class SpkError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"spk error {code}: {message}")
        self.code = code


class MeshDesc(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_void_p), ("num_vertices", ctypes.c_uint32),
                ("indices", ctypes.c_void_p), ("num_triangles", ctypes.c_uint32),
                ("normals", ctypes.c_void_p), ("tangents", ctypes.c_void_p), ("uvs", ctypes.c_void_p)]


class InstanceDesc(ctypes.Structure):
    _fields_ = [("mesh", ctypes.c_uint32), ("material", ctypes.c_uint32), ("world_from_local", ctypes.c_float * 12)]


class MaterialDesc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_uint32), ("base_color", ctypes.c_float * 3), ("emissive", ctypes.c_float * 3),
                ("metallic", ctypes.c_float), ("roughness", ctypes.c_float), ("eta", ctypes.c_float),
                ("base_color_texture", ctypes.c_uint32), ("metallic_roughness_texture", ctypes.c_uint32),
                ("emissive_texture", ctypes.c_uint32), ("normal_texture", ctypes.c_uint32)]  # 0 none, k+1


TEXTURE_RGBA8_UNORM = 0
TEXTURE_RGBA8_SRGB = 1
TEXTURE_RGBA32F = 2


class SphereDesc(ctypes.Structure):
    _fields_ = [("center", ctypes.c_float * 3), ("radius", ctypes.c_float), ("material", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 3)]


class TextureDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("format", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("pixels", ctypes.c_void_p)]


class BuildOptions(ctypes.Structure):
    _fields_ = [("bvh_width", ctypes.c_uint32), ("max_leaf_tris", ctypes.c_uint32), ("num_threads", ctypes.c_uint32),
                ("two_level", ctypes.c_uint32)]


class SceneStats(ctypes.Structure):
    _fields_ = [("num_triangles", ctypes.c_uint64), ("num_nodes", ctypes.c_uint64), ("bvh_width", ctypes.c_uint32),
                ("depth", ctypes.c_uint32), ("node_bytes", ctypes.c_uint64), ("triangle_bytes", ctypes.c_uint64),
                ("device_bytes", ctypes.c_uint64), ("degenerate_tris", ctypes.c_uint64), ("sah_cost", ctypes.c_double),
                ("build_seconds", ctypes.c_double)]


class TraceConfig(ctypes.Structure):
    _fields_ = [("bvh_width", ctypes.c_uint32), ("stack", ctypes.c_uint32), ("block", ctypes.c_uint32),
                ("refill", ctypes.c_uint32), ("ray_sort", ctypes.c_uint32)]

    def as_tuple(self):
        return (self.bvh_width, self.stack, self.block, self.refill, self.ray_sort)


TUNE_WIDTH = 1


class CameraDesc(ctypes.Structure):
    _fields_ = [("position", ctypes.c_float * 3), ("right", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("front", ctypes.c_float * 3), ("lower_left", ctypes.c_float * 3), ("max_u", ctypes.c_float * 3),
                ("max_v", ctypes.c_float * 3), ("lens_radius", ctypes.c_float), ("far_plane", ctypes.c_float)]


class FrameDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("frame_seed", ctypes.c_uint32),
                ("max_bounce", ctypes.c_uint32), ("debug_mode", ctypes.c_uint32), ("sky_enabled", ctypes.c_uint32),
                ("sky_color", ctypes.c_float * 3), ("flags", ctypes.c_uint32), ("output_limit", ctypes.c_float)]


class CacheConfig(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_uint32), ("features", ctypes.c_uint32), ("log2_table", ctypes.c_uint32),
                ("base_resolution", ctypes.c_uint32), ("per_level_scale", ctypes.c_float),
                ("learning_rate", ctypes.c_float), ("batch", ctypes.c_uint32), ("terminate_bounce", ctypes.c_uint32),
                ("seed", ctypes.c_uint32), ("train_fraction", ctypes.c_float), ("train_suffix", ctypes.c_uint32)]


def cache_config(log2_table=19, base_resolution=16, per_level_scale=1.3819, learning_rate=1e-2, batch=1 << 16,
                 terminate_bounce=2, seed=0, train_fraction=0.0, train_suffix=2):
    """The configuration of the learned radiance cache: 16 levels x 2 features, FP16 table.
    train_suffix: training paths end in the cache train_suffix hit indices after terminate_bounce
    (self-training), one in 16 excepted; 0 traces every training path to its end."""
    return CacheConfig(16, 2, log2_table, base_resolution, per_level_scale, learning_rate, batch, terminate_bounce,
                       seed, train_fraction, train_suffix)


FRAME_CACHE = 4
CACHE_ATTRS = 8  # SPK_CACHE_ATTRS: shading normal xyz, albedo rgb, roughness, 1
CACHE_INPUTS = 64  # MLP input width: 32 encoding features, the attributes, zeros
CACHE_ALBEDO_FLOOR = 0.02
FRAME_NO_SORT = 0x80000000


class FrameStats(ctypes.Structure):
    _fields_ = [("rays_traced", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64), ("milliseconds", ctypes.c_double)]


class Schedule(ctypes.Structure):
    _fields_ = [("blocks_per_cu", ctypes.c_uint32), ("material_sort", ctypes.c_uint32), ("refill", ctypes.c_uint32),
                ("tail_bounce", ctypes.c_uint32)]

    def as_tuple(self):
        return (self.blocks_per_cu, self.material_sort, self.refill, self.tail_bounce)


class BatchScene(ctypes.Structure):
    _fields_ = [("meshes", ctypes.c_void_p), ("num_meshes", ctypes.c_uint32), ("instances", ctypes.c_void_p),
                ("num_instances", ctypes.c_uint32), ("materials", ctypes.c_void_p), ("num_materials", ctypes.c_uint32),
                ("camera", CameraDesc), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32)]


_lib = None


def load_library():
    """Loads libspk_pt.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"HIP extension not built: {LIB_PATH} is missing (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    P, U32, U64, I, F = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float
    sig = {
        "spk_create": (I, [I, ctypes.POINTER(P)]),
        "spk_destroy": (I, [P]),
        "spk_last_error": (ctypes.c_char_p, [P]),
        "spk_set_stream": (I, [P, P]),
        "spk_get_stream": (P, [P]),
        "spk_synchronize": (I, [P]),
        "spk_device_alloc": (I, [P, ctypes.c_size_t, ctypes.POINTER(P)]),
        "spk_device_free": (I, [P, P]),
        "spk_copy_to_device": (I, [P, P, P, ctypes.c_size_t]),
        "spk_copy_to_host": (I, [P, P, P, ctypes.c_size_t]),
        "spk_scene_upload": (I, [P, P, U32, P, U32, P, U32, P]),
        "spk_scene_stats_get": (I, [P, ctypes.POINTER(SceneStats)]),
        "spk_textures_upload": (I, [P, P, U32]),
        "spk_scene_set_spheres": (I, [P, P, U32]),
        "spk_set_sky_cubemap": (I, [P, P]),
        "spk_sky_cook": (I, [P, U32, U32, U32, P]),
        "spk_build_check": (I, [P, U32, P, U32, P, ctypes.POINTER(SceneStats), ctypes.c_char_p, ctypes.c_size_t]),
        "spk_trace_closest": (I, [P, P, U64, P]),
        "spk_trace_closest_async": (I, [P, P, U64, P]),
        "spk_trace_closest_host": (I, [P, P, U64, P]),
        "spk_trace_closest_noalloc": (I, [P, P, U64, P]),
        "spk_trace_count": (I, [P, P, U64, P, P]),
        "spk_trace_kernel_timing": (I, [P, U32]),
        "spk_trace_kernel_times": (I, [P, ctypes.POINTER(F), U32, ctypes.POINTER(U32)]),
        "spk_get_trace_config": (I, [P, ctypes.POINTER(TraceConfig)]),
        "spk_set_trace_config": (I, [P, ctypes.POINTER(TraceConfig)]),
        "spk_list_trace_configs": (I, [P, P, U32, ctypes.POINTER(U32)]),
        "spk_autotune": (I, [P, P, U64, U32, ctypes.POINTER(TraceConfig), ctypes.POINTER(ctypes.c_double)]),
        "spk_set_camera": (I, [P, ctypes.POINTER(CameraDesc)]),
        "spk_render_frame": (I, [P, ctypes.POINTER(FrameDesc), P, P, ctypes.POINTER(FrameStats)]),
        "spk_accumulate": (I, [P, P, P, U64, U32, U32]),
        "spk_tonemap": (I, [P, P, P, U32, U32, F]),
        "spk_spatial_fill": (I, [P, P, P, P, U32, U32]),
        "spk_batch_upload": (I, [P, P, U32, ctypes.POINTER(BuildOptions)]),
        "spk_scene_update_instances": (I, [P, ctypes.POINTER(F), U32]),
        "spk_batch_render": (I, [P, ctypes.POINTER(FrameDesc), P, ctypes.POINTER(FrameStats)]),
        "spk_batch_offsets": (I, [P, ctypes.POINTER(U64), U32]),
        "spk_batch_autotune": (I, [P, ctypes.POINTER(FrameDesc), P, ctypes.POINTER(Schedule),
                                   ctypes.POINTER(ctypes.c_double)]),
        "spk_render_autotune": (I, [P, ctypes.POINTER(FrameDesc), P, ctypes.POINTER(Schedule),
                                    ctypes.POINTER(ctypes.c_double)]),
        "spk_get_schedule": (I, [P, I, ctypes.POINTER(Schedule)]),
        "spk_set_schedule": (I, [P, I, ctypes.POINTER(Schedule)]),
        "spk_cache_create": (I, [P, ctypes.POINTER(CacheConfig)]),
        "spk_cache_destroy": (I, [P]),
        "spk_cache_reset": (I, [P]),
        "spk_cache_infer": (I, [P, P, P, U32, I, P]),
        "spk_cache_train": (I, [P, P, P, P, U32, I, ctypes.POINTER(F)]),
        "spk_cache_gradients": (I, [P, P, P, P, U32, I, ctypes.POINTER(F), P, P]),
        "spk_cache_get_params": (I, [P, P, P]),
        "spk_cache_set_params": (I, [P, P, P]),
        "spk_cache_info": (I, [P, ctypes.POINTER(U64), ctypes.POINTER(U64), ctypes.POINTER(U32), ctypes.POINTER(F)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return int(a.data_ptr())  # torch tensor


class Scene:
    """Host description of a scene: meshes (numpy arrays), instances and materials."""

    def __init__(self):
        self.meshes = []
        self.instances = []
        self.materials = []
        self.textures = []  # (pixels (H, W, 4), TEXTURE_* format)
        self.spheres = []   # (center (3,), radius, material): analytic spheres
        self.sky_faces = None  # six texture indices (+X, -X, +Y, -Y, +Z, -Z) of a sky cube map

    def add_mesh(self, positions, indices, normals=None, tangents=None, uvs=None):
        m = dict(positions=np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3),
                 indices=np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3),
                 normals=None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3),
                 tangents=None if tangents is None else np.ascontiguousarray(tangents, dtype=np.float32).reshape(-1, 4),
                 uvs=None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2))
        self.meshes.append(m)
        return len(self.meshes) - 1

    def add_material(self, base_color=(1, 1, 1), emissive=(0, 0, 0), metallic=1.0, roughness=1.0, eta=0.0,
                     type=MATERIAL_PBR, base_color_texture=-1, metallic_roughness_texture=-1, emissive_texture=-1,
                     normal_texture=-1):
        m = dict(type=type, base_color=tuple(base_color), emissive=tuple(emissive), metallic=float(metallic),
                 roughness=float(roughness), eta=float(eta))
        tex = dict(base_color_texture=int(base_color_texture),
                   metallic_roughness_texture=int(metallic_roughness_texture),
                   emissive_texture=int(emissive_texture), normal_texture=int(normal_texture))
        if any(v >= 0 for v in tex.values()):
            m.update(tex)
        self.materials.append(m)
        return len(self.materials) - 1

    def add_sphere(self, center, radius, material):
        self.spheres.append((tuple(float(c) for c in center), float(radius), int(material)))
        return len(self.spheres) - 1

    def add_texture(self, pixels, fmt=TEXTURE_RGBA8_UNORM):
        """pixels (H, W, 4): uint8 for the RGBA8 formats, float32 for TEXTURE_RGBA32F."""
        px = np.ascontiguousarray(pixels, dtype=np.float32 if fmt == TEXTURE_RGBA32F else np.uint8)
        assert px.ndim == 3 and px.shape[2] == 4
        self.textures.append((px, fmt))
        return len(self.textures) - 1

    def set_sky_map(self, equirect, face_size=SKY_FACE_SIZE):
        """A textured sky light: cooks the equirectangular map into six cube faces (sky_cook), adds
        them as RGBA32F textures and makes them the sky cube map. Returns the face texture indices."""
        faces = sky_cook(equirect, face_size)
        self.sky_faces = [self.add_texture(f, TEXTURE_RGBA32F) for f in faces]
        return self.sky_faces

    def add_instance(self, mesh, material, world_from_local=None):
        M = np.eye(3, 4, dtype=np.float32) if world_from_local is None else np.asarray(world_from_local,
                                                                                       dtype=np.float32)
        self.instances.append(dict(mesh=mesh, material=material, M=M.reshape(3, 4)))
        return len(self.instances) - 1

    @property
    def num_triangles(self):
        return sum(self.meshes[i["mesh"]]["indices"].shape[0] for i in self.instances)

    def c_arrays(self):
        ms = (MeshDesc * len(self.meshes))()
        for k, m in enumerate(self.meshes):
            ms[k].positions = _ptr(m["positions"])
            ms[k].num_vertices = m["positions"].shape[0]
            ms[k].indices = _ptr(m["indices"])
            ms[k].num_triangles = m["indices"].shape[0]
            ms[k].normals = _ptr(m["normals"])
            ms[k].tangents = _ptr(m["tangents"])
            ms[k].uvs = _ptr(m["uvs"])
        ins = (InstanceDesc * len(self.instances))()
        for k, i in enumerate(self.instances):
            ins[k].mesh = i["mesh"]
            ins[k].material = i["material"]
            ins[k].world_from_local[:] = [float(x) for x in i["M"].reshape(-1)]
        mats = (MaterialDesc * max(1, len(self.materials)))()
        for k, m in enumerate(self.materials):
            mats[k].type = m["type"]
            mats[k].base_color[:] = m["base_color"]
            mats[k].emissive[:] = m["emissive"]
            mats[k].metallic = m["metallic"]
            mats[k].roughness = m["roughness"]
            mats[k].eta = m["eta"]
            mats[k].base_color_texture = m.get("base_color_texture", -1) + 1
            mats[k].metallic_roughness_texture = m.get("metallic_roughness_texture", -1) + 1
            mats[k].emissive_texture = m.get("emissive_texture", -1) + 1
            mats[k].normal_texture = m.get("normal_texture", -1) + 1
        return ms, ins, mats


def build_check(scene, bvh_width=8, max_leaf_tris=3, num_threads=0):
    """Host-only BVH build + invariant check (no GPU). Returns (rc, stats, message)."""
    lib = load_library()
    ms, ins, _ = scene.c_arrays()
    opts = BuildOptions(bvh_width, max_leaf_tris, num_threads)
    st = SceneStats()
    msg = ctypes.create_string_buffer(512)
    rc = lib.spk_build_check(ms, len(scene.meshes), ins, len(scene.instances), ctypes.byref(opts),
                             ctypes.byref(st), msg, 512)
    return rc, st, msg.value.decode()


class Context:
    """One spk_context on a device; all arrays passed to device calls are torch CUDA tensors."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.spk_create(device, ctypes.byref(h))
        if rc != SPK_OK:
            raise SpkError(rc, "spk_create failed")
        self.h = h
        self.device = device
        self._keep = None

    def close(self):
        if self.h:
            self.lib.spk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != SPK_OK:
            raise SpkError(rc, self.lib.spk_last_error(self.h).decode())
        return rc

    def set_stream(self, stream_ptr):
        return self._check(self.lib.spk_set_stream(self.h, stream_ptr))

    def stream(self):
        return self.lib.spk_get_stream(self.h)

    def synchronize(self):
        return self._check(self.lib.spk_synchronize(self.h))

    def upload_textures(self, textures):
        """spk_textures_upload: textures is a list of (pixels, format) with pixels (H, W, 4) uint8 or
        float32."""
        arr = (TextureDesc * max(1, len(textures)))()
        keep = []
        for k, (px, fmt) in enumerate(textures):
            px = np.ascontiguousarray(px, dtype=np.float32 if fmt == TEXTURE_RGBA32F else np.uint8)
            assert px.ndim == 3 and px.shape[2] == 4
            keep.append(px)
            arr[k].width, arr[k].height, arr[k].format = px.shape[1], px.shape[0], fmt
            arr[k].pixels = px.ctypes.data
        return self._check(self.lib.spk_textures_upload(self.h, arr, len(textures)))

    def upload_scene(self, scene, bvh_width=8, max_leaf_tris=3, num_threads=0, two_level=False):
        """spk_textures_upload (the scene's textures, when it has any or the table is not empty)
        followed by spk_scene_upload; two_level keeps one BLAS per mesh under a TLAS of instances
        (then update_instances moves them)."""
        if scene.textures or getattr(self, "_has_textures", False):
            self.upload_textures(scene.textures)
            self._has_textures = bool(scene.textures)
        ms, ins, mats = scene.c_arrays()
        opts = BuildOptions(bvh_width, max_leaf_tris, num_threads, 1 if two_level else 0)
        self._check(self.lib.spk_scene_upload(self.h, ms, len(scene.meshes), ins, len(scene.instances), mats,
                                              len(scene.materials), ctypes.byref(opts)))
        if getattr(scene, "spheres", None):
            self.set_spheres(scene.spheres)
        self.set_sky_cubemap(getattr(scene, "sky_faces", None))
        return self.stats()

    def update_instances(self, transforms):
        """spk_scene_update_instances: transforms is (num_instances, 3, 4) (or 12 floats each),
        world_from_local of every instance in upload order."""
        m = np.ascontiguousarray(transforms, dtype=np.float32).reshape(-1, 12)
        return self._check(self.lib.spk_scene_update_instances(
            self.h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), m.shape[0]))

    def set_sky_cubemap(self, faces):
        """spk_set_sky_cubemap: six texture indices (+X, -X, +Y, -Y, +Z, -Z) or None."""
        if faces is None:
            return self._check(self.lib.spk_set_sky_cubemap(self.h, None))
        return self._check(self.lib.spk_set_sky_cubemap(self.h, (ctypes.c_uint32 * 6)(*[f + 1 for f in faces])))

    def set_spheres(self, spheres):
        """spk_scene_set_spheres: spheres is a list of (center, radius, material)."""
        arr = (SphereDesc * max(1, len(spheres)))()
        for k, (c, r, m) in enumerate(spheres):
            arr[k].center[:] = list(c)
            arr[k].radius = r
            arr[k].material = m
        return self._check(self.lib.spk_scene_set_spheres(self.h, arr, len(spheres)))

    def stats(self):
        st = SceneStats()
        self._check(self.lib.spk_scene_stats_get(self.h, ctypes.byref(st)))
        return st

    def trace(self, rays_dev, hits_dev, n=None, sync=True):
        """rays_dev: torch uint8/float tensor holding n 32-byte records; hits_dev: n 8-byte records."""
        if n is None:
            n = rays_dev.numel() * rays_dev.element_size() // 32
        fn = self.lib.spk_trace_closest if sync else self.lib.spk_trace_closest_async
        return fn(self.h, _ptr(rays_dev), n, _ptr(hits_dev))

    def trace_checked(self, rays_dev, hits_dev, n=None):
        return self._check(self.trace(rays_dev, hits_dev, n))

    def trace_noalloc(self, rays_dev, hits_dev, n=None):
        """spk_trace_closest_noalloc: the closest hit with the library's fixed scratch (no allocation),
        at most NOALLOC_MAX_RAYS rays. Returns the status code."""
        if n is None:
            n = rays_dev.numel() * rays_dev.element_size() // 32
        return self.lib.spk_trace_closest_noalloc(self.h, _ptr(rays_dev), n, _ptr(hits_dev))

    def trace_host(self, rays, hits=None):
        """rays: numpy RAY_DTYPE array; hits: optional numpy HIT_DTYPE array of the same length to
        write into (a caller that reuses it saves the page faults of a fresh array). Returns (rc,
        hits)."""
        rays = np.ascontiguousarray(rays)
        if hits is None:
            hits = np.zeros(rays.shape[0], dtype=HIT_DTYPE)
        elif hits.dtype != HIT_DTYPE or hits.shape != (rays.shape[0],) or not hits.flags.c_contiguous:
            raise ValueError("hits must be a contiguous HIT_DTYPE array with one record per ray")
        rc = self.lib.spk_trace_closest_host(self.h, rays.ctypes.data, rays.shape[0], hits.ctypes.data)
        return rc, hits

    def trace_count(self, rays_dev, nodes_dev, tris_dev, n=None):
        if n is None:
            n = rays_dev.numel() * rays_dev.element_size() // 32
        return self._check(self.lib.spk_trace_count(self.h, _ptr(rays_dev), n, _ptr(nodes_dev), _ptr(tris_dev)))

    def trace_kernel_timing(self, capacity):
        """Time the closest-hit kernel (not the coherence sort before it) of the next `capacity`
        trace calls with HIP events on the context stream; 0 switches timing off."""
        return self._check(self.lib.spk_trace_kernel_timing(self.h, capacity))

    def trace_kernel_times(self):
        """Kernel milliseconds of the calls timed since trace_kernel_timing, in call order."""
        cnt = ctypes.c_uint32(0)
        self._check(self.lib.spk_trace_kernel_times(self.h, None, 0, ctypes.byref(cnt)))
        out = (ctypes.c_float * max(cnt.value, 1))()
        self._check(self.lib.spk_trace_kernel_times(self.h, out, cnt.value, ctypes.byref(cnt)))
        return [float(out[k]) for k in range(cnt.value)]

    def trace_config(self):
        c = TraceConfig()
        self._check(self.lib.spk_get_trace_config(self.h, ctypes.byref(c)))
        return c.as_tuple()

    def set_trace_config(self, bvh_width, stack, block, refill=0, ray_sort=0):
        """refill 0: one ray per thread; k > 0: persistent lanes refilled once k lanes of a wave are idle;
        ray_sort 1: rays traced in a coherence order (counting sort by direction bin and origin cell)."""
        c = TraceConfig(bvh_width, stack, block, refill, ray_sort)
        return self._check(self.lib.spk_set_trace_config(self.h, ctypes.byref(c)))

    def list_trace_configs(self):
        arr = (TraceConfig * 256)()
        cnt = ctypes.c_uint32()
        self._check(self.lib.spk_list_trace_configs(self.h, arr, 256, ctypes.byref(cnt)))
        return [arr[i].as_tuple() for i in range(min(cnt.value, 256))]

    def autotune(self, rays_dev, n=None, tune_width=False):
        if n is None:
            n = rays_dev.numel() * rays_dev.element_size() // 32
        c = TraceConfig()
        ms = ctypes.c_double()
        self._check(self.lib.spk_autotune(self.h, _ptr(rays_dev), n, TUNE_WIDTH if tune_width else 0,
                                          ctypes.byref(c), ctypes.byref(ms)))
        return c.as_tuple(), ms.value

    def set_camera(self, cam):
        return self._check(self.lib.spk_set_camera(self.h, ctypes.byref(cam)))

    def render_frame(self, frame, rgba_dev, normal_dev=None):
        st = FrameStats()
        self._check(self.lib.spk_render_frame(self.h, ctypes.byref(frame), _ptr(rgba_dev), _ptr(normal_dev),
                                              ctypes.byref(st)))
        return st

    def upload_batch(self, items, bvh_width=8, max_leaf_tris=3, num_threads=0):
        """spk_batch_upload: items are (Scene, CameraDesc, width, height). Returns the pixel offsets."""
        arr = (BatchScene * len(items))()
        keep = []
        textures = []
        for k, (scene, cam, w, h) in enumerate(items):
            ms, ins, mats = scene.c_arrays()
            # one texture table for the batch: this scene's indices shift by the textures before it
            for j in range(len(scene.materials)):
                for f in ("base_color_texture", "metallic_roughness_texture", "emissive_texture", "normal_texture"):
                    if getattr(mats[j], f):
                        setattr(mats[j], f, getattr(mats[j], f) + len(textures))
            textures.extend(scene.textures)
            keep.append((ms, ins, mats))
            arr[k].meshes = ctypes.cast(ms, ctypes.c_void_p)
            arr[k].num_meshes = len(scene.meshes)
            arr[k].instances = ctypes.cast(ins, ctypes.c_void_p)
            arr[k].num_instances = len(scene.instances)
            arr[k].materials = ctypes.cast(mats, ctypes.c_void_p)
            arr[k].num_materials = len(scene.materials)
            arr[k].camera = cam
            arr[k].width = w
            arr[k].height = h
        if textures or getattr(self, "_has_textures", False):
            self.upload_textures(textures)
            self._has_textures = bool(textures)
        opts = BuildOptions(bvh_width, max_leaf_tris, num_threads)
        self._check(self.lib.spk_batch_upload(self.h, arr, len(items), ctypes.byref(opts)))
        offs = (ctypes.c_uint64 * (len(items) + 1))()
        self._check(self.lib.spk_batch_offsets(self.h, offs, len(items) + 1))
        return [int(x) for x in offs]

    def render_batch(self, frame, rgba_dev):
        st = FrameStats()
        self._check(self.lib.spk_batch_render(self.h, ctypes.byref(frame), _ptr(rgba_dev), ctypes.byref(st)))
        return st

    def autotune_batch(self, frame, rgba_dev):
        """spk_batch_autotune: returns ((blocks_per_cu, material_sort, refill, tail_bounce), best frame ms)."""
        sc, ms = Schedule(), ctypes.c_double()
        self._check(self.lib.spk_batch_autotune(self.h, ctypes.byref(frame), _ptr(rgba_dev), ctypes.byref(sc),
                                                ctypes.byref(ms)))
        return sc.as_tuple(), ms.value

    def autotune_render(self, frame, rgba_dev):
        """spk_render_autotune: returns ((blocks_per_cu, material_sort, refill, tail_bounce), best frame ms)."""
        sc, ms = Schedule(), ctypes.c_double()
        self._check(self.lib.spk_render_autotune(self.h, ctypes.byref(frame), _ptr(rgba_dev), ctypes.byref(sc),
                                                 ctypes.byref(ms)))
        return sc.as_tuple(), ms.value

    def schedule(self, batch=False):
        sc = Schedule()
        self._check(self.lib.spk_get_schedule(self.h, 1 if batch else 0, ctypes.byref(sc)))
        return sc.as_tuple()

    def set_schedule(self, blocks_per_cu, material_sort, refill, tail_bounce=0, batch=False):
        """spk_set_schedule; tail_bounce > 0 runs the segments from that index on one path per lane."""
        sc = Schedule(blocks_per_cu, material_sort, refill, tail_bounce)
        return self._check(self.lib.spk_set_schedule(self.h, 1 if batch else 0, ctypes.byref(sc)))

    def accumulate(self, rgba_dev, accum_dev, num_pixels, cumulated, spp=1):
        return self._check(self.lib.spk_accumulate(self.h, _ptr(rgba_dev), _ptr(accum_dev), num_pixels,
                                                   cumulated, spp))

    def spatial_fill(self, rgba_dev, normal_dev, out_dev, width, height):
        return self._check(self.lib.spk_spatial_fill(self.h, _ptr(rgba_dev), _ptr(normal_dev), _ptr(out_dev),
                                                     width, height))

    def tonemap(self, accum_dev, out_dev, width, height, exposure=1.0):
        return self._check(self.lib.spk_tonemap(self.h, _ptr(accum_dev), _ptr(out_dev), width, height, exposure))

    # ---- learned radiance cache ----
    def cache_create(self, cfg=None):
        cfg = cfg or cache_config()
        return self._check(self.lib.spk_cache_create(self.h, ctypes.byref(cfg)))

    def cache_reset(self):
        return self._check(self.lib.spk_cache_reset(self.h))

    def cache_destroy(self):
        return self._check(self.lib.spk_cache_destroy(self.h))

    def cache_info(self):
        nw, nt, st, loss = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_float()
        self._check(self.lib.spk_cache_info(self.h, ctypes.byref(nw), ctypes.byref(nt), ctypes.byref(st),
                                            ctypes.byref(loss)))
        return dict(num_weights=nw.value, num_table_params=nt.value, steps=st.value, last_loss=loss.value)

    def cache_infer(self, points_dev, out_dev, n, world=False, attrs_dev=None):
        """attrs_dev: CACHE_ATTRS floats per query (normal, albedo, roughness, 1) or None (zeros)."""
        return self._check(self.lib.spk_cache_infer(self.h, _ptr(points_dev), _ptr(attrs_dev), n, 1 if world else 0,
                                                    _ptr(out_dev)))

    def cache_train(self, points_dev, targets_dev, n, world=False, attrs_dev=None):
        loss = ctypes.c_float()
        self._check(self.lib.spk_cache_train(self.h, _ptr(points_dev), _ptr(attrs_dev), _ptr(targets_dev), n,
                                             1 if world else 0, ctypes.byref(loss)))
        return loss.value

    def cache_gradients(self, points_dev, targets_dev, n, world=False, table=False, attrs_dev=None):
        info = self.cache_info()
        dw = np.zeros(info["num_weights"], dtype=np.float32)
        dt = np.zeros(info["num_table_params"], dtype=np.float32) if table else None
        loss = ctypes.c_float()
        self._check(self.lib.spk_cache_gradients(self.h, _ptr(points_dev), _ptr(attrs_dev), _ptr(targets_dev), n,
                                                 1 if world else 0, ctypes.byref(loss), _ptr(dw), _ptr(dt)))
        return loss.value, dw, dt

    def cache_get_params(self):
        info = self.cache_info()
        w = np.zeros(info["num_weights"], dtype=np.float32)
        t = np.zeros(info["num_table_params"], dtype=np.float32)
        self._check(self.lib.spk_cache_get_params(self.h, _ptr(w), _ptr(t)))
        return w, t

    def cache_set_params(self, weights=None, table=None):
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        t = None if table is None else np.ascontiguousarray(table, dtype=np.float32)
        return self._check(self.lib.spk_cache_set_params(self.h, _ptr(w), _ptr(t)))


_HALF_MAX = 65504.0


def sky_source_texels(image):
    """The texels a sky map is sampled from, as the reference's image loader keeps them
    (Image2D::LoadFromFile, io/Image.cpp:150-197): an 8-bit image as c / 255 (R8G8B8A8Unorm), a float
    (.hdr) image clamped to the FP16 range and rounded to FP16 (RGBAFloat16). (H, W, 4) float32."""
    a = np.asarray(image)
    if a.ndim != 3 or a.shape[2] not in (3, 4):
        raise ValueError("sky map must be (H, W, 3 or 4)")
    if a.dtype == np.uint8:
        out = a.astype(np.float32) / np.float32(255.0)
    else:
        out = np.clip(a.astype(np.float32), -_HALF_MAX, _HALF_MAX).astype(np.float16).astype(np.float32)
    if out.shape[2] == 3:
        out = np.concatenate([out, np.ones(out.shape[:2] + (1,), np.float32)], 2)
    return np.ascontiguousarray(out)


def sky_cook(image, face_size=SKY_FACE_SIZE):
    """spk_sky_cook on an equirectangular sky map (uint8 or float, (H, W, 3 or 4)): six (face_size,
    face_size, 4) float32 faces +X, -X, +Y, -Y, +Z, -Z with FP16-rounded values."""
    src = sky_source_texels(image)
    h, w = src.shape[:2]
    faces = np.empty((6, face_size, face_size, 4), np.float32)
    rc = load_library().spk_sky_cook(src.ctypes.data, w, h, face_size, faces.ctypes.data)
    if rc != SPK_OK:
        raise SpkError(rc, f"spk_sky_cook failed for a {w}x{h} map and {face_size}^2 faces")
    return list(faces)


def camera_desc(position, front, up_hint=(0.0, 0.0, 1.0), vertical_fov=None, focus_distance=1.0,
                aperture_radius=0.0, aspect=1.0, far_plane=1000.0):
    """Builds the camera record SetupViewRay consumes (CameraRenderProxy.cpp:24-27, 113-136) from a
    position, a viewing direction and the physical attributes, in float32 in the reference's order."""
    f32 = np.float32
    P = np.asarray(position, dtype=f32)
    F = np.asarray(front, dtype=np.float64)
    F = (F / np.linalg.norm(F)).astype(f32)
    R = np.cross(F.astype(np.float64), np.asarray(up_hint, dtype=np.float64))
    R = (R / np.linalg.norm(R)).astype(f32)
    U = np.cross(R.astype(np.float64), F.astype(np.float64))
    U = (U / np.linalg.norm(U)).astype(f32)
    if vertical_fov is None:
        # CameraComponent defaults: 35 mm focal length, 24 mm sensor height (CameraComponent.h:18-22)
        vertical_fov = 2.0 * np.arctan(0.024 / (2.0 * 0.035))
    h = f32(np.tan(f32(vertical_fov) / f32(2.0)))
    height = f32(2.0) * h * f32(focus_distance)
    width = f32(aspect) * height
    mu = R * width
    mv = U * height
    ll = P + F * f32(focus_distance) - mu * f32(0.5) - mv * f32(0.5)
    c = CameraDesc()
    c.position[:] = [float(x) for x in P]
    c.right[:] = [float(x) for x in R]
    c.up[:] = [float(x) for x in U]
    c.front[:] = [float(x) for x in F]
    c.lower_left[:] = [float(x) for x in ll]
    c.max_u[:] = [float(x) for x in mu]
    c.max_v[:] = [float(x) for x in mv]
    c.lens_radius = float(aperture_radius)
    c.far_plane = float(far_plane)
    return c


def camera_from_basis(position, right, front, up, focal_length, sensor_height, aperture, focus_distance, aspect,
                      far_plane=1000.0):
    """The render camera of a camera node (CameraComponent::UpdateRenderData and the focus plane of
    CameraRenderProxy.cpp:24-27, 113-136) from its position, its (right, front, up) basis and its
    physical attributes (metres, f-number), in float32."""
    f32 = np.float32
    fov = f32(2.0) * f32(math.atan(f32(sensor_height) / (f32(2.0) * f32(focal_length))))
    lens_radius = f32(focal_length) / f32(aperture) * f32(0.5)
    h = f32(math.tan(fov / f32(2.0)))
    height = f32(2.0) * h * f32(focus_distance)
    width = f32(aspect) * height
    P, R, F, U = (np.asarray(v, f32) for v in (position, right, front, up))
    mu = R * width
    mv = U * height
    ll = P + F * f32(focus_distance) - mu * f32(0.5) - mv * f32(0.5)
    cd = CameraDesc()
    cd.position[:] = [float(x) for x in P]
    cd.right[:] = [float(x) for x in R]
    cd.up[:] = [float(x) for x in U]
    cd.front[:] = [float(x) for x in F]
    cd.lower_left[:] = [float(x) for x in ll]
    cd.max_u[:] = [float(x) for x in mu]
    cd.max_v[:] = [float(x) for x in mv]
    cd.lens_radius = float(lens_radius)
    cd.far_plane = float(far_plane)
    return cd


def orbit_basis(center, radius, pitch, yaw):
    """Position and (right, front, up) of an orbit camera (OrbitCameraComponent::Setup and
    UpdateTransform, OrbitCameraComponent.cpp:16-30): the node rotation is
    AngleAxis(pitch, Right) * AngleAxis(yaw, Up) (EulerRotationToRotationAxis of (pitch, 0, yaw) in
    radians, Utilities.h:392-396, with Right = +X, Front = +Y, Up = +Z), the camera basis is its rows
    (Transform::ExtractLocalBasis), and the camera sits `radius` behind `center` along its front.
    Angles in degrees; positive pitch looks down."""
    p, y = math.radians(pitch), math.radians(yaw)
    cp, sp, cy, sy = math.cos(p), math.sin(p), math.cos(y), math.sin(y)
    rx = np.array([[1.0, 0.0, 0.0], [0.0, cp, -sp], [0.0, sp, cp]])
    rz = np.array([[cy, -sy, 0.0], [sy, cy, 0.0], [0.0, 0.0, 1.0]])
    rot = rx @ rz
    right, front, up = rot[0], rot[1], rot[2]
    position = np.asarray(center, np.float64) - float(radius) * front
    return position, right, front, up


def orbit_camera(center=(0.0, 0.0, 0.0), radius=10.0, pitch=0.0, yaw=0.0, aspect=1.0, focal_length=0.035,
                 sensor_height=0.024, aperture=22.0, far_plane=1000.0):
    """Render camera of an orbit camera focused on its centre (SetFocusDistance(radius)), with the
    CameraComponent defaults (35 mm, 24 mm sensor, f/22; CameraComponent.h:18-22). The defaults of
    this function are the camera every loaded scene starts with (SceneManager.cpp:142-147:
    Setup(Zeros, 10, 0, 0): at (0, -10, 0) looking along +Y)."""
    position, right, front, up = orbit_basis(center, radius, pitch, yaw)
    return camera_from_basis(position, right, front, up, focal_length, sensor_height, aperture, radius, aspect,
                             far_plane)
