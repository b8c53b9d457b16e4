"""Seeded synthetic workloads: the procedural heightfield with its coherent and incoherent ray sets,
the Cornell-style room, and the many-instance scene. All numbers are generated in float64 with
numpy's PCG64 and rounded once to float32, so every run and every machine builds identical inputs.
"""
import math

import numpy as np

from . import (MATERIAL_DIELECTRIC, MATERIAL_PBR, RAY_DTYPE, TEXTURE_RGBA8_SRGB, TEXTURE_RGBA8_UNORM, TEXTURE_RGBA32F,
               Scene, camera_desc)


# ---------------------------------------------------------------------------------------------
# heightfield: 1024 x 1024 cells over the unit square, 6-octave value noise, amplitude 0.2
# ---------------------------------------------------------------------------------------------
def value_noise(n_vert, octaves=6, base_freq=4, seed=0):
    """Value noise on an n_vert x n_vert grid over [0,1]^2: lattice values U[0,1), smoothstep
    interpolation, lacunarity 2, gain 0.5, normalised to [0, 1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.linspace(0.0, 1.0, n_vert)
    total = np.zeros((n_vert, n_vert))
    amp = 1.0
    for o in range(octaves):
        f = base_freq * (2 ** o)
        lattice = rng.random((f + 1, f + 1))
        g = x * f
        i0 = np.minimum(np.floor(g).astype(np.int64), f - 1)
        t = g - i0
        s = t * t * (3.0 - 2.0 * t)
        # rows: y, cols: x
        v00 = lattice[np.ix_(i0, i0)]
        v10 = lattice[np.ix_(i0, i0 + 1)]
        v01 = lattice[np.ix_(i0 + 1, i0)]
        v11 = lattice[np.ix_(i0 + 1, i0 + 1)]
        sx = s[None, :]
        sy = s[:, None]
        total += amp * ((v00 * (1 - sx) + v10 * sx) * (1 - sy) + (v01 * (1 - sx) + v11 * sx) * sy)
        amp *= 0.5
    return (total - total.min()) / (total.max() - total.min())


def heightfield_mesh(cells=1024, amplitude=0.2, seed=0):
    """Vertices (i/cells, j/cells, amplitude*N); two triangles per cell, fixed diagonal;
    triangle id = 2*(j*cells + i) + k."""
    nv = cells + 1
    h = value_noise(nv, seed=seed) * amplitude
    jj, ii = np.meshgrid(np.arange(nv), np.arange(nv), indexing="ij")
    pos = np.stack([ii / cells, jj / cells, h], axis=-1).reshape(-1, 3)
    # normals from central differences, tangents along +x
    dzdx = np.gradient(h, 1.0 / cells, axis=1)
    dzdy = np.gradient(h, 1.0 / cells, axis=0)
    nrm = np.stack([-dzdx, -dzdy, np.ones_like(h)], axis=-1).reshape(-1, 3)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    tan = np.stack([np.ones_like(h), np.zeros_like(h), dzdx], axis=-1).reshape(-1, 3)
    tan /= np.linalg.norm(tan, axis=1, keepdims=True)
    tan = np.concatenate([tan, np.ones((tan.shape[0], 1))], axis=1)
    uv = pos[:, :2].copy()
    ci, cj = np.meshgrid(np.arange(cells), np.arange(cells), indexing="xy")
    a = (cj * nv + ci).reshape(-1)
    b = a + 1
    c = a + nv + 1
    d = a + nv
    idx = np.empty((cells * cells * 2, 3), dtype=np.uint32)
    idx[0::2] = np.stack([a, b, c], axis=1)
    idx[1::2] = np.stack([a, c, d], axis=1)
    return (pos.astype(np.float32), idx, nrm.astype(np.float32), tan.astype(np.float32), uv.astype(np.float32))


def heightfield_scene(cells=1024, amplitude=0.2, seed=0):
    pos, idx, nrm, tan, uv = heightfield_mesh(cells, amplitude, seed)
    sc = Scene()
    m = sc.add_mesh(pos, idx, nrm, tan, uv)
    mat = sc.add_material(base_color=(0.6, 0.6, 0.6), metallic=0.0, roughness=1.0)
    sc.add_instance(m, mat)
    return sc


def scene_aabb(scene):
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    for ins in scene.instances:
        p = scene.meshes[ins["mesh"]]["positions"].astype(np.float64)
        M = ins["M"].astype(np.float64)
        w = p @ M[:, :3].T + M[:, 3]
        lo = np.minimum(lo, w.min(0))
        hi = np.maximum(hi, w.max(0))
    return lo, hi


def incoherent_rays(aabb_lo, aabb_hi, n=8 * 1024 * 1024, seed=0):
    """Origins uniform in the AABB, directions uniform on the sphere (z = 1 - 2u, phi = 2*pi*v),
    tmin = 0, tmax = +inf."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo = np.asarray(aabb_lo, dtype=np.float64)
    ext = np.asarray(aabb_hi, dtype=np.float64) - lo
    rays = np.empty(n, dtype=RAY_DTYPE)
    chunk = 1 << 22
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        m = e - s
        o = lo + rng.random((m, 3)) * ext
        u = rng.random(m)
        v = rng.random(m)
        z = 1.0 - 2.0 * u
        r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        phi = 2.0 * math.pi * v
        d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
        rays["org"][s:e] = o.astype(np.float32)
        rays["dir"][s:e] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    return rays


def coherent_rays(width=1920, height=1080, vfov_deg=60.0, pitch_deg=35.0, target=(0.5, 0.5, 0.1), distance=1.35):
    """Pixel-centre pinhole rays from a camera pitched down at the heightfield centre."""
    p = math.radians(pitch_deg)
    front = np.array([0.0, math.cos(p), -math.sin(p)])
    eye = np.asarray(target, dtype=np.float64) - distance * front
    right = np.array([1.0, 0.0, 0.0])
    up = np.cross(right, front)
    th = math.tan(math.radians(vfov_deg) / 2.0)
    aspect = width / height
    i = (np.arange(width) + 0.5) / width
    j = (np.arange(height) + 0.5) / height
    jj, ii = np.meshgrid(j, i, indexing="ij")
    d = (front[None, None, :] + ((2 * ii - 1) * th * aspect)[..., None] * right + ((1 - 2 * jj) * th)[..., None] * up)
    d = d.reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.empty(width * height, dtype=RAY_DTYPE)
    rays["org"] = eye.astype(np.float32)
    rays["dir"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    return rays


# ---------------------------------------------------------------------------------------------
# meshes for the room scenes
# ---------------------------------------------------------------------------------------------
def _quad(p0, p1, p2, p3, normal, tangent):
    pos = np.array([p0, p1, p2, p3], dtype=np.float64)
    idx = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)
    nrm = np.tile(np.asarray(normal, dtype=np.float64), (4, 1))
    tan = np.tile(np.concatenate([np.asarray(tangent, dtype=np.float64), [1.0]]), (4, 1))
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float64)
    return pos, idx, nrm, tan, uv


def box_mesh(center, half, angle_z):
    """Axis-aligned box rotated about z, 12 triangles with flat normals."""
    c, s = math.cos(angle_z), math.sin(angle_z)
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
    faces = []
    hx, hy, hz = half
    corners = lambda sx, sy, sz: np.array([sx * hx, sy * hy, sz * hz])
    # +x, -x, +y, -y, +z, -z
    defs = [((1, 0, 0), (0, 1, 0)), ((-1, 0, 0), (0, -1, 0)), ((0, 1, 0), (-1, 0, 0)), ((0, -1, 0), (1, 0, 0)),
            ((0, 0, 1), (1, 0, 0)), ((0, 0, -1), (1, 0, 0))]
    pos_all, idx_all, nrm_all, tan_all, uv_all = [], [], [], [], []
    for n, t in defs:
        n = np.array(n, dtype=np.float64)
        t = np.array(t, dtype=np.float64)
        b = np.cross(n, t)
        ext = np.array([hx, hy, hz])
        ctr = n * ext
        ut = t * ext
        vb = b * ext
        quad = [ctr - ut - vb, ctr + ut - vb, ctr + ut + vb, ctr - ut + vb]
        base = len(pos_all) * 4
        p, i, nn, tt, uv = _quad(*quad, normal=n, tangent=t)
        pos_all.append(p)
        idx_all.append(i + 4 * len(idx_all))
        nrm_all.append(nn)
        tan_all.append(tt)
        uv_all.append(uv)
    pos = np.concatenate(pos_all) @ R.T + np.asarray(center)
    nrm = np.concatenate(nrm_all) @ R.T
    tan = np.concatenate(tan_all)
    tan[:, :3] = tan[:, :3] @ R.T
    return pos, np.concatenate(idx_all), nrm, tan, np.concatenate(uv_all)


def icosphere(subdiv, radius=1.0, center=(0, 0, 0)):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in verts]
    f = faces
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[key] = len(v) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    V = np.array(v)
    nrm = V.copy()
    # tangent: d/dphi direction, fall back to +x at the poles
    tan = np.cross(np.array([0.0, 0.0, 1.0]), nrm)
    ln = np.linalg.norm(tan, axis=1, keepdims=True)
    tan = np.where(ln > 1e-6, tan / np.maximum(ln, 1e-12), np.array([1.0, 0.0, 0.0]))
    tan = np.concatenate([tan, np.ones((tan.shape[0], 1))], axis=1)
    uv = np.stack([np.arctan2(V[:, 1], V[:, 0]) / (2 * math.pi) + 0.5, np.arccos(np.clip(V[:, 2], -1, 1)) / math.pi],
                  axis=1)
    return V * radius + np.asarray(center), np.array(f, dtype=np.uint32), nrm, tan, uv


def torus(R, r, n_major, n_minor, center=(0, 0, 0)):
    u = np.linspace(0, 2 * math.pi, n_major, endpoint=False)
    v = np.linspace(0, 2 * math.pi, n_minor, endpoint=False)
    uu, vv = np.meshgrid(u, v, indexing="ij")
    x = (R + r * np.cos(vv)) * np.cos(uu)
    y = (R + r * np.cos(vv)) * np.sin(uu)
    z = r * np.sin(vv)
    pos = np.stack([x, y, z], -1).reshape(-1, 3) + np.asarray(center)
    nrm = np.stack([np.cos(vv) * np.cos(uu), np.cos(vv) * np.sin(uu), np.sin(vv)], -1).reshape(-1, 3)
    tan = np.stack([-np.sin(uu), np.cos(uu), np.zeros_like(uu)], -1).reshape(-1, 3)
    tan = np.concatenate([tan, np.ones((tan.shape[0], 1))], axis=1)
    uv = np.stack([uu / (2 * math.pi), vv / (2 * math.pi)], -1).reshape(-1, 2)
    i = np.arange(n_major)[:, None]
    j = np.arange(n_minor)[None, :]
    a = (i * n_minor + j)
    b = (((i + 1) % n_major) * n_minor + j)
    c = (((i + 1) % n_major) * n_minor + (j + 1) % n_minor)
    d = (i * n_minor + (j + 1) % n_minor)
    idx = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)])
    return pos, idx.astype(np.uint32), nrm, tan, uv


def _add(sc, mesh, material):
    pos, idx, nrm, tan, uv = mesh
    m = sc.add_mesh(pos, idx, nrm, tan, uv)
    return sc.add_instance(m, material)


def cornell_scene(seed=0):
    """Room of 5 walls (10 triangles), a 1x1 area light (2 triangles, radiance 15) below the ceiling,
    two boxes rotated by seed-drawn angles in +-30 degrees (24 triangles) and a 3x subdivided
    icosphere of radius 0.3 (1280 triangles). Lambertian albedos from U[0.2, 0.8]^3 and one GGX metal
    with roughness 0.3. The room spans [-1, 1]^2 x [0, 2]; z is up, the camera looks along +y."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sc = Scene()

    def lambert():
        return sc.add_material(base_color=tuple(rng.uniform(0.2, 0.8, 3)), metallic=0.0, roughness=1.0)

    m_floor, m_ceil, m_back, m_left, m_right = lambert(), lambert(), lambert(), lambert(), lambert()
    m_box1, m_box2 = lambert(), lambert()
    m_metal = sc.add_material(base_color=(0.9, 0.75, 0.5), metallic=1.0, roughness=0.3)
    m_light = sc.add_material(base_color=(1, 1, 1), emissive=(15.0, 15.0, 15.0), metallic=0.0, roughness=1.0)
    L = 1.0
    _add(sc, _quad((-L, -L, 0), (L, -L, 0), (L, L, 0), (-L, L, 0), (0, 0, 1), (1, 0, 0)), m_floor)
    _add(sc, _quad((-L, -L, 2), (-L, L, 2), (L, L, 2), (L, -L, 2), (0, 0, -1), (1, 0, 0)), m_ceil)
    _add(sc, _quad((-L, L, 0), (L, L, 0), (L, L, 2), (-L, L, 2), (0, -1, 0), (1, 0, 0)), m_back)
    _add(sc, _quad((-L, -L, 0), (-L, L, 0), (-L, L, 2), (-L, -L, 2), (1, 0, 0), (0, 1, 0)), m_left)
    _add(sc, _quad((L, -L, 0), (L, -L, 2), (L, L, 2), (L, L, 0), (-1, 0, 0), (0, 1, 0)), m_right)
    _add(sc, _quad((-0.5, -0.5, 1.99), (-0.5, 0.5, 1.99), (0.5, 0.5, 1.99), (0.5, -0.5, 1.99), (0, 0, -1),
                   (1, 0, 0)), m_light)
    a1, a2 = np.radians(rng.uniform(-30, 30, 2))
    _add(sc, box_mesh((-0.4, 0.3, 0.6), (0.3, 0.3, 0.6), a1), m_box1)
    _add(sc, box_mesh((0.45, -0.2, 0.3), (0.3, 0.3, 0.3), a2), m_box2)
    _add(sc, icosphere(3, radius=0.3, center=(0.4, -0.25, 0.9)), m_metal)
    return sc


def textured_cornell_scene(seed=0):
    """The Cornell room with textures on every material channel the reference samples: an sRGB
    checker base colour on the floor (uv repeated 3.5 times, so wrapping is exercised), a
    metallic-roughness texture on the back wall, an emissive texture (factor x texture) on the
    light and a bumpy normal map on the left wall."""
    sc = cornell_scene(seed)
    rng = np.random.Generator(np.random.PCG64(seed + 101))
    n = 16
    jj, ii = np.mgrid[0:n, 0:n]
    checker = np.where((ii // 4 + jj // 4) % 2 == 0, 230, 40).astype(np.uint8)
    base = np.stack([checker, checker // 2 + 60, 255 - checker, np.full_like(checker, 255)], -1)
    t_base = sc.add_texture(base, TEXTURE_RGBA8_SRGB)
    mr = np.zeros((8, 8, 4), np.uint8)
    mr[..., 1] = rng.integers(40, 255, (8, 8))   # roughness in g
    mr[..., 2] = rng.integers(0, 255, (8, 8))    # metallic in b
    mr[..., 3] = 255
    t_mr = sc.add_texture(mr, TEXTURE_RGBA8_UNORM)
    em = np.ones((4, 4, 4), np.float32)
    em[..., :3] = rng.uniform(0.3, 1.0, (4, 4, 3))
    t_em = sc.add_texture(em, TEXTURE_RGBA32F)
    x = np.linspace(0, 4 * np.pi, 32)
    bump = np.stack([0.5 + 0.35 * np.sin(x)[None, :].repeat(32, 0), 0.5 + 0.35 * np.cos(x)[:, None].repeat(32, 1),
                     np.full((32, 32), 0.9), np.ones((32, 32))], -1)
    t_nrm = sc.add_texture((bump * 255).astype(np.uint8), TEXTURE_RGBA8_UNORM)
    # instance order of cornell_scene: floor, ceiling, back, left, right, light, boxes, sphere
    floor, back, left, light = (sc.instances[k]["material"] for k in (0, 2, 3, 5))
    sc.materials[floor]["base_color_texture"] = t_base
    sc.materials[back]["metallic_roughness_texture"] = t_mr
    sc.materials[back]["metallic"] = 0.8
    sc.materials[back]["roughness"] = 0.9
    sc.materials[light]["emissive_texture"] = t_em
    sc.materials[left]["normal_texture"] = t_nrm
    for k in (0, 2, 3, 5):
        fill = {"base_color_texture": -1, "metallic_roughness_texture": -1, "emissive_texture": -1,
                "normal_texture": -1}
        fill.update({f: v for f, v in sc.materials[sc.instances[k]["material"]].items() if f.endswith("_texture")})
        sc.materials[sc.instances[k]["material"]].update(fill)
    sc.meshes[sc.instances[0]["mesh"]]["uvs"] = (sc.meshes[sc.instances[0]["mesh"]]["uvs"] * 3.5 - 0.25).astype(
        np.float32)
    return sc


def cornell_camera(width=256, height=256):
    # camera outside the open front of the room, looking along +y
    # CameraComponent defaults: 35 mm lens on a 24 mm sensor at f/22 -> 37.85 deg, radius 7.95e-4
    return camera_desc(position=(0.0, -3.4, 1.0), front=(0.0, 1.0, 0.0), vertical_fov=None, focus_distance=3.4,
                       aperture_radius=0.035 / 22.0 * 0.5, aspect=width / height)


# ---------------------------------------------------------------------------------------------
# many-instance scene: seed-placed icospheres, tori and noise-displaced boxes on a 20 x 20 ground
# ---------------------------------------------------------------------------------------------
_ICO_CACHE = {}


def _unit_icosphere(subdiv):
    if subdiv not in _ICO_CACHE:
        _ICO_CACHE[subdiv] = icosphere(subdiv, 1.0, (0, 0, 0))
    return _ICO_CACHE[subdiv]


def displaced_box(n, amp, rng):
    """Box [-1,1]^3 with each face split into n x n cells and displaced along its normal by value
    noise of amplitude `amp`; normals are the face normals, tangents the face u axis."""
    pos_all, idx_all, nrm_all, tan_all, uv_all = [], [], [], [], []
    faces = [((1, 0, 0), (0, 1, 0)), ((-1, 0, 0), (0, -1, 0)), ((0, 1, 0), (-1, 0, 0)), ((0, -1, 0), (1, 0, 0)),
             ((0, 0, 1), (1, 0, 0)), ((0, 0, -1), (-1, 0, 0))]
    g = np.linspace(-1.0, 1.0, n + 1)
    base = 0
    for nv, tv in faces:
        nv = np.array(nv, float)
        tv = np.array(tv, float)
        bv = np.cross(nv, tv)
        uu, vv = np.meshgrid(g, g, indexing="xy")
        disp = amp * (value_noise(n + 1, octaves=3, base_freq=2, seed=int(rng.integers(1 << 30))) - 0.5)
        # keep face borders closed: no displacement on the rim
        rim = (np.abs(uu) > 1 - 1e-9) | (np.abs(vv) > 1 - 1e-9)
        disp[rim] = 0.0
        p = nv[None, None, :] * (1.0 + disp)[..., None] + uu[..., None] * tv + vv[..., None] * bv
        pos_all.append(p.reshape(-1, 3))
        nrm_all.append(np.tile(nv, ((n + 1) ** 2, 1)))
        tan_all.append(np.tile(np.concatenate([tv, [1.0]]), ((n + 1) ** 2, 1)))
        uv_all.append(np.stack([(uu + 1) / 2, (vv + 1) / 2], -1).reshape(-1, 2))
        ci, cj = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
        a = (cj * (n + 1) + ci).reshape(-1) + base
        b, c, d = a + 1, a + n + 2, a + n + 1
        idx_all.append(np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)]))
        base += (n + 1) ** 2
    return (np.concatenate(pos_all), np.concatenate(idx_all).astype(np.uint32), np.concatenate(nrm_all),
            np.concatenate(tan_all), np.concatenate(uv_all))


def many_instance_scene(seed=0, count=400, target_tris=5_000_000, light_radiance=25.0, aspect=1920 / 1080,
                        shared=False):
    """`count` seed-placed instances flattened to world space (about `target_tris` triangles):
    icospheres (subdivision 2-5, capped by the per-instance budget), tori (R/r = 3) and noise-displaced
    boxes on a 20 x 20 ground plane; materials 60% Lambertian, 30% GGX metal (roughness U[0.05, 0.6]),
    10% dielectric (IOR 1.5); three emissive area lights. Returns (scene, camera with the given
    aspect ratio, 1920x1080 by default).
    shared: the same placement as instances instead: one object-space mesh per icosphere subdivision
    and one torus, shared by every instance of that kind, and each displaced box as its own mesh,
    all placed by world_from_local transforms (scale, rotation, translation)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sc = Scene()
    ground = sc.add_material(base_color=(0.5, 0.5, 0.5), metallic=0.0, roughness=1.0)
    _add(sc, _quad((-10, -10, 0), (10, -10, 0), (10, 10, 0), (-10, 10, 0), (0, 0, 1), (1, 0, 0)), ground)
    # per-instance triangle budget steering the mesh resolutions toward target_tris
    per = target_tris / count
    max_sub = 2
    while max_sub < 5 and 20 * 4 ** (max_sub + 1) <= 2 * per:
        max_sub += 1
    side = int(math.ceil(math.sqrt(count)))
    cell = 19.0 / side
    protos = {}

    def proto(key, make):
        if key not in protos:
            protos[key] = sc.add_mesh(*make())
        return protos[key]

    for k in range(count):
        gx, gy = k % side, k // side
        cx = -9.5 + (gx + 0.5) * cell + rng.uniform(-0.2, 0.2) * cell
        cy = -9.5 + (gy + 0.5) * cell + rng.uniform(-0.2, 0.2) * cell
        size = rng.uniform(0.18, 0.38) * cell
        kind = rng.integers(3)
        M = None  # world_from_local of the shared form
        if kind == 0:
            sub = min(int(rng.integers(2, 6)), max_sub)
            if shared:
                m = proto(("ico", sub), lambda: _unit_icosphere(sub))
                M = np.array([[size, 0, 0, cx], [0, size, 0, cy], [0, 0, size, size]])
            else:
                p, i, n, t, uv = _unit_icosphere(sub)
                p = p * size + np.array([cx, cy, size])
        elif kind == 1:
            r = size / 4.0
            nmaj = int(max(16, math.sqrt(per * 2)))
            nmin = max(8, nmaj // 2)
            if shared:
                m = proto(("torus", nmaj, nmin), lambda: torus(3.0, 1.0, nmaj, nmin, center=(0.0, 0.0, 0.0)))
                M = np.array([[r, 0, 0, cx], [0, r, 0, cy], [0, 0, r, r * 1.05]])
            else:
                p, i, n, t, uv = torus(3 * r, r, nmaj, nmin, center=(cx, cy, r * 1.05))
        else:
            nb = int(max(4, math.sqrt(per / 12)))
            p, i, n, t, uv = displaced_box(nb, 0.25, rng)
            ang = rng.uniform(0, math.pi)
            c, s = math.cos(ang), math.sin(ang)
            R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
            if shared:
                m = sc.add_mesh(p, i, n, t, uv)
                M = np.concatenate([R * size, np.array([[cx], [cy], [size * 1.3]])], 1)
            else:
                p = (p * size) @ R.T + np.array([cx, cy, size * 1.3])
                n = n @ R.T
                t = t.copy()
                t[:, :3] = t[:, :3] @ R.T
        u = rng.random()
        if u < 0.6:
            mat = sc.add_material(base_color=tuple(rng.uniform(0.2, 0.8, 3)), metallic=0.0, roughness=1.0)
        elif u < 0.9:
            mat = sc.add_material(base_color=tuple(rng.uniform(0.5, 0.95, 3)), metallic=1.0,
                                  roughness=float(rng.uniform(0.05, 0.6)))
        else:
            mat = sc.add_material(base_color=(1.0, 1.0, 1.0), metallic=0.0, roughness=0.0, eta=1.5,
                                  type=MATERIAL_DIELECTRIC)
        if shared:
            sc.add_instance(m, mat, M.astype(np.float32))
        else:
            _add(sc, (p, i, n, t, uv), mat)
    for (lx, ly) in ((-5.0, -3.0), (4.0, -4.0), (0.0, 5.0)):
        light = sc.add_material(base_color=(1, 1, 1), emissive=(light_radiance,) * 3, metallic=0.0)
        h = 6.0
        _add(sc, _quad((lx - 1, ly - 1, h), (lx - 1, ly + 1, h), (lx + 1, ly + 1, h), (lx + 1, ly - 1, h),
                       (0, 0, -1), (1, 0, 0)), light)
    cam = camera_desc(position=(0.0, -16.0, 7.0), front=(0.0, 1.0, -0.42), vertical_fov=math.radians(50.0),
                      focus_distance=16.0, aperture_radius=0.0, aspect=aspect)
    return sc, cam


def batch_scenes(seed=0, count=64, min_tris=1_000, max_tris=500_000, sizes=(64, 128, 256, 512)):
    """`count` independent scenes of the many-instance generator: triangle budgets log-uniform in
    [min_tris, max_tris], square images with sides drawn from `sizes`. Returns a list of
    (scene, camera, width, height)."""
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    out = []
    for k in range(count):
        target = float(np.exp(rng.uniform(math.log(min_tris), math.log(max_tris))))
        side = int(sizes[int(rng.integers(len(sizes)))])
        inst = int(min(400, max(4, target // 1500)))
        sc, cam = many_instance_scene(seed=seed * 1000 + k, count=inst, target_tris=target, aspect=1.0)
        out.append((sc, cam, side, side))
    return out
