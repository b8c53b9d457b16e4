"""Shared helpers for the tests (device arrays, random triangle soups)."""
import numpy as np

import sparkle_amd as spk


def to_dev(arr):
    import torch

    return torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).reshape(-1)).cuda()


def empty_dev(nbytes):
    import torch

    return torch.empty(max(1, nbytes), dtype=torch.uint8, device="cuda")


def hits_from_dev(t, n):
    return t.cpu().numpy()[: n * 8].view(spk.HIT_DTYPE).copy()


def gpu_trace(ctx, rays):
    n = rays.shape[0]
    d_rays = to_dev(rays)
    d_hits = empty_dev(n * 8)
    rc = ctx.trace(d_rays, d_hits, n)
    return rc, hits_from_dev(d_hits, n)


def soup_scene(n_tris, seed=0, scale=1.0, spread=0.05):
    """Random triangle soup in the unit cube (small triangles), one mesh, identity instance."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.random((n_tris, 1, 3))
    v = c + rng.normal(0.0, spread, (n_tris, 3, 3))
    pos = (v * scale).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * n_tris, dtype=np.uint32).reshape(-1, 3)
    sc = spk.Scene()
    m = sc.add_mesh(pos, idx)
    mat = sc.add_material(base_color=(0.5, 0.5, 0.5), metallic=0.0)
    sc.add_instance(m, mat)
    return sc


def multi_mesh_scene(seed=0):
    """Several meshes and instances (with transforms baked at upload) for id numbering tests."""
    from sparkle_amd import scenes

    sc = spk.Scene()
    mat = sc.add_material(base_color=(0.5, 0.5, 0.5), metallic=0.0)
    p, i, n, t, uv = scenes.icosphere(2, radius=0.3, center=(0.5, 0.5, 0.5))
    m0 = sc.add_mesh(p, i, n, t, uv)
    p, i, n, t, uv = scenes.box_mesh((0.2, 0.2, 0.2), (0.1, 0.15, 0.2), 0.3)
    m1 = sc.add_mesh(p, i, n, t, uv)
    sc.add_instance(m0, mat)
    sc.add_instance(m1, mat)
    sc.add_instance(m0, mat, np.array([[1, 0, 0, 0.3], [0, 1, 0, -0.1], [0, 0, 1, 0.2]], dtype=np.float32))
    return sc


def rotation(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    x, y, z = axis
    c, s_ = np.cos(angle), np.sin(angle)
    C = 1 - c
    return np.array([[c + x * x * C, x * y * C - z * s_, x * z * C + y * s_],
                     [y * x * C + z * s_, c + y * y * C, y * z * C - x * s_],
                     [z * x * C - y * s_, z * y * C + x * s_, c + z * z * C]])


def instanced_scene(seed=0, count=40, mirror=True):
    """Three meshes (icosphere, box, heightfield patch) placed `count` times with random rotations,
    non-uniform scales (one mirrored instance when `mirror`) and translations in [0, 4]^3, plus an
    emissive quad; returns the scene and the transforms of its instances."""
    from sparkle_amd import scenes

    rng = np.random.Generator(np.random.PCG64(seed))
    sc = spk.Scene()
    mats = [sc.add_material(base_color=tuple(rng.uniform(0.2, 0.8, 3)), metallic=0.0, roughness=1.0)
            for _ in range(3)]
    light = sc.add_material(base_color=(1, 1, 1), emissive=(8.0, 8.0, 8.0), metallic=0.0)
    meshes = [sc.add_mesh(*scenes.icosphere(2, radius=0.3, center=(0.0, 0.0, 0.0))),
              sc.add_mesh(*scenes.box_mesh((0.0, 0.0, 0.0), (0.25, 0.15, 0.2), 0.4))]
    hf = scenes.heightfield_scene(cells=12)
    meshes.append(sc.add_mesh(**{k: v for k, v in hf.meshes[0].items()}))
    xfs = []
    for k in range(count):
        R = rotation(rng.normal(size=3), rng.uniform(0, 2 * np.pi))
        S = np.diag(rng.uniform(0.5, 1.6, 3))
        if mirror and k == 3:
            S[0, 0] = -S[0, 0]
        M = np.zeros((3, 4))
        M[:, :3] = R @ S
        M[:, 3] = rng.uniform(0.0, 4.0, 3)
        M = M.astype(np.float32)
        sc.add_instance(meshes[k % 3], mats[k % 3], M)
        xfs.append(M)
    q = np.array([[1.5, 1.5, 4.6], [2.5, 1.5, 4.6], [2.5, 2.5, 4.6], [1.5, 2.5, 4.6]], np.float32)
    lm = sc.add_mesh(q, np.array([[0, 2, 1], [0, 3, 2]], np.uint32))
    M = np.eye(3, 4, dtype=np.float32)
    sc.add_instance(lm, light, M)
    xfs.append(M)
    return sc, np.stack(xfs)
