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
