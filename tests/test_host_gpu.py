"""The C++ host adapter (GpuPathRenderer, CPURenderer's role over the C ABI) renders a room."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_render_demo_produces_an_image(tmp_path):
    demo = os.path.join(ROOT, "sparkle_amd", "lib", "render_demo")
    out = str(tmp_path / "room.pfm")
    r = subprocess.run([demo, out, "96", "64", "8"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    stats = json.loads(r.stdout.strip().splitlines()[-1])
    assert stats["frames"] == 8 and stats["cumulated"] == 7 and stats["rays"] > 96 * 64 * 8
    with open(out, "rb") as f:
        assert f.readline() == b"PF\n"
        w, h = map(int, f.readline().split())
        f.readline()
        img = np.frombuffer(f.read(), dtype="<f4").reshape(h, w, 3)
    assert (w, h) == (96, 64)
    assert 0.05 < img.mean() < 0.95 and np.isfinite(img).all()


def test_render_demo_with_spatial_fill(tmp_path):
    demo = os.path.join(ROOT, "sparkle_amd", "lib", "render_demo")
    out = str(tmp_path / "room_fill.pfm")
    r = subprocess.run([demo, out, "64", "48", "4", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with open(out, "rb") as f:
        f.readline()
        w, h = map(int, f.readline().split())
        f.readline()
        img = np.frombuffer(f.read(), dtype="<f4").reshape(h, w, 3)
    assert (w, h) == (64, 48) and np.isfinite(img).all()


def _demo_room():
    """The room render_demo builds (same vertices, normals, tangents and materials, float32)."""
    import sparkle_amd as spk

    quads = [
        ([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], [0, 0, 1], [1, 0, 0]),
        ([[-1, -1, 2], [-1, 1, 2], [1, 1, 2], [1, -1, 2]], [0, 0, -1], [1, 0, 0]),
        ([[-1, 1, 0], [1, 1, 0], [1, 1, 2], [-1, 1, 2]], [0, -1, 0], [1, 0, 0]),
        ([[-1, -1, 0], [-1, 1, 0], [-1, 1, 2], [-1, -1, 2]], [1, 0, 0], [0, 1, 0]),
        ([[1, -1, 0], [1, -1, 2], [1, 1, 2], [1, 1, 0]], [-1, 0, 0], [0, 1, 0]),
        ([[-0.5, -0.5, 1.99], [-0.5, 0.5, 1.99], [0.5, 0.5, 1.99], [0.5, -0.5, 1.99]], [0, 0, -1], [1, 0, 0]),
    ]
    albedo = [(0.7, 0.7, 0.7), (0.7, 0.7, 0.7), (0.7, 0.7, 0.7), (0.7, 0.2, 0.2), (0.2, 0.6, 0.2), (1, 1, 1)]
    sc = spk.Scene()
    for k, (p, n, t) in enumerate(quads):
        m = sc.add_mesh(np.array(p, np.float32), np.array([[0, 1, 2], [0, 2, 3]], np.uint32),
                        np.tile(np.array(n, np.float32), (4, 1)), np.tile(np.array(t + [1], np.float32), (4, 1)))
        mat = sc.add_material(base_color=albedo[k], emissive=(15.0,) * 3 if k == 5 else (0, 0, 0), metallic=0.0,
                              roughness=1.0)
        sc.add_instance(m, mat)
    return sc


def test_render_demo_accumulation_matches_oracle(tmp_path):
    """GpuPathRenderer's sample bookkeeping (frame seed = dispatched samples, running mean with the
    cumulated count, CPURenderer::Render / DenoisePass, CPURenderer.cpp:119-208, 527-542) against the
    oracle's frames accumulated the same way: equal on the pixels whose sample paths are the same,
    within 3 sigma elsewhere, and the cumulated count the reference keeps."""
    import sparkle_amd as spk
    from oracle import refcpu

    demo = os.path.join(ROOT, "sparkle_amd", "lib", "render_demo")
    w, h, frames = 64, 48, 12
    out, acc_path = str(tmp_path / "room.pfm"), str(tmp_path / "acc.pfm")
    r = subprocess.run([demo, out, str(w), str(h), str(frames), "0", acc_path], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    stats = json.loads(r.stdout.strip().splitlines()[-1])
    assert stats["cumulated"] == frames - 1
    with open(acc_path, "rb") as f:
        assert f.readline() == b"PF\n"
        assert tuple(map(int, f.readline().split())) == (w, h)
        f.readline()
        dev = np.frombuffer(f.read(), dtype="<f4").reshape(h, w, 3)
    cam = spk.CameraDesc()
    for k, v in stats["camera"].items():
        if isinstance(v, list):
            getattr(cam, k)[:] = v
        else:
            setattr(cam, k, v)
    o = refcpu.OracleScene(_demo_room())
    acc = np.zeros((h, w, 4), np.float32)
    samples = []
    for k in range(frames):
        rgba, _ = o.render_frame(cam, w, h, frame_seed=k, max_bounce=4)
        refcpu.accumulate(rgba, acc, k, 1)
        samples.append(rgba[..., :3].astype(np.float64))
    same = (np.abs(dev - acc[..., :3]).max(-1) <= 1e-5 * np.abs(acc[..., :3]).max(-1) + 1e-6).mean()
    s = np.stack(samples)
    se = s.std(0, ddof=1) / np.sqrt(frames)
    within = (np.abs(dev - acc[..., :3]) <= 3 * se + 1e-5).all(-1).mean()
    print("demo accumulation: equal pixels", same, "within 3 sigma", within)
    assert same > 0.9 and within > 0.99
