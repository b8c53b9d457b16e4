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
