"""Writes inputs_sha256.json: the sha256 of every synthetic input buffer the benchmarks and parity
tests generate (heightfield vertices and indices, both ray sets, the many-instance and Cornell
scenes), so a change to a generator is caught. Run from the repository root:

    python tests/golden/make_input_hashes.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from sparkle_amd import scenes  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def scene_hash(sc):
    h = hashlib.sha256()
    for m in sc.meshes:
        h.update(m["positions"].tobytes())
        h.update(m["indices"].tobytes())
    for i in sc.instances:
        h.update(np.asarray([i["mesh"], i["material"]], dtype=np.uint32).tobytes())
        h.update(i["M"].astype(np.float32).tobytes())
    for m in sc.materials:
        h.update(json.dumps(m, sort_keys=True).encode())
    return h.hexdigest()


def compute():
    pos, idx, _, _, _ = scenes.heightfield_mesh(1024)
    sc = scenes.heightfield_scene(1024)
    lo, hi = scenes.scene_aabb(sc)
    out = {
        "heightfield_1024_positions": sha(pos.astype(np.float32)),
        "heightfield_1024_indices": sha(idx.astype(np.uint32)),
        "incoherent_rays_8m_seed0": sha(scenes.incoherent_rays(lo, hi)),
        "coherent_rays_1080p": sha(scenes.coherent_rays()),
        "cornell_scene_seed0": scene_hash(scenes.cornell_scene(0)),
        "many_instance_scene_seed0": scene_hash(scenes.many_instance_scene(0)[0]),
    }
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "inputs_sha256.json"), "w") as f:
        json.dump(compute(), f, indent=1, sort_keys=True)
        f.write("\n")
