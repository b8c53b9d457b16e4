"""The C-ABI library loads and exports every symbol include/spk_pt.h declares (no GPU needed)."""
import ctypes
import os
import re

import sparkle_amd as spk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "spk_pt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spk_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_all_declared_symbols():
    lib = spk.load_library()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_has_no_cpp_or_torch_types():
    text = open(os.path.join(ROOT, "include", "spk_pt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    for bad in ("std::", "torch", "at::", "hipStream_t", "class "):
        assert bad not in text


# ctypes mirror of each struct the header declares
STRUCTS = {"spk_mesh_desc": "MeshDesc", "spk_instance_desc": "InstanceDesc", "spk_material_desc": "MaterialDesc",
           "spk_sphere_desc": "SphereDesc", "spk_texture_desc": "TextureDesc", "spk_build_options": "BuildOptions",
           "spk_scene_stats": "SceneStats", "spk_trace_config": "TraceConfig", "spk_camera_desc": "CameraDesc",
           "spk_frame_desc": "FrameDesc", "spk_cache_config": "CacheConfig", "spk_frame_stats": "FrameStats",
           "spk_schedule": "Schedule", "spk_batch_scene": "BatchScene"}


def test_struct_sizes_match_header_layout(tmp_path):
    """A C program compiled against the header prints every struct's size; the ctypes mirrors
    must agree (a field added on one side only fails here)."""
    import subprocess

    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "spk_pt.h"\nint main(void) {\n' +
                   "".join(f'    printf("%s %zu\\n", "{c}", sizeof({c}));\n' for c in STRUCTS) + "    return 0;\n}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    sizes = dict(zip(out[0::2], map(int, out[1::2])))
    assert set(sizes) == set(STRUCTS)
    for c, py in STRUCTS.items():
        assert ctypes.sizeof(getattr(spk, py)) == sizes[c], (c, ctypes.sizeof(getattr(spk, py)), sizes[c])
    assert sizes["spk_cache_config"] == 44
    assert spk.RAY_DTYPE.itemsize == 32 and spk.HIT_DTYPE.itemsize == 8


def test_create_without_device_fails_cleanly():
    # On a machine without a GPU spk_create must return an error code, never crash.
    import torch

    if torch.cuda.is_available():
        return
    lib = spk.load_library()
    h = ctypes.c_void_p()
    rc = lib.spk_create(0, ctypes.byref(h))
    assert rc in (spk.SPK_ERR_NO_DEVICE, spk.SPK_ERR_HIP)


def test_null_arguments_are_rejected():
    lib = spk.load_library()
    assert lib.spk_destroy(None) == spk.SPK_ERR_INVALID_ARGUMENT
    assert lib.spk_set_camera(None, None) == spk.SPK_ERR_INVALID_ARGUMENT
    assert lib.spk_create(0, None) == spk.SPK_ERR_INVALID_ARGUMENT


def test_host_adapter_demo_is_built_and_fails_cleanly_without_a_device():
    import subprocess

    import torch

    demo = os.path.join(ROOT, "sparkle_amd", "lib", "render_demo")
    assert os.path.exists(demo), "run __graft_entry__.build()"
    if torch.cuda.is_available():
        return
    r = subprocess.run([demo, "/tmp/spk_demo_nodev.pfm", "8", "8", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "init failed" in r.stderr
