"""Builds the HIP extension libspk_pt.so (gfx950) in-tree.

Device code: hipcc --offload-arch=gfx950, no FMA contraction (the closest-hit arithmetic must round
exactly like the CPU oracle), IEEE single-precision division and square root, packed FP32 disabled.
Host-only sources (.cpp) are compiled with g++. Objects go to sparkle_amd/_build/, the library to
sparkle_amd/lib/.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libspk_pt.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread", "-Wall", "-Wextra",
             "-Wno-unused-parameter"]


def _sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def _headers_mtime():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    hs.append(os.path.join(HERE, "..", "include", "spk_pt.h"))
    return max(os.path.getmtime(h) for h in hs if os.path.exists(h))


def _compile(cmd, src, obj, force):
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(src), _headers_mtime()):
        return None
    tmp = obj + ".tmp.o"
    r = subprocess.run(cmd + ["-c", src, "-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        return f"{os.path.basename(src)}:\n{r.stdout}\n{r.stderr}"
    os.replace(tmp, obj)
    return None


def build(force=True, verbose=True):
    """Compiles every source (always when force) and links libspk_pt.so. Raises on failure."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    hip, cpp = _sources()
    jobs = []
    objs = []
    for f in hip:
        obj = os.path.join(BUILD, f + ".o")
        jobs.append(([HIPCC] + HIP_FLAGS, os.path.join(CSRC, f), obj))
        objs.append(obj)
    for f in cpp:
        obj = os.path.join(BUILD, f + ".o")
        jobs.append((["g++"] + CXX_FLAGS, os.path.join(CSRC, f), obj))
        objs.append(obj)
    workers = min(len(jobs), max(1, (os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        errs = [e for e in ex.map(lambda j: _compile(j[0], j[1], j[2], force), jobs) if e]
    if errs:
        raise RuntimeError("HIP build failed:\n" + "\n".join(errs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_host(verbose)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


DEMO = os.path.join(LIBDIR, "render_demo")


def build_host(verbose=True):
    """The C++ host adapter (GpuPathRenderer) and its demo, linked against libspk_pt.so."""
    host = os.path.join(HERE, "host")
    tmp = DEMO + ".tmp"
    cmd = (["g++"] + CXX_FLAGS + ["-I", os.path.join(HERE, "..", "include"),
                                  os.path.join(host, "gpu_renderer.cpp"), os.path.join(host, "render_demo.cpp"),
                                  "-L", LIBDIR, "-lspk_pt", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host adapter build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, DEMO)
    return DEMO


if __name__ == "__main__":
    build(force="--incremental" not in sys.argv)
