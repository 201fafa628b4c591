"""Build libptcore.so (HIP kernels for gfx950 + C++ host code) in-tree.

Usage: python gpu-path-tracer_amd/build.py [--jobs N] [--verbose]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libptcore.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

SOURCES = ["kernels.hip", "guide.hip", "bvh_gpu.hip", "bvh_build.cpp", "scene.cpp", "renderer.cpp", "server.cpp", "guide_host.cpp", "obj_import.cpp",
           "tangent_space.cpp", "image_io.cpp", "capi.cpp"]

# -ffp-contract=off: no FMA contraction, so float expressions round the same
# way on the CPU and the GPU.  The packed-fp32 feature is switched off for all
# device code.
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function", "-fvisibility=hidden",
    f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(HERE, '..', 'include')}",
]


def _compile(src, verbose, build_dir=BUILD, defines=()):
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    cmd = [HIPCC] + COMMON + [f"-D{d}" for d in defines] + (["-x", "hip"] if src.endswith(".hip") else []) + \
        ["-c", path, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    warn = (r.stderr or "").strip()
    return obj, time.time() - t0, warn


def build(jobs=None, verbose=False, out=LIB, defines=()):
    """Compile every source and link `out`; `defines` are extra -D macros
    (variant builds for A/B timing go to their own build directory)."""
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    build_dir = BUILD if out == LIB else out + ".build"
    os.makedirs(build_dir, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_compile, s, verbose, build_dir, defines): s for s in SOURCES}
        for f in cf.as_completed(futs):
            obj, dt, warn = f.result()
            objs.append(obj)
            if verbose:
                print(f"  {os.path.basename(obj)}: {dt:.1f}s", flush=True)
            if warn and verbose:
                print(warn, flush=True)
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + sorted(objs) + ["-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--out", default=LIB)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    a = ap.parse_args()
    try:
        print(build(a.jobs, a.verbose, os.path.abspath(a.out), a.defines))
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
