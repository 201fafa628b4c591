"""Build the CPU oracle (test infrastructure) into oracle/_build/liboracle.so,
and, where the reference's sources are present, the one part of them that
compiles on its own: the vendored mikktspace library
(External/mikktspace/mikktspace.c), built where it lies together with
mikk_ref_driver.c into oracle/_ref/libmikk_ref.so.  The rest of the reference
(CUDA kernels, cuRAND, <expected>, glm, entt) cannot be compiled here.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "liboracle.so")
REF_OUT = os.path.join(HERE, "_ref", "libmikk_ref.so")
REF_MIKK = "/root/reference/External/mikktspace"


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-pthread",
           "-Wall", "-o", tmp, os.path.join(HERE, "oracle.cpp")]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


def build_ref(verbose=False):
    """Compile the reference's mikktspace.c (release flags: -O2 -DNDEBUG, no
    FMA contraction) with the driver; returns the library path, or None when
    the reference's sources are not present."""
    src = os.path.join(REF_MIKK, "mikktspace.c")
    if not os.path.exists(src):
        return None
    os.makedirs(os.path.dirname(REF_OUT), exist_ok=True)
    tmp = REF_OUT + ".tmp"
    cmd = ["gcc", "-O2", "-std=c99", "-DNDEBUG", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
           f"-I{REF_MIKK}", "-o", tmp, os.path.join(HERE, "mikk_ref_driver.c"), src, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"reference mikktspace build failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, REF_OUT)
    return REF_OUT


if __name__ == "__main__":
    try:
        print(build(verbose="--verbose" in sys.argv))
        print(build_ref(verbose="--verbose" in sys.argv))
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
