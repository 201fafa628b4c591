"""Build the CPU oracle (test infrastructure) into oracle/_build/liboracle.so.

The reference's own sources cannot be compiled here (CUDA, cuRAND, <expected>,
glm, entt), so there is no oracle/_ref build; see the design notes.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "liboracle.so")


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


if __name__ == "__main__":
    try:
        print(build(verbose="--verbose" in sys.argv))
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
