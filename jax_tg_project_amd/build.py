"""Build recipe for libtgdns.so (gfx950), in-tree next to this file.

    python -m jax_tg_project_amd.build
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "tgdns.hip")
OUT = os.path.join(HERE, "libtgdns.so")
DEPS = [os.path.join(HERE, "csrc", f) for f in ("tgdns.hip", "tg_passes.hpp", "tg_fft.hpp")] + [
    os.path.join(ROOT, "include", "tgdns.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # keep the compiler from emitting packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=True):
    """Compile csrc/tgdns.hip into libtgdns.so (always recompiles)."""
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with code {res.returncode}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
