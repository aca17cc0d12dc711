"""Build libnecpp_amd.so (host C++ + HIP kernels for gfx950) in-tree.

Usage: python -m necpp_amd.build   (or call build()).  Every source is
compiled with hipcc for --offload-arch=gfx950; the fill and the host
geometry/basis code are compiled without FP contraction so that they keep
nec2++'s rounding.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnecpp_amd.so")
ARCH = os.environ.get("NECPP_AMD_ARCH", "gfx950")

SOURCES = ["geometry.cpp", "kclock.cpp", "engine.cpp", "batch_host.cpp", "adjoint.cpp", "context.cpp", "api.cpp",
           "deck.cpp",
           "fill.hip", "lu.hip", "solve.hip", "refine.hip", "batch.hip", "current.hip", "pattern.hip"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
          f"--offload-arch={ARCH}",
          # packed FP32 instructions stay off: some of their forms have given
          # wrong results on this hardware without any error
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-fvisibility=hidden", "-I", os.path.join(HERE, "..", "include")]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"]
    cmd = [hipcc()] + lang + COMMON + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose=False):
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
