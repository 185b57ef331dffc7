"""Build libtempo_hip.so for gfx950 (in-tree: tempo_amd/lib/).

    python -m tempo_amd.build

Device code is written for CDNA4 only.  All HIP sources are compiled with
packed-FP32 instructions disabled; the FP64 twin and the FP64 table builder
with -ffp-contract=off so they round exactly as the reference FP64 build;
the FP32 path with contraction (outside the scopes that turn it off),
reciprocal and approximate-function math.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(LIBDIR, "libtempo_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]

# (source, extra flags)
UNITS = [
    # FP32 production path: contracted multiply-adds (except in the scopes a
    # `#pragma clang fp contract(off)` marks: the checks whose cancellations
    # must be exact, and the sedimentation, so that its kernels agree bit for
    # bit), hardware reciprocal and approximate exp/log/sqrt (errors of a few
    # ulps, within the FP32 tolerances)
    ("tempo_step_f32.hip", DEVICE + ["-ffp-contract=fast-honor-pragmas", "-fapprox-func", "-freciprocal-math"]),
    ("tempo_step_f64.hip", DEVICE + ["-ffp-contract=off"]),
    ("tempo_tables.hip", DEVICE + ["-ffp-contract=off"]),
    ("tempo_capi.cpp", ["-x", "hip"] + DEVICE + ["-ffp-contract=off"]),
    ("tempo_noalloc.hip", DEVICE),
    ("tempo_host.cpp", ["-ffp-contract=off"]),
]


def _compile(unit):
    src, flags = unit
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        objs = list(ex.map(_compile, UNITS))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
