"""In-tree build of the native libraries (no JIT cache, nothing installed).

* opensubdiv_amd/lib/libosdhip.so -- HIP kernels + C ABI, gfx950 only
  (hipcc --offload-arch=gfx950; packed-FP32 instructions disabled).
* opensubdiv_amd/lib/libosdfar.so -- host table compiler (C++17, g++),
  when its sources exist.

``python -m opensubdiv_amd.build`` or ``build_all()``.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall",
             "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_hip(verbose=False):
    """Every csrc/*.hip to an object (in parallel), then one shared library."""
    os.makedirs(os.path.join(LIB, "obj"), exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    objs = [os.path.join(LIB, "obj", os.path.basename(s)[:-4] + ".o") for s in srcs]
    procs = [(subprocess.Popen([HIPCC, *HIP_FLAGS, "-c", "-o", o, s], stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True), s) for s, o in zip(srcs, objs)]
    for p, s in procs:
        text, _ = p.communicate()
        if p.returncode != 0:
            text = "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)
            raise RuntimeError(f"build failed: {s}\n{text}")
    out = os.path.join(LIB, "libosdhip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_far(verbose=False):
    srcs = sorted(glob.glob(os.path.join(HERE, "far", "*.cpp")))
    if not srcs:
        return None
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libosdfar.so")
    tmp = out + ".tmp"
    _run([os.environ.get("CXX", "g++"), *CXX_FLAGS, "-o", tmp, *srcs])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_all(verbose=False):
    return [build_hip(verbose), build_far(verbose)]


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
