"""Build the HIP library rbmd_amd/_build/librbmd_mi355x.so for gfx950.

Every source is compiled by hipcc into an object in rbmd_amd/_build/ and the
objects are linked into one shared library with a plain C ABI
(include/rbmd_mi355x.h).  Packed-FP32 instruction selection is switched off
(-target-feature -packed-fp32-ops) for every device compile.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "librbmd_mi355x.so")
ARCH = os.environ.get("RBMD_OFFLOAD_ARCH", "gfx950")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

SOURCES = ["k_cells.hip", "k_force.hip", "k_pair.hip", "k_kspace.hip", "k_integrate.hip", "noalloc.hip", "api.cpp", "sampler.cpp"]
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return h


def _compile(src):
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
           "-I", os.path.join(HERE, "..", "include")]
    if src.endswith(".hip"):
        cmd += [f"--offload-arch={ARCH}", "-x", "hip"] + DEVICE_FLAGS + ["-munsafe-fp-atomics"]
    else:
        cmd += ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include")]
    cmd += ["-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for o, err in results:
            if err.strip():
                print(f"[{os.path.basename(o)}]\n{err}", file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
