"""Build libstfem_amd.so in-tree with hipcc for gfx950.

Usage: python dealii-stfem_amd/build.py [-j N]
Each source is compiled to an object file under dealii-stfem_amd/_build/
(in parallel), then linked into dealii-stfem_amd/libstfem_amd.so.
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libstfem_amd.so")
BUILD = os.path.join(HERE, "_build")

SOURCES = ["k_operator.hip", "k_noalloc.hip", "k_error.hip", "k_setup.hip", "k_eigen.hip", "k_adjoint.hip", "k_vanka.hip", "k_transfer.hip", "k_blas.hip",
           "time_fe.cpp", "spec.cpp", "mesh.cpp", "stfem_host.cpp", "capi.cpp", "capi_objects.cpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
COMMON = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
          "-munsafe-fp-atomics",
          # packed FP32 VALU instructions are kept out of all generated code
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-variable", "-Wno-unused-function"]


def compile_one(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-fopenmp", "-I" + CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{msgs}")
    return obj, msgs


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, SOURCES))
    if verbose:
        for obj, msgs in results:
            if msgs.strip():
                print(msgs)
    objs = [o for o, _ in results]
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-fopenmp", "-o", tmp] + objs + \
        ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas",
         "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/llvm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))
    sys.exit(0)
