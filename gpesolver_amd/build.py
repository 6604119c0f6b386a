"""Build the gfx950 shared library libgpehip.so in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libgpehip.so")
EXE = os.path.join(HERE, "gpesolver")
SOURCES = [os.path.join(HERE, "csrc", "gpehip.hip"), os.path.join(HERE, "csrc", "gpe_fno.hip")]
HOST_SOURCES = [os.path.join(HERE, "host", "gpe_host.cpp"), os.path.join(HERE, "host", "gpe_host_abi.cpp")]
HEADERS = [os.path.join(HERE, "csrc", "gpe_fft.hpp"), os.path.join(HERE, "csrc", "gpe_fft_ct.hpp"),
           os.path.join(HERE, "csrc", "gpe_sincos.hpp"),
           os.path.join(HERE, "csrc", "gpe_gemm.hpp"),
           os.path.join(HERE, "host", "gpe_host.hpp"), os.path.join(ROOT, "include", "gpehip.h"),
           os.path.join(ROOT, "include", "gpesolver_host.h")]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Packed FP32 instructions are kept out of the generated code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def _run(cmd, verbose, what):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build failed: {what}")


def build(verbose: bool = False) -> str:
    """libgpehip.so = HIP kernels + C ABI (hipcc) + host layer (g++ objects);
    gpesolver = command-line front end linked against it."""
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(HERE, "host", os.path.basename(src) + ".o")
        _run(["g++", *HOST_FLAGS, "-c", src, "-o", obj], verbose, src)
        objs.append(obj)
    for src in SOURCES:
        obj = os.path.join(HERE, "csrc", os.path.basename(src) + ".o")
        _run([hipcc(), *HIPCC_FLAGS, "-c", src, "-o", obj], verbose, src)
        objs.append(obj)
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    _run(["g++", "-shared", "-o", LIB + ".tmp", *objs, f"-L{rocm}/lib", "-lamdhip64", f"-Wl,-rpath,{rocm}/lib"],
         verbose, "libgpehip.so")
    os.replace(LIB + ".tmp", LIB)
    _run(["g++", *HOST_FLAGS, os.path.join(HERE, "host", "main.cpp"), "-o", EXE + ".tmp", "-L" + HERE, "-lgpehip",
          "-Wl,-rpath,$ORIGIN"], verbose, "gpesolver")
    os.replace(EXE + ".tmp", EXE)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
