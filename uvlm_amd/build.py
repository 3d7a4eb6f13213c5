"""In-tree build of the gfx950 library ``uvlm_amd/lib/libuvlm_gpu.so``.

hipcc cross-compiles for gfx950 without a GPU.  Every HIP source is built
with ``-ffp-contract=off`` (kernels that want FMA call it explicitly, the
F64_EXACT kernels must not be contracted) and with packed FP32 instructions
disabled.  The C++ host library (the reference-shaped ``UVLM`` API) is built
with g++ by ``uvlm_amd/host/Makefile`` and links the GPU library.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
GPU_LIB = os.path.join(LIBDIR, "libuvlm_gpu.so")
HOST_LIB = os.path.join(LIBDIR, "libuvlm.so")
SOURCES = ["capi.hip", "nbody.hip", "chain.hip", "dense.hip", "step.hip", "batch.hip",
           "farfield.hip", "inv_solve.hip", "mlp.hip", "gemm.hip"]
ARCH = os.environ.get("UVLM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_gpu(verbose: bool = False) -> str:
    os.makedirs(os.path.join(LIBDIR, "obj"), exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, "obj", src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(4, len(cmds))) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = GPU_LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp,
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


def build_host(verbose: bool = False) -> str:
    mk = os.path.join(HERE, "host", "Makefile")
    if os.path.exists(mk):
        out = _run(["make", "-s", "-C", os.path.join(HERE, "host")])
        if verbose and out.strip():
            print(out)
    return HOST_LIB


def build(verbose: bool = False) -> None:
    build_gpu(verbose)
    build_host(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", GPU_LIB)
