"""Builds the in-tree HIP library torch_amd/libtorchgpu.so for gfx950.

Each .hip translation unit is compiled in parallel with hipcc, then linked
into one shared library exporting the C ABI of include/torchgpu.h.

Flags:
  -ffp-contract=off        no FMA contraction, so the FP64 path rounds every
                           operation like the reference's x86-64 build
                           (the mixed-precision kernel, which has no bitwise
                           contract, is compiled with -ffp-contract=fast)
  -packed-fp32-ops off     no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtorchgpu.so")
SOURCES = ["hydro.hip", "hydro_mixed.hip", "thermo.hip", "raytrace.hip", "radiation.hip", "mlp.hip", "api.hip",
           "units.cpp", "mlp_api.cpp", "snapshot.cpp"]
ARCH = os.environ.get("TORCHGPU_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
# per-file replacements of the contraction flag
CONTRACT = {"hydro_mixed.hip": "-ffp-contract=fast"}


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        flags = [CONTRACT.get(src, f) if f.startswith("-ffp-contract") else f for f in CFLAGS]
        cmd = [HIPCC, *flags, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        msg = "\n".join(f"--- {s} ---\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lz"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


HOST_SOURCES = ["lua_config.cpp", "lua_setup.cpp", "torch_gpu.cpp", "torch_gpu_main.cpp"]
HOST_EXE = os.path.join(HERE, "torch_gpu")


def build_host(verbose: bool = False) -> str:
    """The C++ host layer and command-line driver, linked against the C ABI."""
    srcs = [os.path.join(HERE, "host", s) for s in HOST_SOURCES]
    rocm = os.path.dirname(os.path.dirname(HIPCC))
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{os.path.dirname(HERE)}/include",
           *srcs, "-o", HOST_EXE + ".tmp", f"-L{HERE}", "-ltorchgpu", f"-Wl,-rpath,$ORIGIN",
           f"-Wl,-rpath,{rocm}/lib", f"-L{rocm}/lib", "-lamdhip64"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(HOST_EXE + ".tmp", HOST_EXE)
    return HOST_EXE


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
