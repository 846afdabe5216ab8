"""Compile the gfx950 library (lib/libekgsim.so) and the CLI (bin/ekgSim).

    python -m ekgsim_amd.build

Kernels are compiled for gfx950 only, with packed-FP32 instructions disabled.
Host code is compiled without floating-point contraction so that the FP64
host arithmetic (lead displacement, lag tables, time grids) rounds like the
reference build.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_obj")
LIB = os.path.join(PKG, "lib", "libekgsim.so")
CLI = os.path.join(PKG, "bin", "ekgSim")

ARCH = "gfx950"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
# MFMA results and accumulators in ordinary VGPRs: the kernels convert or
# read them right after the instruction, and an accumulation register would
# cost a copy per value.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-mllvm", "-amdgpu-mfma-vgpr-form"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HOST = ["-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]

KERNEL_SRC = ["ekg_ca.hip", "ekg_ecg.hip", "ekg_crit.hip", "ekg_scratch.hip"]
# The C++ facade (sim_lib.cpp with its ini reader) is part of the library, so
# a caller written against the reference's SimLib::EkgSim links libekgsim.so
# alone.
HOST_SRC = ["ekg_engine.cpp", "ekg_capi.cpp", "ekg_server.cpp", "ekg_model.cpp", "ini.cpp", "sim_lib.cpp"]
CLI_SRC = ["ekgsim_app.cpp", "ekgsim_main.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def _compile(src, device):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + (DEVICE if device else HOST + ["-x", "c++"]) + ["-c", os.path.join(CSRC, src), "-o", obj]
    _run(cmd)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    cli_src = [s for s in CLI_SRC if os.path.exists(os.path.join(CSRC, s))]
    jobs = [(s, True) for s in KERNEL_SRC] + [(s, False) for s in HOST_SRC + cli_src]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = dict(zip([j[0] for j in jobs], ex.map(lambda j: _compile(*j), jobs)))
    lib_objs = [objs[s] for s in KERNEL_SRC + HOST_SRC]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + lib_objs + ["-lpthread"])
    os.replace(tmp, LIB)
    if len(cli_src) == len(CLI_SRC):
        _run([HIPCC, "--offload-arch=" + ARCH, "-o", CLI] + [objs[s] for s in cli_src]
             + ["-L" + os.path.dirname(LIB), "-lekgsim", "-Wl,-rpath,$ORIGIN/../lib", "-lpthread"])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
