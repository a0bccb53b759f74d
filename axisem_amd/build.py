"""Build steps of the GPU time loop.

* ``build_library()``  -- hipcc for gfx950 -> ``axisem_amd/libaxihip.so``
  (the C-ABI of ``include/axihip.h``).
* ``build_host(src)`` -- the drop-in Fortran host: the reference SOLVER's own
  sources (read in place from ``src``, never copied) with its
  ``time_evol_wave.F90`` replaced by ``fortran/time_evol_wave_gpu.F90``,
  linked against ``libaxihip.so`` -> ``axisem_amd/fortran/_build/axisem_gpu``.
"""
from __future__ import annotations

import glob
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libaxihip.so")
HOST_DIR = os.path.join(PKG, "fortran", "_build")
HOST_EXE = os.path.join(HOST_DIR, "axisem_gpu")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # reference arithmetic: no contraction into FMA, FP32 denormals flushed
    # like SOLVER/ftz.c
    "-ffp-contract=off", "-fgpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# Flags of the Fortran host; the oracle's reference build uses the same so
# that the precomputed terms are bit-identical between the two programs.
FFLAGS = ["-O2", "-cpp", "-Dserial", "-Dsolver"]


def build_library(out: str = LIB) -> str:
    srcs = [os.path.join(PKG, "csrc", "axh_host.cpp"),
            os.path.join(PKG, "csrc", "axh_device.hip")]
    tmp = out + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(REPO, "include"), *srcs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, out)
    return out


def build_host(ref_root: str = "/root/reference", jobs: int = 8) -> str:
    """Drop-in host; needs the reference tree (present only where it is built)."""
    from .fortran import fortran_build as fb

    solver = os.path.join(ref_root, "SOLVER")
    if not os.path.isdir(solver):
        raise FileNotFoundError(solver)
    if not os.path.exists(LIB):
        build_library()
    srcs = sorted(glob.glob(os.path.join(solver, "*.f90")) +
                  glob.glob(os.path.join(solver, "*.F90")))
    srcs = [s for s in srcs if os.path.basename(s) != "time_evol_wave.F90"]
    fdir = os.path.join(PKG, "fortran")
    srcs += [os.path.join(fdir, "axh_iface.f90"), os.path.join(fdir, "time_evol_wave_gpu.F90")]
    obj = os.path.join(HOST_DIR, "obj")
    objs = fb.compile_all(srcs, obj, fflags=FFLAGS, jobs=jobs)
    flang_inc = "/opt/rocm/llvm/include/flang"
    objs += fb.compile_c([os.path.join(solver, "ftz.c"), os.path.join(solver, "pthread.c")],
                         obj, cflags=["-O2", "-w"])
    objs += fb.compile_c([os.path.join(fdir, "axh_host_glue.c")], obj,
                         cflags=["-O2", "-I", flang_inc, "-I", os.path.join(REPO, "include")])
    fb.link(objs, HOST_EXE, ldflags=["-L" + PKG, "-laxihip", "-lpthread",
                                     "-Wl,-rpath,$ORIGIN/../.."])
    return HOST_EXE
