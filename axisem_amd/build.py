"""Build steps of the GPU time loop.

* ``build_library()``  -- hipcc for gfx950 -> ``axisem_amd/libaxihip.so``
  (the C-ABI of ``include/axihip.h``).
* ``build_host(src)`` -- the drop-in Fortran host: the reference SOLVER's own
  sources (read in place from ``src``) with its ``time_evol_wave.F90``
  replaced by ``fortran/time_evol_wave_gpu.F90``, linked against
  ``libaxihip.so`` -> ``axisem_amd/fortran/_build/axisem_gpu``.  The private
  helpers the drop-in needs (``create_fluid_free_surface_mask``,
  ``create_absorbing_gamma``, ``energy`` of ``time_evol_wave.F90``;
  ``fast_correct`` of ``attenuation.f90``) are the reference's own: both
  files are compiled a second time under the module names
  ``axh_ref_time_evol_wave`` / ``axh_ref_attenuation`` with every entity
  public (a rename of the module and of its default accessibility, written
  into the build directory only).
* ``build_post(src)`` -- the reference's seismogram post-processing program
  with its convolution replaced by the GPU's (``fortran/_build/xpost_gpu``).
"""
from __future__ import annotations

import glob
import os
import re
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libaxihip.so")
LIB_FMA = os.path.join(PKG, "libaxihip_fma.so")
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


def build_library(out: str = LIB, fused: bool = False) -> str:
    """The library; fused=True builds the variant whose kernels contract
    a*b+c into fused multiply-adds (libaxihip_fma.so): not the reference's
    arithmetic bit for bit, within its seismogram misfit."""
    flags = HIP_FLAGS
    if fused:
        flags = [("-ffp-contract=fast" if f == "-ffp-contract=off" else f) for f in HIP_FLAGS]
        flags.append("-DAXH_FUSED_MULTIPLY_ADD")
    srcs = [os.path.join(PKG, "csrc", "axh_host.cpp"),
            os.path.join(PKG, "csrc", "axh_meshdb.cpp"),
            os.path.join(PKG, "csrc", "axh_device.hip")]
    inc = ["-I", os.path.join(REPO, "include")]

    def run(cmd):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    # the post-processing kernels follow a program without the solver's
    # flush-to-zero (SOLVER/ftz.c is linked into the SOLVER only): their
    # object keeps FP32 denormals
    post_obj = out + ".post.o"
    post_flags = [f for f in HIP_FLAGS if f not in ("-shared", "-fgpu-flush-denormals-to-zero")]
    run([HIPCC, *post_flags, *inc, "-c", os.path.join(PKG, "csrc", "axh_post.hip"),
         "-o", post_obj])
    tmp = out + ".tmp"
    try:
        run([HIPCC, *flags, *inc, *srcs, "-x", "none", post_obj, "-o", tmp])
    finally:
        if os.path.exists(post_obj):
            os.remove(post_obj)
    os.replace(tmp, out)
    return out


def public_module_copy(src: str, new_name: str, out_dir: str) -> str:
    """The reference module in src as module new_name with its default
    accessibility public, written to out_dir (the build directory); the
    statements are otherwise the reference's, unchanged."""
    txt = open(src).read()
    old = re.search(r"^[ \t]*module[ \t]+(\w+)[ \t]*$", txt, re.I | re.M)
    if not old:
        raise RuntimeError("no module statement in " + src)
    name = old.group(1)
    txt, n = re.subn(r"^([ \t]*module[ \t]+)%s([ \t]*)$" % name, r"\g<1>%s\g<2>" % new_name, txt,
                     count=1, flags=re.I | re.M)
    txt = re.sub(r"^([ \t]*end[ \t]+module)[ \t]+%s[ \t]*$" % name, r"\g<1> %s" % new_name, txt,
                 flags=re.I | re.M)
    # the module's default-accessibility statement (a bare `private` before
    # the first `contains`)
    head, sep, rest = txt.partition("\ncontains")
    head, m = re.subn(r"^[ \t]*private[ \t]*$", "  public", head, count=1, flags=re.I | re.M)
    if n != 1 or m != 1 or not sep:
        raise RuntimeError("unexpected layout of " + src)
    os.makedirs(out_dir, exist_ok=True)
    ext = os.path.splitext(src)[1]
    out = os.path.join(out_dir, new_name + ext)
    with open(out, "w") as f:
        f.write(head + sep + rest)
    return out


POST_EXE = os.path.join(HOST_DIR, "xpost_gpu")


def routine_renamed(src: str, name: str, new_name: str, out_dir: str, out_name: str) -> str:
    """The program in src with its subroutine `name` renamed `new_name`
    (definition and end statement; the calls keep `name` and so reach a
    replacement linked beside it), written to out_dir (the build directory);
    the statements are otherwise the reference's, unchanged."""
    txt = open(src).read()
    txt, n = re.subn(r"^([ \t]*subroutine[ \t]+)%s\b" % name, r"\g<1>%s" % new_name, txt,
                     flags=re.I | re.M)
    txt, m = re.subn(r"^([ \t]*end[ \t]+subroutine[ \t]+)%s\b" % name, r"\g<1>%s" % new_name,
                     txt, flags=re.I | re.M)
    if n != 1 or m != 1:
        raise RuntimeError("unexpected layout of " + src)
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, out_name)
    with open(out, "w") as f:
        f.write(txt)
    return out


def build_post(ref_root: str = "/root/reference", jobs: int = 8) -> str:
    """The reference's seismogram post-processing program
    (SOLVER/UTILS/post_processing.F90) with its convolve_with_stf replaced by
    the GPU convolution (fortran/convolve_gpu.f90 -> axh_convolve_traces):
    axisem_amd/fortran/_build/xpost_gpu."""
    from .fortran import fortran_build as fb

    utils = os.path.join(ref_root, "SOLVER", "UTILS")
    if not os.path.isdir(utils):
        raise FileNotFoundError(utils)
    if not os.path.exists(LIB):
        build_library()
    gen = os.path.join(HOST_DIR, "src")
    fdir = os.path.join(PKG, "fortran")
    srcs = [os.path.join(utils, "nc_postroutines.F90"),
            routine_renamed(os.path.join(utils, "post_processing.F90"), "convolve_with_stf",
                            "ref_convolve_with_stf", gen, "post_processing_gpu.F90"),
            os.path.join(fdir, "axh_iface.f90"), os.path.join(fdir, "convolve_gpu.f90")]
    obj = os.path.join(HOST_DIR, "obj_post")
    objs = fb.compile_all(srcs, obj, fflags=["-O2", "-cpp"], jobs=jobs)
    fb.link(objs, POST_EXE, ldflags=["-L" + PKG, "-laxihip", "-Wl,-rpath,$ORIGIN/../.."])
    return POST_EXE


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
    gen = os.path.join(HOST_DIR, "src")
    srcs += [public_module_copy(os.path.join(solver, "time_evol_wave.F90"),
                                "axh_ref_time_evol_wave", gen),
             public_module_copy(os.path.join(solver, "attenuation.f90"),
                                "axh_ref_attenuation", gen)]
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
