"""Recipe: build the oracle and the unmodified reference programs.

TEST INFRASTRUCTURE ONLY (the checker, never the product path).

* ``build_oracle()`` -- gcc on ``oracle/axisem_oracle.c`` ->
  ``oracle/liboracle.so`` (the CPU restatement; -ffp-contract=off, so the
  arithmetic is the reference's operation for operation).
* ``build_reference(src)`` -- amdflang on the reference's own SOLVER and
  MESHER sources, compiled where they lie under ``src`` (nothing is copied),
  serial, no NetCDF, realkind = 4 as shipped -> ``oracle/_ref/axisem`` and
  ``oracle/_ref/xmesh``.  Same flags as the drop-in host build.  Also the
  reference's seismogram post-processing program
  (SOLVER/UTILS/post_processing.F90, no NetCDF) -> ``oracle/_ref/xpost``.
* ``build_reference_fp64(src)`` -- the same programs in double precision,
  as the reference's own comment directs ("Choose solver precision here",
  SOLVER/global_parameters.f90:40, MESHER/global_parameters.f90:40, and
  smallval "Change for dp", :47).  The reference's double-precision build
  does not pass a strict compiler's argument checks in five output-only
  places, so copies of the changed files, written to
  ``oracle/_ref/src_fp64/``, carry these edits (FP64_EDITS) and every other
  source compiles in place:
    - global_parameters.f90 (SOLVER, MESHER): realkind = dp, smallval =
      smallval_dble;
    - SOLVER/nc_routines.F90:473,533: the arguments of nc_dump_stf and
      nc_dump_rec (bodies empty without NetCDF) are real(realkind), as their
      callers pass (SOLVER/source.f90:184, seismograms.f90:818);
    - SOLVER/wavefields_io.f90:134-136: the xdmf snapshot work arrays are
      real(realkind), as the routines they are passed to declare;
    - SOLVER/parameters.F90:1542: a NetCDF attribute is passed as real();
    - MESHER/parallelization.f90:226: the domain-decomposition VTK work array
      is real(sp), as the VTK writers it is passed to declare.
  None of these changes what a seismogram-only run without NetCDF computes.
  -> ``oracle/_ref/axisem_fp64``, ``oracle/_ref/xmesh_fp64``.

Usage: ``python oracle/build_ref.py [--reference /root/reference] [--fp64]``.
"""
from __future__ import annotations

import argparse
import glob
import os
import re
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REF_OUT = os.path.join(HERE, "_ref")
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
FFLAGS = ["-O2", "-cpp", "-Dserial", "-Dsolver"]

_MOD = re.compile(r"^\s*module\s+(\w+)\s*(?:!.*)?$", re.I | re.M)
_USE = re.compile(r"^\s*use\s*(?:,\s*\w+\s*)?(?:::)?\s*(\w+)", re.I | re.M)


def _levels(paths):
    prov, uses = {}, {}
    for p in paths:
        t = open(p, errors="replace").read()
        for m in _MOD.finditer(t):
            if m.group(1).lower() != "procedure":
                prov[m.group(1).lower()] = p
        uses[p] = {u.lower() for u in _USE.findall(t)}
    deps = {p: {prov[u] for u in uses[p] if u in prov and prov[u] != p} for p in paths}
    out, done, rest = [], set(), list(paths)
    while rest:
        ready = [p for p in rest if deps[p] <= done]
        if not ready:
            raise RuntimeError("circular module dependencies")
        out.append(ready)
        done.update(ready)
        rest = [p for p in rest if p not in done]
    return out


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))


def _fortran_program(srcs, csrcs, objdir, exe, fflags, ldflags=(), jobs=8):
    os.makedirs(objdir, exist_ok=True)
    objs = []

    def one(p):
        o = os.path.join(objdir, os.path.splitext(os.path.basename(p))[0] + ".o")
        _run(["amdflang", *fflags, "-J", objdir, "-I", objdir, "-c", p, "-o", o])
        return o

    for lev in _levels(srcs):
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            objs.extend(ex.map(one, lev))
    for c in csrcs:
        o = os.path.join(objdir, os.path.splitext(os.path.basename(c))[0] + ".o")
        _run(["gcc", "-O2", "-w", "-c", c, "-o", o])
        objs.append(o)
    _run(["amdflang", *objs, "-o", exe, *ldflags])
    return exe


def build_oracle() -> str:
    src = os.path.join(HERE, "axisem_oracle.c")
    tmp = ORACLE_LIB + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          "-std=c11", src, "-o", tmp, "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_reference(ref_root: str = "/root/reference", jobs: int = 8):
    solver = os.path.join(ref_root, "SOLVER")
    mesher = os.path.join(ref_root, "MESHER")
    if not os.path.isdir(solver):
        raise FileNotFoundError(solver)
    os.makedirs(REF_OUT, exist_ok=True)
    ssrc = sorted(glob.glob(os.path.join(solver, "*.f90")) + glob.glob(os.path.join(solver, "*.F90")))
    axisem = _fortran_program(ssrc, [os.path.join(solver, "ftz.c"), os.path.join(solver, "pthread.c")],
                              os.path.join(REF_OUT, "obj_solver"), os.path.join(REF_OUT, "axisem"),
                              FFLAGS, ["-lpthread"], jobs)
    msrc = sorted(glob.glob(os.path.join(mesher, "*.f90")) + glob.glob(os.path.join(mesher, "*.F90")))
    xmesh = _fortran_program(msrc, [], os.path.join(REF_OUT, "obj_mesher"),
                             os.path.join(REF_OUT, "xmesh"), ["-O2", "-cpp", "-fopenmp"],
                             ["-fopenmp"], jobs)
    utils = os.path.join(solver, "UTILS")
    xpost = _fortran_program([os.path.join(utils, "nc_postroutines.F90"),
                              os.path.join(utils, "post_processing.F90")], [],
                             os.path.join(REF_OUT, "obj_post"), os.path.join(REF_OUT, "xpost"),
                             ["-O2", "-cpp"], [], jobs)
    return axisem, xmesh, xpost


# (file relative to the reference root, [(pattern, replacement)]), each pattern
# matching exactly once
FP64_EDITS = [
    ("SOLVER/global_parameters.f90",
     [(r"(integer, parameter\s*::\s*realkind\s*=\s*)sp\b", r"\g<1>dp"),
      (r"(real\(kind=realkind\), parameter\s*::\s*smallval\s*=\s*)smallval_sngl\b",
       r"\g<1>smallval_dble")]),
    ("MESHER/global_parameters.f90",
     [(r"(integer, parameter\s*::\s*realkind\s*=\s*)sp\b", r"\g<1>dp"),
      (r"(real\(kind=realkind\), parameter\s*::\s*smallval\s*=\s*)smallval_sngl\b",
       r"\g<1>smallval_dble")]),
    ("SOLVER/nc_routines.F90",
     [(r"real\(kind=sp\)(, intent\(in\), dimension\(:\) :: stf)", r"real(kind=realkind)\g<1>"),
      (r"real\(sp\)(, intent\(in\), dimension\(3,num_rec\) :: recfield)",
       r"real(realkind)\g<1>")]),
    ("SOLVER/wavefields_io.f90",
     [(r"real\(sp\)(, allocatable\s+:: u\(:,:\), usz_fl)", r"real(realkind)\g<1>"),
      (r"real\(sp\)(, allocatable\s+:: straintrace\(:,:,:,:\))", r"real(realkind)\g<1>"),
      (r"real\(sp\)(, allocatable\s+:: curlinplane\(:,:,:,:\))", r"real(realkind)\g<1>")]),
    ("SOLVER/parameters.F90",
     [(r"call nc_write_att_real\(shift_fact,", r"call nc_write_att_real(real(shift_fact),")]),
    ("MESHER/parallelization.f90",
     [(r"real\(kind=realkind\)(, dimension\(:\), allocatable :: wel2proc)", r"real(kind=sp)\g<1>")]),
]


def _fp64_sources(ref_root: str, part: str) -> dict:
    """basename -> path of the edited copies of one part (SOLVER or MESHER)."""
    out = {}
    for rel, edits in FP64_EDITS:
        if not rel.startswith(part + "/"):
            continue
        txt = open(os.path.join(ref_root, rel)).read()
        for pat, rep in edits:
            txt, n = re.subn(pat, rep, txt, count=1)
            if n != 1:
                raise RuntimeError("unexpected layout of %s (%s)" % (rel, pat))
        d = os.path.join(REF_OUT, "src_fp64", part.lower())
        os.makedirs(d, exist_ok=True)
        path = os.path.join(d, os.path.basename(rel))
        with open(path, "w") as f:
            f.write(txt)
        out[os.path.basename(rel)] = path
    return out


def build_reference_fp64(ref_root: str = "/root/reference", jobs: int = 8):
    solver = os.path.join(ref_root, "SOLVER")
    mesher = os.path.join(ref_root, "MESHER")
    if not os.path.isdir(solver):
        raise FileNotFoundError(solver)
    progs = []
    for part, exe, cs, fflags, ldflags in (
            ("SOLVER", "axisem_fp64", [os.path.join(solver, "ftz.c"), os.path.join(solver, "pthread.c")],
             FFLAGS, ["-lpthread"]),
            ("MESHER", "xmesh_fp64", [], ["-O2", "-cpp", "-fopenmp"], ["-fopenmp"])):
        d = os.path.join(ref_root, part)
        edited = _fp64_sources(ref_root, part)
        srcs = sorted(glob.glob(os.path.join(d, "*.f90")) + glob.glob(os.path.join(d, "*.F90")))
        srcs = [edited.get(os.path.basename(p), p) for p in srcs]
        progs.append(_fortran_program(srcs, cs, os.path.join(REF_OUT, "obj_%s_fp64" % part.lower()),
                                      os.path.join(REF_OUT, exe), fflags, ldflags, jobs))
    return tuple(progs)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--fp64", action="store_true", help="also the double-precision programs")
    a = ap.parse_args()
    print(build_oracle())
    print(*build_reference(a.reference))
    if a.fp64:
        print(*build_reference_fp64(a.reference))
