"""Recipe: build the oracle and the unmodified reference programs.

TEST INFRASTRUCTURE ONLY (the checker, never the product path).

* ``build_oracle()`` -- gcc on ``oracle/axisem_oracle.c`` ->
  ``oracle/liboracle.so`` (the CPU restatement; -ffp-contract=off, so the
  arithmetic is the reference's operation for operation).
* ``build_reference(src)`` -- amdflang on the reference's own SOLVER and
  MESHER sources, compiled where they lie under ``src`` (nothing is copied),
  serial, no NetCDF, realkind = 4 as shipped -> ``oracle/_ref/axisem`` and
  ``oracle/_ref/xmesh``.  Same flags as the drop-in host build.

Usage: ``python oracle/build_ref.py [--reference /root/reference]``.
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
    return axisem, xmesh


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    a = ap.parse_args()
    print(build_oracle())
    print(*build_reference(a.reference))
