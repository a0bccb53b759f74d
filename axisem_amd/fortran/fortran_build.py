"""Compile a set of Fortran sources in module-dependency order.

Small build driver used by the drop-in host build (the reference SOLVER's
own setup code with the GPU time_evol_wave module) and by the oracle's
reference build.  It scans `module` / `use` statements, compiles
independent files in parallel, and links.  No cmake or external build system.
"""
from __future__ import annotations

import os
import re
import subprocess
from concurrent.futures import ThreadPoolExecutor

_MOD = re.compile(r"^\s*module\s+(\w+)\s*(?:!.*)?$", re.I | re.M)
_USE = re.compile(r"^\s*use\s*(?:,\s*\w+\s*)?(?:::)?\s*(\w+)", re.I | re.M)


def dependency_levels(paths):
    """Group source paths into levels; every file only uses modules of
    earlier levels (or intrinsic/unknown modules)."""
    provides, uses = {}, {}
    for p in paths:
        txt = open(p, errors="replace").read()
        for m in _MOD.finditer(txt):
            name = m.group(1).lower()
            if name != "procedure":
                provides[name] = p
        uses[p] = {u.lower() for u in _USE.findall(txt)}
    deps = {p: {provides[u] for u in uses[p] if u in provides and provides[u] != p}
            for p in paths}
    level, done = [], set()
    remaining = list(paths)
    while remaining:
        ready = [p for p in remaining if deps[p] <= done]
        if not ready:
            raise RuntimeError("circular module dependencies: %s" % remaining)
        level.append(ready)
        done.update(ready)
        remaining = [p for p in remaining if p not in done]
    return level


def compile_all(paths, objdir, fc="amdflang", fflags=(), jobs=8, log=print):
    os.makedirs(objdir, exist_ok=True)
    objs = []

    def one(p):
        base = os.path.splitext(os.path.basename(p))[0]
        obj = os.path.join(objdir, base + ".o")
        cmd = [fc, *fflags, "-J", objdir, "-I", objdir, "-c", p, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
        return obj

    for lev in dependency_levels(paths):
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            objs.extend(ex.map(one, lev))
    return objs


def compile_c(paths, objdir, cc="gcc", cflags=()):
    objs = []
    for p in paths:
        obj = os.path.join(objdir, os.path.splitext(os.path.basename(p))[0] + ".o")
        cmd = [cc, *cflags, "-c", p, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stderr))
        objs.append(obj)
    return objs


def link(objs, exe, fc="amdflang", ldflags=()):
    cmd = [fc, *objs, "-o", exe, *ldflags]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    return exe
