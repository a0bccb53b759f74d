"""Misfit of the single-precision runs against the reference built in double
precision (realkind = dp), per golden case.

Runs in the build container only (needs the reference programs of
oracle/build_ref.py, including ``build_reference_fp64``):

1. the double-precision MESHER (oracle/_ref/xmesh_fp64) writes each golden
   mesh again (the mesh database carries the spectral operators in the
   solver's precision, so the double-precision SOLVER needs its own);
2. the double-precision SOLVER (oracle/_ref/axisem_fp64) runs every case of
   cases.CASES on the same inputs;
3. its seismograms are compared with the single-precision reference's
   (tests/golden/<case>/ref_seis.npz, the samples kept there): relative L2
   misfit ||s32 - s64|| / ||s64|| over every kept sample, receiver and
   component, and the largest per-trace relative L2 misfit.

The GPU loop writes the single-precision reference's seismograms bit for bit
(tests/test_gpu_parity.py), so its misfit against the double-precision
reference is the same number.  Results: tests/golden/fp64_misfit.json.

Usage: python tests/golden/make_fp64_misfit.py [case ...]
"""
from __future__ import annotations

import json
import os
import resource
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)

from axisem_amd import inputs  # noqa: E402
import cases  # noqa: E402

XMESH64 = os.path.join(REPO, "oracle", "_ref", "xmesh_fp64")
AXISEM64 = os.path.join(REPO, "oracle", "_ref", "axisem_fp64")
OUT = os.path.join(HERE, "fp64_misfit.json")


def _unlimited_stack():
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def mesh64(work: str, name: str) -> str:
    """The double-precision MESHER's version of the case's mesh."""
    src = cases.mesh_of(name)
    key = os.path.basename(src)
    out = os.path.join(work, key)
    if os.path.exists(os.path.join(out, "meshdb.dat0000")):
        return out
    period = {cases.MESH_DIR: cases.MESH_PERIOD, cases.REGIONAL_MESH_DIR: cases.MESH_PERIOD,
              cases.MESH50_DIR: cases.MESH50_PERIOD, cases.ANI_MESH_DIR: cases.MESH_PERIOD}[src]
    colat = cases.REGIONAL_MAX_COLAT if src == cases.REGIONAL_MESH_DIR else None
    model = "prem_ani" if src == cases.ANI_MESH_DIR else "prem_iso"
    inputs.run_mesher(XMESH64, out, period, model, max_colat=colat)
    return out


def read_seis64(data_dir: str, nrec: int, ncomp: int) -> np.ndarray:
    cols = []
    for k in range(nrec):
        a = np.loadtxt(os.path.join(data_dir, "R%03d_AX_disp.dat" % k), dtype=np.float64, ndmin=2)
        cols.append(a[:, :ncomp])
    return np.stack(cols, axis=1)


def misfit(name: str, work: str) -> dict:
    cfg = cases.CASES[name]
    mdir = mesh64(work, name)
    with tempfile.TemporaryDirectory(dir=work) as tmp:
        inputs.write_run_dir(tmp, mdir, cfg)
        with open(os.path.join(tmp, "OUTPUT"), "w") as log:
            subprocess.run([AXISEM64], cwd=tmp, stdout=log, stderr=subprocess.STDOUT, check=True,
                           preexec_fn=_unlimited_stack, timeout=7200)
        nc = cases.NCOMP[cfg.source_type]
        s64 = read_seis64(os.path.join(tmp, "Data"), len(cfg.stations_deg), nc)
    g = np.load(os.path.join(cases.case_dir(name), "ref_seis.npz"))
    stride = int(g["stride"])
    s32 = g["seis"].astype(np.float64)
    s64 = s64[1::stride][: len(s32)]
    if s64.shape != s32.shape:
        raise RuntimeError("%s: sample layout differs %s %s" % (name, s64.shape, s32.shape))
    d = s32 - s64
    rel = float(np.linalg.norm(d) / np.linalg.norm(s64))
    per = np.sqrt((d ** 2).sum(axis=0)) / np.maximum(np.sqrt((s64 ** 2).sum(axis=0)), 1e-300)
    live = np.sqrt((s64 ** 2).sum(axis=0)) > 1e-6 * np.sqrt((s64 ** 2).sum(axis=0)).max()
    return {"rel_l2": rel, "max_trace_rel_l2": float(per[live].max()),
            "samples_compared": int(s32.shape[0]), "stride": stride,
            "receivers": int(s32.shape[1]), "components": int(s32.shape[2])}


def main(argv):
    names = argv or sorted(cases.CASES)
    res = json.load(open(OUT)) if os.path.exists(OUT) else {}
    work = os.path.join(REPO, "data", "fp64")
    os.makedirs(work, exist_ok=True)
    # meshes first (shared), then the runs in parallel
    for n in names:
        mesh64(work, n)
    with ThreadPoolExecutor(max_workers=min(8, len(names))) as ex:
        for n, r in zip(names, ex.map(lambda n: misfit(n, work), names)):
            res[n] = r
            print(n, json.dumps(r), flush=True)
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:])
