"""Write the kernel-wavefield fixtures: one xdmf snapshot of the reference
program per case of cases.KWF_CASES.

Runs in the build container only (needs oracle/_ref/axisem, the unmodified
reference SOLVER built by oracle/build_ref.py).  The reference writes its
xdmf snapshots (SOLVER/wavefields_io.f90:119-200, glob_snapshot_xdmf) as
big-endian float32 streams, one array of npoint_plot values per snapshot and
component (xdmf_snap_s/p/z_0000.dat): the solid displacement in (s, phi, z)
and the fluid displacement inv_rho_fluid * grad chi, the values the
displ_only kernel-wavefield dump gathers.  Stored per case in
tests/golden/<case>/ref_xdmf.npz: the last snapshot k (the fields after
k * snap_it steps, snap_it = floor(SNAPSHOT_DT / deltat); snapshot 0 holds
the zero initial fields), components s, (p,) z.

Usage: python tests/golden/make_kwf_golden.py [case ...]
"""
from __future__ import annotations

import os
import resource
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)

from axisem_amd import inputs  # noqa: E402
import cases  # noqa: E402

AXISEM = os.path.join(REPO, "oracle", "_ref", "axisem")


def _unlimited_stack():
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run_case(name: str):
    cfg = cases.KWF_CASES[name]
    out = cases.case_dir(name)
    os.makedirs(out, exist_ok=True)
    comps = ["s", "z"] if cases.NCOMP[cfg.source_type] == 2 else ["s", "p", "z"]
    with tempfile.TemporaryDirectory() as tmp:
        inputs.write_run_dir(tmp, cases.mesh_of(name), cfg)
        with open(os.path.join(tmp, "OUTPUT"), "w") as log:
            subprocess.run([AXISEM], cwd=tmp, stdout=log, stderr=subprocess.STDOUT, check=True,
                           preexec_fn=_unlimited_stack, timeout=3600)
        arrs = {c: np.fromfile(os.path.join(tmp, "Data", "xdmf_snap_%s_0000.dat" % c), ">f4")
                for c in comps}
        # two float32 coordinates per plotted point (meshes_io.F90:224-300)
        npoint = os.path.getsize(os.path.join(tmp, "Data", "xdmf_points_0000.dat")) // 8
    n = arrs[comps[0]].size
    nsnap = n // npoint
    if nsnap * npoint != n or nsnap < 2:
        raise RuntimeError("cannot split the snapshot stream: %d values, %d points" % (n, npoint))
    # snapshot k is the fields after k * snap_it steps, snap_it = floor(SNAPSHOT_DT /
    # deltat) (SOLVER/parameters.F90:944); snapshot 0 holds the zero initial fields
    k = nsnap - 1
    data = {"xdmf_" + c: arrs[c].reshape(nsnap, npoint)[k].astype(np.float32) for c in comps}
    np.savez_compressed(os.path.join(out, "ref_xdmf.npz"), snap_index=np.int64(k),
                        snapshot_dt=np.float64(cfg.snapshot_dt), **data)
    print(name, "snapshot", k, "of", nsnap, ";", npoint, "points;", comps)


if __name__ == "__main__":
    names = sys.argv[1:] or sorted(cases.KWF_CASES)
    for nm in names:
        run_case(nm)
