"""Write the golden fixtures: reference seismograms of the cases in cases.py.

Runs in the build container only (needs /root/reference for the reference
programs built by oracle/build_ref.py):

1. the reference MESHER -> tests/golden/mesh_prem_iso_200s/ (meshdb.dat0000),
   the regional mesh (LOCAL_MAX_COLAT 60) -> mesh_prem_iso_200s_colat60/ and
   the 50 s mesh -> mesh_prem_iso_50s/ and the transversely isotropic
   prem_ani 200 s mesh -> mesh_prem_ani_200s/;
2. for every case, the unmodified reference SOLVER (oracle/_ref/axisem) on the
   case's inputs -> receiver seismograms, stored as float32 (the reference
   prints them with round-trip precision) every sample_stride(case) samples in
   tests/golden/<case>/ref_seis.npz, with niter, deltat and the seismogram
   sampling of the run;
3. for every case of OUTPUT_CASES (snapshots, energies), the SHA-256 of the
   reference's output files in tests/golden/<case>/outputs.json.

Usage: python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import os
import resource
import shutil
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

XMESH = os.path.join(REPO, "oracle", "_ref", "xmesh")
AXISEM = os.path.join(REPO, "oracle", "_ref", "axisem")


def _unlimited_stack():
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def make_mesh(mesh_dir, max_colat=None, period=cases.MESH_PERIOD, model="prem_iso"):
    os.makedirs(mesh_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        inputs.run_mesher(XMESH, tmp, period, model, max_colat=max_colat)
        for fn in ("meshdb.dat0000", "mesh_params.h"):
            shutil.copy(os.path.join(tmp, fn), mesh_dir)


def run_reference(name: str):
    cfg = cases.CASES[name]
    out = cases.case_dir(name)
    os.makedirs(out, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        inputs.write_run_dir(tmp, cases.mesh_of(name), cfg)
        with open(os.path.join(tmp, "OUTPUT"), "w") as log:
            subprocess.run([AXISEM], cwd=tmp, stdout=log, stderr=subprocess.STDOUT, check=True,
                           preexec_fn=_unlimited_stack, timeout=3600)
        ncomp = cases.NCOMP[cfg.source_type]
        seis = inputs.read_seismograms(os.path.join(tmp, "Data"), len(cfg.stations_deg), ncomp)
        text = open(os.path.join(tmp, "OUTPUT"), errors="replace").read()
    # the first row is the t = 0 sample written before the loop
    nsamp = seis.shape[0] - 1
    stride = cases.sample_stride(name)
    np.savez_compressed(os.path.join(out, "ref_seis.npz"), seis=seis[1::stride],
                        nsamples=np.int64(nsamp), stride=np.int64(stride))
    for line in text.splitlines():
        if "Time loop routine" in line:
            print(name, line.strip())
    print(name, "samples", nsamp, "receivers", seis.shape[1], "components", ncomp)


def output_hashes(data_dir: str) -> dict:
    import hashlib
    out = {}
    for fn in sorted(os.listdir(data_dir)):
        if fn.startswith(cases.OUTPUT_FILES):
            with open(os.path.join(data_dir, fn), "rb") as f:
                out[fn] = hashlib.sha256(f.read()).hexdigest()
    return out


def run_outputs(name: str):
    """The reference run with snapshots or energies: SHA-256 of every such
    output file."""
    import json
    cfg = cases.OUTPUT_CASES[name]
    out = cases.case_dir(name)
    os.makedirs(out, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        inputs.write_run_dir(tmp, cases.mesh_of(name), cfg)
        with open(os.path.join(tmp, "OUTPUT"), "w") as log:
            subprocess.run([AXISEM], cwd=tmp, stdout=log, stderr=subprocess.STDOUT, check=True,
                           preexec_fn=_unlimited_stack, timeout=3600)
        hashes = output_hashes(os.path.join(tmp, "Data"))
    with open(os.path.join(out, "outputs.json"), "w") as f:
        json.dump(hashes, f, indent=1, sort_keys=True)
    print(name, len(hashes), "output files")


def main(argv):
    if not os.path.exists(os.path.join(cases.MESH_DIR, "meshdb.dat0000")):
        make_mesh(cases.MESH_DIR)
    if not os.path.exists(os.path.join(cases.REGIONAL_MESH_DIR, "meshdb.dat0000")):
        make_mesh(cases.REGIONAL_MESH_DIR, cases.REGIONAL_MAX_COLAT)
    if not os.path.exists(os.path.join(cases.MESH50_DIR, "meshdb.dat0000")):
        make_mesh(cases.MESH50_DIR, period=cases.MESH50_PERIOD)
    if not os.path.exists(os.path.join(cases.ANI_MESH_DIR, "meshdb.dat0000")):
        make_mesh(cases.ANI_MESH_DIR, model="prem_ani")
    names = argv or list(cases.CASES) + list(cases.OUTPUT_CASES)
    for n in names:
        if n in cases.OUTPUT_CASES:
            run_outputs(n)
        else:
            run_reference(n)


if __name__ == "__main__":
    main(sys.argv[1:])
