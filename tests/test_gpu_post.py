"""Seismogram post-processing with the GPU convolution (needs an MI355X).

The reference's post-processing program (SOLVER/UTILS/post_processing.F90:
moment-tensor sum with radiation factors, convolution with a source time
function, rotation of the components) runs on the outputs of drop-in GPU
runs twice: as the reference built it (oracle/_ref/xpost) and with its
convolve_with_stf replaced by the GPU convolution
(axisem_amd/fortran/_build/xpost_gpu).  Every output file must be
byte-identical.  Also axh_convolve_traces against a direct evaluation of the
reference's formula on random traces.
"""
import ctypes as C
import hashlib
import os
import resource
import shutil
import subprocess

import numpy as np
import pytest

import cases
from axisem_amd import inputs
from axisem_amd import solver as S

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
XPOST_REF = os.path.join(REPO, "oracle", "_ref", "xpost")
XPOST_GPU = os.path.join(REPO, "axisem_amd", "fortran", "_build", "xpost_gpu")
# the four runs of a moment tensor and the directories the post-processing
# reads them from (SOLVER/UTILS/post_processing.F90:536-541)
MT_DIRS = (("MZZ", "mrr"), ("MXX_P_MYY", "mtt_p_mpp"), ("MXZ_MYZ", "mtr"),
           ("MXY_MXX_M_MYY", "mtp"))
CMT = ("PDE  2011  8 23 17 51  4.60  37.9400  -77.9300   6.0 5.9 5.8 TEST\n"
       "event name:     TEST\ntime shift:      0.0000\nhalf duration:   1.8000\n"
       "latitude:       90.0000\nlongitude:       0.0000\ndepth:         100.0000\n"
       "Mrr:       4.710000e+26\nMtt:       3.810000e+25\nMpp:      -4.740000e+26\n"
       "Mrt:       3.990000e+25\nMrp:      -8.050000e+25\nMtp:      -1.230000e+26\n")


def _stack():
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def _run(exe, cwd, log):
    with open(os.path.join(cwd, log), "w") as f:
        r = subprocess.run([exe], cwd=cwd, stdout=f, stderr=subprocess.STDOUT, timeout=300,
                           preexec_fn=_stack)
    assert r.returncode == 0, open(os.path.join(cwd, log), errors="replace").read()[-2000:]


def _tree_hashes(root):
    out = {}
    for d, _, files in os.walk(root):
        for fn in files:
            p = os.path.join(d, fn)
            with open(p, "rb") as f:
                out[os.path.relpath(p, root)] = hashlib.sha256(f.read()).hexdigest()
    return out


def _post_params(path, comp, conv_period):
    with open(path, "w") as f:
        f.write("REC_COMP_SYS    %s\nCONV_PERIOD     %.1f\nCONV_STF        gauss_0\n"
                "DATA_DIR        './Data_Postprocessing'\nNEGATIVE_TIME   .true.\n"
                "DETAILED_OUTPUT .false.\nLOAD_SNAPS      .false.\n" % (comp, conv_period))


def _both_programs(top, tmp_path):
    """Run the reference and the GPU post-processing on copies of top; the
    output trees must be identical."""
    if not (os.path.exists(XPOST_REF) and os.path.exists(XPOST_GPU)):
        pytest.skip("post-processing programs not built (need the reference tree)")
    trees = {}
    for name, exe in (("ref", XPOST_REF), ("gpu", XPOST_GPU)):
        d = str(tmp_path / name)
        shutil.copytree(top, d, symlinks=True)
        os.makedirs(os.path.join(d, "Data_Postprocessing", "SEISMOGRAMS", "UNPROCESSED"))
        _run(exe, d, "OUTPUT_post")
        trees[name] = _tree_hashes(os.path.join(d, "Data_Postprocessing"))
    assert trees["ref"] and sorted(trees["ref"]) == sorted(trees["gpu"])
    assert any("_conv" in k for k in trees["ref"])
    for k in trees["ref"]:
        assert trees["gpu"][k] == trees["ref"][k], k


@pytest.mark.parametrize("comp", ["enz", "sph"])
def test_moment_tensor_postprocessing_matches_reference(comp, host_exe, tmp_path):
    print("post-processing of a moment tensor (GPU runs), components", comp, flush=True)
    top = str(tmp_path / "mt")
    os.makedirs(top)
    for sub, run in MT_DIRS:
        cfg = inputs.RunConfig(run, 100.0, 1e20, "newmark2", False, 900.0, cases.STATIONS,
                               simulation_type="moment")
        rundir = os.path.join(top, sub)
        inputs.write_run_dir(rundir, cases.MESH_DIR, cfg)
        _run(host_exe, rundir, "OUTPUT")
    with open(os.path.join(top, "inparam_basic"), "w") as f:
        f.write("SIMULATION_TYPE   moment\n")
    with open(os.path.join(top, "CMTSOLUTION"), "w") as f:
        f.write(CMT)
    _post_params(os.path.join(top, "param_post_processing"), comp, 250.0)
    _both_programs(top, tmp_path)


def test_single_source_postprocessing_matches_reference(host_exe, tmp_path):
    print("post-processing of a single explosion (GPU run)", flush=True)
    top = str(tmp_path / "single")
    cfg = inputs.RunConfig("explosion", 300.0, 1e20, "newmark2", False, 900.0, cases.STATIONS)
    inputs.write_run_dir(top, cases.MESH_DIR, cfg)
    _run(host_exe, top, "OUTPUT")
    _post_params(os.path.join(top, "param_post_processing"), "enz", 400.0)
    _both_programs(top, tmp_path)


def _convolve_direct(seis, stf, dt, scale):
    """The reference's convolution loop, restated term by term with numpy
    scalars of the reference's precisions (REAL(4) products, REAL(8) scaling
    and sum, REAL(4) partial sums)."""
    nt, nj = seis.shape[0], stf.shape[0]
    out = np.zeros_like(seis)
    for i in range(1, nt + 1):
        acc = np.zeros(seis.shape[1], np.float32)
        for j in range(1, min(nj, i - 1) + 1):
            p = (seis[i - j - 1] * stf[j - 1]).astype(np.float32)
            acc = (acc.astype(np.float64) + p.astype(np.float64) * dt).astype(np.float32)
        out[i - 1] = (acc.astype(np.float64) * scale).astype(np.float32)
    return out


def test_convolve_traces_equals_the_formula():
    rng = np.random.default_rng(3)
    nt, ntr, nj = 300, 5, 57
    seis = (rng.standard_normal((ntr, nt)) * 1e-7).astype(np.float32)
    stf = (rng.standard_normal(nj) * 0.03).astype(np.float32)
    stf[::7] = 0
    dt, scale = 0.137, 3.1415927410125732
    out = np.zeros_like(seis)
    L = S.library()
    S._check(L.axh_convolve_traces(nt, ntr, S._p(seis), nj, S._p(stf), dt, scale, S._p(out)))
    ref = _convolve_direct(seis.T.copy(), stf, dt, scale).T
    np.testing.assert_array_equal(out, ref)
    # more terms than samples: the sum stops at i - 1
    out2 = np.zeros_like(seis)
    S._check(L.axh_convolve_traces(nt, ntr, S._p(seis), nt + 40,
                                   S._p(np.resize(stf, nt + 40).astype(np.float32)), dt, scale,
                                   S._p(out2)))
    ref2 = _convolve_direct(seis.T.copy(), np.resize(stf, nt + 40).astype(np.float32), dt,
                            scale).T
    np.testing.assert_array_equal(out2, ref2)
