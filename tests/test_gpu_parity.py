"""GPU path against the reference and the oracle (needs an MI355X).

* The drop-in Fortran host (reference setup + GPU time loop) on every golden
  case must write the reference's seismograms bit for bit.
* Driving the library through the C-ABI (individual uploads, and state files)
  must reproduce the oracle's fields bit for bit after N steps, from the
  reference initial state and from random initial fields that exercise the
  fluid, the solid-fluid coupling and the axis in the first steps.
"""
import os
import resource
import subprocess
import zlib

import numpy as np
import pytest

import cases
from axisem_amd import inputs
from axisem_amd.state import read_dims, read_state
from axisem_amd.solver import BlowUp, Solver
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _unlimited_stack():
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_dropin_host_matches_reference(name, host_exe, tmp_path):
    print("drop-in GPU run:", name, flush=True)
    cfg = cases.CASES[name]
    rundir = str(tmp_path / name)
    inputs.write_run_dir(rundir, cases.mesh_of(name), cfg)
    with open(os.path.join(rundir, "OUTPUT"), "w") as log:
        r = subprocess.run([host_exe], cwd=rundir, stdout=log, stderr=subprocess.STDOUT,
                           timeout=300, preexec_fn=_unlimited_stack)
    assert r.returncode == 0, open(os.path.join(rundir, "OUTPUT"), errors="replace").read()[-2000:]
    nc = cases.NCOMP[cfg.source_type]
    seis = inputs.read_seismograms(os.path.join(rundir, "Data"), len(cfg.stations_deg), nc)
    g = np.load(cases.case_dir(name) + "/ref_seis.npz")
    assert seis.shape[0] - 1 == int(g["nsamples"])
    np.testing.assert_array_equal(seis[1::int(g["stride"])], g["seis"])


def _compare_fields(gpu, orc):
    names = ("disp", "velo", "chi", "dchi", "ddchi")
    for nm, a, b in zip(names, gpu, orc):
        np.testing.assert_array_equal(a, b, err_msg=nm)


@pytest.mark.parametrize("name", sorted(cases.OUTPUT_CASES))
def test_dropin_output_files_match_reference(name, host_exe, tmp_path):
    """SAVE_SNAPSHOTS (xdmf, vtk) and SAVE_ENERGY: the drop-in on the GPU
    writes the reference program's snapshot and energy files byte for byte
    (SHA-256 recorded from the reference run by tests/golden/make_golden.py)."""
    import hashlib
    import json
    print("drop-in GPU run with snapshots / energies:", name, flush=True)
    rundir = str(tmp_path / name)
    inputs.write_run_dir(rundir, cases.mesh_of(name), cases.OUTPUT_CASES[name])
    with open(os.path.join(rundir, "OUTPUT"), "w") as log:
        r = subprocess.run([host_exe], cwd=rundir, stdout=log, stderr=subprocess.STDOUT,
                           timeout=300, preexec_fn=_unlimited_stack)
    assert r.returncode == 0, open(os.path.join(rundir, "OUTPUT"), errors="replace").read()[-2000:]
    want = json.load(open(os.path.join(cases.case_dir(name), "outputs.json")))
    data = os.path.join(rundir, "Data")
    got = {}
    for fn in sorted(os.listdir(data)):
        if fn.startswith(cases.OUTPUT_FILES):
            with open(os.path.join(data, fn), "rb") as f:
                got[fn] = hashlib.sha256(f.read()).hexdigest()
    assert sorted(got) == sorted(want)
    for fn in want:
        assert got[fn] == want[fn], fn


# the reference-run cases plus the McLachlan-scheme cases only the oracle runs
ORACLE_RUNS = sorted({**cases.CASES, **cases.ORACLE_CASES})


@pytest.mark.parametrize("name", ORACLE_RUNS)
def test_uploads_match_oracle(name, state_cache):
    print("C-ABI uploads vs oracle:", name, flush=True)
    _, st = state_cache(name)
    d = read_dims(st)
    n = min(d["niter"], 2000)
    with Solver.from_state(read_state(st)) as s:
        seis = s.run(n)
        gf = s.fields()
    o = Oracle(st)
    oseis = o.run(1, n, d["num_rec"], d["seis_it"])
    np.testing.assert_array_equal(seis, oseis)
    _compare_fields(gf, o.fields())


def _random_fields(st, seed):
    """Random fields with every copy of a global point equal (element-local
    layout), as the reference's assembled fields are."""
    rng = np.random.default_rng(seed)
    d = st.dims
    ig_s = st["igloc_solid"] - 1
    ig_f = st["igloc_fluid"] - 1
    nc = 3
    gs = rng.standard_normal((nc, d["nglob_solid"])).astype(np.float32) * 1e-3
    gv = rng.standard_normal((nc, d["nglob_solid"])).astype(np.float32) * 1e-3
    disp = np.concatenate([gs[c][ig_s] for c in range(nc)])
    velo = np.concatenate([gv[c][ig_s] for c in range(nc)])
    if d["src_type"] == 0:      # monopole: no phi component
        ns = 25 * d["nel_solid"]
        disp[ns:2 * ns] = 0
        velo[ns:2 * ns] = 0
    f = [rng.standard_normal(max(d["nglob_fluid"], 1)).astype(np.float32)[ig_f] * 1e-3
         for _ in range(3)]
    return disp, velo, f[0], f[1], f[2]


@pytest.mark.parametrize("name", ORACLE_RUNS)
def test_random_fields_match_oracle(name, state_cache):
    print("random initial fields vs oracle:", name, flush=True)
    _, st_path = state_cache(name)
    st = read_state(st_path)
    fields = _random_fields(st, seed=zlib.crc32(name.encode()))
    n = 25
    with Solver.from_state_file(st_path) as s:
        s.set_fields(*fields)
        s.run(n)
        gf = s.fields()
    o = Oracle(st_path)
    o.set_fields(*fields)
    o.run(1, n, st.dims["num_rec"], st.dims["seis_it"])
    of = o.fields()
    # the fields must have moved away from the initial state everywhere
    assert not np.array_equal(gf[0], fields[0])
    if st.dims["nel_fluid"]:
        assert not np.array_equal(gf[2], fields[2])
    _compare_fields(gf, of)


def test_chunked_run_equals_single_run(state_cache):
    _, st = state_cache("dipole_phiforce_symplec4_anel")
    with Solver.from_state_file(st) as a, Solver.from_state_file(st) as b:
        sa = a.run(60)
        sb = np.concatenate([b.run(7) for _ in range(8)] + [b.run(4)])
        np.testing.assert_array_equal(sa, sb)
        _compare_fields(a.fields(), b.fields())


def test_blowup_is_reported(state_cache):
    """runtime_info's check (SOLVER/time_evol_wave.F90:1042): |disp(1,1,:,:)|
    above 10 |magnitude| stops the run, forced here from the host."""
    _, st_path = state_cache("mono_vertforce_newmark")
    st = read_state(st_path)
    disp = np.zeros(75 * st.dims["nel_solid"], np.float32)
    g = st["igloc_solid"][6] - 1              # point (1,1) of element 1
    mask = (st["igloc_solid"] - 1) == g
    disp[:25 * st.dims["nel_solid"]][mask] = 1e22
    with Solver.from_state_file(st_path) as s:
        s.set_fields(disp=disp)
        with pytest.raises(BlowUp):
            s.run(3)


def test_state_roundtrip_on_device(state_cache, tmp_path):
    """A state written by the library and read back runs identically."""
    _, st_path = state_cache("quad_mtp_newmark")
    st = read_state(st_path)
    with Solver.from_state(st) as a, Solver.from_state_file(st_path) as b:
        np.testing.assert_array_equal(a.run(30), b.run(30))
        assert a.device_bytes() > 0


def test_run_begin_end_equals_run(state_cache):
    """axh_run_begin / axh_run_end (the drop-in host's overlapped chunks) give
    axh_run's samples and fields; the context refuses other calls while a run
    is queued."""
    from axisem_amd import solver as S
    _, st = state_cache("dipole_phiforce_symplec4_anel")
    with Solver.from_state_file(st) as a, Solver.from_state_file(st) as b:
        ref = a.run(61)
        first = b.run_begin(25)
        disp = np.zeros(75 * b.dims["nel_solid"], np.float32)
        rc = b._L.axh_download_fields(b._h, S._p(disp), None, None, None, None)
        assert rc == S.AXH_ERR_STATE
        b.run_end()
        second = b.run_begin(36)
        b.run_end()
        np.testing.assert_array_equal(np.concatenate([first, second]), ref)
        _compare_fields(b.fields(), a.fields())
