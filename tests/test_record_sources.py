"""Record mode with several sources (no GPU).

One run of the drop-in host sets up the mesh and model once and records the
state of several sources of one source type (AXH_EXTRA_SOURCES); each state
file must be byte-identical to the one a separate run of the host records for
that source alone (up to the entries the reference leaves uninitialised).
"""
import os

import numpy as np
import pytest

import cases
from axisem_amd import inputs
from axisem_amd.state import read_state


def _cfg(src, depth, amp=1e20, scheme="newmark2", anel=False):
    return inputs.RunConfig(src, depth, amp, scheme, anel, 300.0, cases.STATIONS)


@pytest.mark.parametrize("cfgs", [
    # monopole: moment-tensor runs at several depths and a force
    [_cfg("mrr", 100.0), _cfg("mtt_p_mpp", 300.0, 2e20), _cfg("vertforce", 0.0),
     _cfg("explosion", 600.0)],
    # dipole, anelastic symplectic
    [_cfg("mtr", 10.0, scheme="symplec4", anel=True),
     _cfg("thetaforce", 700.0, 3e20, scheme="symplec4", anel=True)],
], ids=["monopole", "dipole_anel"])
def test_extra_sources_equal_separate_runs(cfgs, host_exe, tmp_path):
    paths = [str(tmp_path / ("multi_%d.bin" % k)) for k in range(len(cfgs))]
    inputs.record_states(host_exe, str(tmp_path / "multi"), cases.MESH_DIR, cfgs, paths)
    for k, cfg in enumerate(cfgs):
        rundir = str(tmp_path / ("solo_%d" % k))
        inputs.write_run_dir(rundir, cases.MESH_DIR, cfg)
        solo = inputs.record_state(host_exe, rundir, os.path.join(rundir, "state.bin"))
        _same_state(read_state(solo), read_state(paths[k]), "source %d (%s)" % (k, cfg.source_type))


def _same_state(a, b, what):
    """Equal dims and arrays; M0_w_fl only on the axial fluid elements (the
    reference leaves the other entries uninitialised,
    SOLVER/def_precomp_terms.f90:2429-2455)."""
    assert a.dims == b.dims, what
    assert sorted(a.arrays) == sorted(b.arrays), what
    for n in a.arrays:
        x, y = np.asarray(a[n]), np.asarray(b[n])
        if n == "M0_w_fl":
            ax = np.asarray(a["axis_fluid"]) != 0
            x, y = x.reshape(-1, 5)[ax], y.reshape(-1, 5)[ax]
        assert x.dtype == y.dtype and x.shape == y.shape, (what, n)
        assert x.tobytes() == y.tobytes(), (what, n)


def test_extra_sources_refuse_mixed_types(host_exe, tmp_path):
    with pytest.raises(ValueError):
        inputs.record_states(host_exe, str(tmp_path / "m"), cases.MESH_DIR,
                             [_cfg("mrr", 100.0), _cfg("mtr", 100.0)],
                             [str(tmp_path / "a"), str(tmp_path / "b")])


def test_recorder_on_a_pipe_equals_separate_runs(host_exe, tmp_path):
    """A host kept running (sources fed through a pipe as requests come)
    records states equal to separate runs."""
    cfgs = [_cfg("mrr", 100.0), _cfg("vertforce", 0.0, 2e20), _cfg("mtt_p_mpp", 600.0)]
    rec = inputs.SourceRecorder(host_exe, str(tmp_path / "rec"), cases.MESH_DIR, cfgs[0],
                                str(tmp_path / "r0.bin"))
    try:
        paths = [rec.first_state] + [rec.record(c, str(tmp_path / ("r%d.bin" % k)))
                                     for k, c in enumerate(cfgs[1:], start=1)]
        with pytest.raises(ValueError):
            rec.record(_cfg("mtr", 100.0), str(tmp_path / "bad.bin"))
    finally:
        rec.close()
    for k, cfg in enumerate(cfgs):
        rundir = str(tmp_path / ("solo_%d" % k))
        inputs.write_run_dir(rundir, cases.MESH_DIR, cfg)
        solo = inputs.record_state(host_exe, rundir, os.path.join(rundir, "state.bin"))
        _same_state(read_state(solo), read_state(paths[k]), "source %d" % k)


@pytest.mark.parametrize("scheme,anel", [("newmark2", False), ("symplec4", True)])
def test_longer_recording_runs_the_shorter_run(scheme, anel, host_exe, tmp_path):
    """A request for a shorter seismogram runs the first steps of a state
    recorded for a longer one: seismograms and fields equal the shorter run's
    (the oracle on both states)."""
    from axisem_amd.serve import Preparer
    from oracle.oracle import Oracle
    short = inputs.RunConfig("mtr", 100.0, 1e20, scheme, anel, 300.0, cases.STATIONS)
    long_ = inputs.RunConfig("mtr", 100.0, 1e20, scheme, anel, 700.0, cases.STATIONS)
    paths = []
    for nm, c in (("short", short), ("long", long_)):
        rundir = str(tmp_path / nm)
        inputs.write_run_dir(rundir, cases.MESH_DIR, c)
        paths.append(inputs.record_state(host_exe, rundir, os.path.join(rundir, "state.bin")))
    ds, dl = read_state(paths[0]).dims, read_state(paths[1]).dims
    n = Preparer.steps_for(300.0, dl["deltat"])
    assert n == ds["niter"] < dl["niter"]
    out = []
    for p in paths:
        o = Oracle(p)
        out.append((o.run(1, n, ds["num_rec"], ds["seis_it"]), o.fields()))
        o.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        np.testing.assert_array_equal(a, b)
