"""The CPU oracle against the reference SOLVER's own outputs.

tests/golden/<case>/ref_seis.npz hold the receiver seismograms written by the
unmodified reference program (oracle/_ref/axisem, realkind = 4) on the
golden cases; the oracle, restated from the Fortran, must reproduce them bit
for bit (the reference prints single-precision values with round-trip
precision, so equality of the parsed float32 values is equality of bits).
"""
import numpy as np
import pytest

import cases
from axisem_amd.state import read_dims
from oracle.oracle import Oracle

# full-length comparisons for a subset, the first steps for the rest, so the
# CPU suite stays within a few minutes
FULL = {"mono_vertforce_newmark", "dipole_thetaforce_newmark_anel", "quad_mttmmpp_symplec4_anel",
        "p50_mono_vertforce_newmark"}
PREFIX_STEPS = 1200


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_oracle_matches_reference(name, state_cache):
    _, st = state_cache(name)
    d = read_dims(st)
    nsteps = d["niter"] if name in FULL else min(d["niter"], PREFIX_STEPS)
    o = Oracle(st)
    seis = o.run(1, nsteps, d["num_rec"], d["seis_it"])
    cfg = cases.CASES[name]
    nc = cases.NCOMP[cfg.source_type]
    ref = np.load(cases.case_dir(name) + "/ref_seis.npz")["seis"]
    mine = seis[::cases.sample_stride(name), :len(cfg.stations_deg), :nc]
    n = min(len(mine), len(ref))
    assert n > 0
    assert np.abs(ref[:n]).max() > 0
    np.testing.assert_array_equal(mine[:n], ref[:n])


def test_recorded_state_dims(state_cache):
    """The drop-in host hands the library the reference's own sizes."""
    _, st = state_cache("dipole_thetaforce_newmark_anel")
    d = read_dims(st)
    assert d["npol"] == 4 and d["src_type"] == 1 and d["anel"] == 1
    assert d["coarse_grained"] == 1 and d["n_sls"] == 5
    # 19 receivers + epicentre, antipode, equator traces
    assert d["num_rec"] == 22
    for line in open(cases.MESH_DIR + "/mesh_params.h"):
        if "nel_solid =" in line:
            assert int(line.split("=")[1].split("!")[0]) == d["nel_solid"]
        if "nel_fluid =" in line:
            assert int(line.split("=")[1].split("!")[0]) == d["nel_fluid"]
