"""A source in the fluid (add_source_fl) on the GPU (needs an MI355X).

The reference program cannot place a fluid source on its meshes
(SOLVER/source.f90:531), so the states are recorded ones with their source
moved into fluid elements (tests/fluid_source.py); the oracle's add_source_fl
is pinned by the closed-form check in test_oracle_math.py.  The GPU runs the
injection in the fluid tiles of its stiffness launch and must equal the
oracle bit for bit, alone and as a member of a batch.
"""
import numpy as np
import pytest

from axisem_amd.state import read_dims, read_state, write_state
from axisem_amd.solver import Batch, Solver
from fluid_source import fluid_source_state, pick_fluid_elements
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

NAMES = ("disp", "velo", "chi", "dchi", "ddchi")


def _same(a, b, what):
    for nm, x, y in zip(NAMES, a, b):
        np.testing.assert_array_equal(x, y, err_msg="%s %s" % (what, nm))


def _fluid_state(state_cache, name, tmp_path, nel=3, seed=0):
    _, path = state_cache(name)
    st = read_state(path)
    fs = fluid_source_state(st, pick_fluid_elements(st, nel), seed=seed)
    out = str(tmp_path / ("%s_fluidsrc.bin" % name))
    write_state(fs, out)
    return out


@pytest.mark.parametrize("name", ["mono_vertforce_newmark", "b_mono_explosion_100",
                                  "dipole_thetaforce_newmark_anel"])
def test_fluid_source_matches_oracle(name, state_cache, tmp_path):
    print("fluid source vs oracle:", name, flush=True)
    p = _fluid_state(state_cache, name, tmp_path, nel=3 if name != "b_mono_explosion_100" else 8)
    d = read_dims(p)
    assert d["fluid_src"] == 1
    n = 300
    with Solver.from_state(read_state(p)) as s:
        seis = s.run(n)
        gf = s.fields()
    o = Oracle(p)
    oseis = o.run(1, n, d["num_rec"], d["seis_it"])
    of = o.fields()
    np.testing.assert_array_equal(seis, oseis)
    _same(gf, of, name)
    # the source reached the fluid and, through the coupling, the solid
    assert np.abs(of[2]).max() > 0 and np.abs(of[0]).max() > 0


def test_fluid_source_member_in_a_batch(state_cache, tmp_path):
    """A fluid-source wavefield and two solid-source ones in one group: each
    equals the oracle."""
    paths = [_fluid_state(state_cache, "b_mono_explosion_100", tmp_path, nel=2, seed=4),
             state_cache("b_mono_explosion_100")[1], state_cache("b_mono_vertforce_0")[1]]
    n = 200
    solvers = [Solver.from_state_file(p) for p in paths]
    try:
        with Batch(solvers) as b:
            out = b.run(n)
        got = [s.fields() for s in solvers]
    finally:
        for s in solvers:
            s.close()
    for k, p in enumerate(paths):
        d = read_dims(p)
        o = Oracle(p)
        np.testing.assert_array_equal(out[k], o.run(1, n, d["num_rec"], d["seis_it"]))
        _same(got[k], o.fields(), "member %d" % k)
