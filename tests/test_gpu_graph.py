"""Time steps replayed from a captured HIP graph (needs an MI355X).

With axh_set_graph_steps(S) the library captures S time steps once and
replays the graph for every S steps of a run; every kernel reads its step
from device memory.  Runs of any length, split into calls at any point and
mixed with eager work between replays, must give the seismograms and fields
of the direct launches bit for bit.
"""
import numpy as np
import pytest

import cases
from axisem_amd.solver import Batch, Solver

pytestmark = pytest.mark.gpu

NAMES = ("disp", "velo", "chi", "dchi", "ddchi")


def _same(a, b, what):
    for nm, x, y in zip(NAMES, a, b):
        np.testing.assert_array_equal(x, y, err_msg="%s %s" % (what, nm))


@pytest.mark.parametrize("name,steps", [("mono_vertforce_newmark", 8),
                                        ("dipole_phiforce_symplec4_anel", 5),
                                        ("quad_mtp_newmark_anelfull", 3),
                                        ("dipole_thetaforce_symplec4_anelfull", 4)])
def test_graph_replays_equal_direct_launches(name, steps, state_cache):
    print("graph vs direct:", name, steps, flush=True)
    _, st = state_cache(name)
    with Solver.from_state_file(st) as a:
        ref = np.concatenate([a.run(37), a.run(50), a.run(13)])
        ref_f = a.fields()
    with Solver.from_state_file(st) as g:
        g.set_graph_steps(steps)
        parts = [g.run(37)]
        # eager work between replays: a download of the fields
        g.fields()
        parts.append(g.run(50))
        g.set_graph_steps(0)          # direct launches again
        parts.append(g.run(13))
        np.testing.assert_array_equal(np.concatenate(parts), ref)
        _same(g.fields(), ref_f, name)


def test_graph_replays_in_a_batch(state_cache):
    """A batch of mixed source types (one graph per group) against the same
    batch launched directly; 3 calls of 29 steps, graphs of 6 steps."""
    names = cases.BATCHES["mt4_newmark"]
    paths = [state_cache(n)[1] for n in names]
    out = []
    for gs in (0, 6):
        solvers = [Solver.from_state_file(p) for p in paths]
        try:
            with Batch(solvers, graph_steps=gs) as b:
                seis = [b.run(29) for _ in range(3)]
            out.append(([np.concatenate([s[k] for s in seis]) for k in range(len(names))],
                        [s.fields() for s in solvers]))
        finally:
            for s in solvers:
                s.close()
    for k, nm in enumerate(names):
        np.testing.assert_array_equal(out[1][0][k], out[0][0][k], err_msg=nm)
        _same(out[1][1][k], out[0][1][k], nm)


def test_graph_ring_wraps(state_cache):
    """The full 50 s benchmark run (12 569 steps, a sample every step: the
    2048-sample staging ring wraps six times, copied out between replays)
    with graphs of 64 steps."""
    _, st = state_cache("p50_mono_vertforce_newmark")
    with Solver.from_state_file(st) as a, Solver.from_state_file(st) as g:
        g.set_graph_steps(64)
        n = a.dims["niter"]
        np.testing.assert_array_equal(g.run(n), a.run(n))
        _same(g.fields(), a.fields(), "full run")
