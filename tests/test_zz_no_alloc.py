"""The entry point that allocates nothing (axh_run_noalloc; needs an MI355X).

It runs the loop with its step counter, blow-up flag and receiver-sample
staging in a fixed device array instead of the context's own buffers.  Two
calls in a row, the second shorter, must give the seismograms and fields of
axh_run bit for bit; also with graph replays, and after a batch has used the
same array's GPU.
"""
import ctypes as C

import numpy as np
import pytest

import cases
from axisem_amd import solver as S
from axisem_amd.solver import Batch, Solver

pytestmark = pytest.mark.gpu

NAMES = ("disp", "velo", "chi", "dchi", "ddchi")


def _noalloc(s, nsteps):
    """axh_run_noalloc through the C-ABI for the next nsteps of solver s."""
    L = S.library()
    it0 = s.iter + 1
    it1 = s.iter + nsteps
    n = L.axh_seis_samples(s._h, it0, it1)
    out = np.zeros((max(n, 1), max(s.dims["num_rec"], 1), 3), np.float32)
    S._check(L.axh_run_noalloc(s._h, it0, it1, out.ctypes.data_as(C.c_void_p)))
    s.iter = it1
    return out[:n]


@pytest.mark.parametrize("name,graph", [("dipole_phiforce_symplec4_anel", 0),
                                        ("quad_mtp_newmark", 0),
                                        ("mono_vertforce_newmark", 7)])
def test_noalloc_equals_run(name, graph, state_cache):
    print("no-alloc run vs axh_run:", name, graph, flush=True)
    _, st = state_cache(name)
    with Solver.from_state_file(st) as a, Solver.from_state_file(st) as b:
        # both contexts first run through axh_run (the no-alloc entry point
        # needs a context whose device part exists)
        ref = [a.run(3)]
        got = [b.run(3)]
        b.set_graph_steps(graph)
        ref += [a.run(120), a.run(45)]
        got += [_noalloc(b, 120), _noalloc(b, 45)]   # second call smaller
        for k, (x, y) in enumerate(zip(got, ref)):
            np.testing.assert_array_equal(x, y, err_msg="%s call %d" % (name, k))
        for nm, x, y in zip(NAMES, b.fields(), a.fields()):
            np.testing.assert_array_equal(x, y, err_msg="%s %s" % (name, nm))


def test_noalloc_needs_a_run_first(state_cache):
    _, st = state_cache("mono_vertforce_newmark")
    with Solver.from_state_file(st) as s:
        rc = S.library().axh_run_noalloc(s._h, 1, 5, None)
        assert rc == S.AXH_ERR_STATE
        assert b"has run" in S.library().axh_last_error()


def test_noalloc_contexts_share_the_array(state_cache):
    """Two contexts take turns on the one array, with a batch between: the
    array keeps no state a later call depends on."""
    names = cases.BATCHES["mt4_newmark"]
    paths = [state_cache(n)[1] for n in names]
    _, st = state_cache("dipole_phiforce_symplec4_anel")
    _, st2 = state_cache("mono_vertforce_newmark")
    with Solver.from_state_file(st) as a, Solver.from_state_file(st) as b, \
            Solver.from_state_file(st2) as c, Solver.from_state_file(st2) as d:
        ref, ref2 = a.run(60), d.run(50)
        got, got2 = [b.run(1), _noalloc(b, 20)], [c.run(2)]
        got2.append(_noalloc(c, 30))
        solvers = [Solver.from_state_file(p) for p in paths]
        try:
            with Batch(solvers) as bt:
                bt.run(10)
        finally:
            for s in solvers:
                s.close()
        got.append(_noalloc(b, 39))
        got2.append(_noalloc(c, 18))
        np.testing.assert_array_equal(np.concatenate(got), ref)
        np.testing.assert_array_equal(np.concatenate(got2), ref2)
        for nm, x, y in zip(NAMES, b.fields(), a.fields()):
            np.testing.assert_array_equal(x, y, err_msg=nm)
