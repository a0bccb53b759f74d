"""The build with fused multiply-adds (libaxihip_fma.so; needs an MI355X).

Its kernels contract a*b+c into one FMA, so its results are not the
reference's bit for bit; they must stay within the seismogram misfit budget
against the reference FP32 runs: relative L2 misfit <= 1e-4 over every
receiver, component and stored sample of every golden case (the reference's
own seismograms), and the same against the bit-exact build for the members
of a batch of four.
"""
import numpy as np
import pytest

import cases
from axisem_amd.solver import Batch, Solver

pytestmark = pytest.mark.gpu

TOL = 1e-4      # relative L2 misfit budget against the reference FP32 run


def _rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_fused_build_within_misfit_of_reference(name, state_cache):
    _, st = state_cache(name)
    g = np.load(cases.case_dir(name) + "/ref_seis.npz")
    n, stride = int(g["nsamples"]), int(g["stride"])
    nc = cases.NCOMP[cases.CASES[name].source_type]
    with Solver.from_state_file(st, fused=True) as s:
        seis = s.run()
    # ref_seis holds every stride-th of the n samples after t = 0, starting
    # with the first; the run returns those n samples
    assert seis.shape[0] == n
    got = seis[::stride, :g["seis"].shape[1], :nc]
    ref = g["seis"][:, :, :nc]
    assert got.shape == ref.shape
    m = _rel_l2(got, ref)
    print("%s: relative L2 misfit of the fused build %.2e" % (name, m), flush=True)
    assert m <= TOL


def test_fused_batch_within_misfit_of_exact_build(state_cache):
    names = cases.BATCHES["dipole_symplec4_anel"]
    paths = [state_cache(n)[1] for n in names]
    exact, fused = [], []
    for fl, out in ((False, exact), (True, fused)):
        members = [Solver.from_state_file(p, fused=fl) for p in paths]
        try:
            with Batch(members) as b:
                out.extend(b.run())
        finally:
            for m in members:
                m.close()
    misfits = []
    for nm, a, b in zip(names, fused, exact):
        m = _rel_l2(a, b)
        print("%s: relative L2 misfit, fused vs exact batch %.2e" % (nm, m), flush=True)
        assert m <= TOL
        misfits.append(m)
    # the fused build does round differently (it is not the exact one loaded twice)
    assert max(misfits) > 0.0
