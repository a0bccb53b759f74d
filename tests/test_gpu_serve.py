"""The request loop on the GPU (needs an MI355X): requests on two meshes arrive
at different times, are co-batched while they overlap, and each result equals
the request's solo run bit for bit."""
import numpy as np
import pytest

import cases
from axisem_amd.serve import RequestLoop
from axisem_amd.solver import Solver

pytestmark = pytest.mark.gpu


def test_served_results_equal_solo_runs(state_cache):
    names = (cases.BATCHES["dipole_symplec4_anel"][:3] + cases.BATCHES["mono_newmark_anel"][:2])
    paths = [state_cache(n)[1] for n in names]
    solo = []
    for p in paths:
        with Solver.from_state_file(p) as s:
            solo.append(s.run())
    print("serving", len(paths), "requests", flush=True)
    loop = RequestLoop(hbm_budget_bytes=8e9, chunk_steps=40,
                       log=lambda m: print("[loop]", m, flush=True))
    reqs = [loop.submit(p, arrival_s=0.05 * k) for k, p in enumerate(paths)]
    st = loop.run()
    print(st, flush=True)
    assert st["completed"] == len(paths) and st["rejected"] == 0
    assert st["max_cobatched"] >= 2
    for r, ref, nm in zip(reqs, solo, names):
        np.testing.assert_array_equal(r.seis, ref, err_msg=nm)
    assert st["time_to_seismogram_p99_s"] >= st["time_to_seismogram_p50_s"] > 0
