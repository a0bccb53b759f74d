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


def test_stream_prepared_on_arrival_matches_oracle(host_exe, tmp_path):
    """The serving configuration's kinds of requests on the 50 s and 20 s
    meshes (elastic newmark2: vertical force, horizontal force, explosion, a
    moment tensor as its four runs), bounded to short durations, prepared by
    the loop when they arrive (one host kept running per mesh and source
    type, recording every source for the longest duration), co-batched across
    source types, members retiring while others run; every served seismogram
    equals the oracle's bit for bit, and the steps each request runs are the
    ones its own duration needs."""
    import os
    from axisem_amd import inputs
    from axisem_amd.serve import Preparer
    from axisem_amd.state import read_dims
    from oracle.oracle import Oracle

    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    meshes = {m: os.path.join(repo, "data", "meshes", m) for m in ("prem_iso_50s", "prem_iso_20s")}
    for m in meshes.values():
        if not os.path.exists(os.path.join(m, "meshdb.dat0000")):
            pytest.skip("benchmark meshes not built")
    stations = [float(k) for k in range(0, 181, 6)]

    def cfg(src, depth, length, amp=1e20):
        return inputs.RunConfig(src, depth, amp, "newmark2", False, length, stations)
    stream = [(0.0, "prem_iso_50s", cfg("vertforce", 0.0, 60.0)),
              (0.0, "prem_iso_50s", cfg("explosion", 210.0, 45.0)),
              (0.3, "prem_iso_20s", cfg("thetaforce", 29.0, 40.0))]
    # a moment tensor's four runs, prepared ahead (these runs take the GPU a
    # few milliseconds, so runs prepared on arrival seconds apart would not
    # overlap): they arrive together and must share one batch
    mt = [(0.5, "prem_iso_50s", cfg(r, 425.0, 50.0, a))
          for r, a in (("mrr", 0.3e20), ("mtt_p_mpp", 1.2e20), ("mtr", 0.7e20),
                       ("mtp", 2.1e20))]
    stream += [(0.8, "prem_iso_50s", cfg("vertforce", 511.0, 30.0))]
    prep = Preparer(host_exe, str(tmp_path / "prep"), max_length_s=120.0)
    loop = RequestLoop(hbm_budget_bytes=20e9, chunk_steps=37, preparer=prep,
                       log=lambda m: print("[loop]", m, flush=True))
    try:
        ready = [prep.prepare(meshes[m], c) for _, m, c in mt]
        reqs = [loop.submit_config(meshes[m], c, arrival_s=t) for t, m, c in stream]
        reqs += [loop.submit(p, arrival_s=t, steps=n) for (t, _, _), (p, n) in zip(mt, ready)]
        stream += mt
        st = loop.run()
    finally:
        prep.close()
    print(st, flush=True)
    assert st["completed"] == len(stream) and st["rejected"] == 0
    assert st["max_cobatched"] >= 4          # the moment tensor's runs together
    assert st["setup_p50_s"] is not None
    for r, (_, _, c) in zip(reqs, stream):
        d = read_dims(r.state_path)
        # the state is recorded for the recorders' longest duration; the
        # request runs the steps its own duration needs
        assert r.niter == Preparer.steps_for(c.seismogram_length_s, d["deltat"]) < d["niter"]
        o = Oracle(r.state_path)
        ref = o.run(1, r.niter, d["num_rec"], d["seis_it"])
        o.close()
        np.testing.assert_array_equal(r.seis, ref, err_msg="request %d" % r.rid)
