"""Batched wavefields on one mesh (needs an MI355X).

A batch steps several sources of one type on one mesh with one launch per
kernel.  Every member must produce, bit for bit, the seismograms and fields of
its own solo run, and the oracle's.
"""
import numpy as np
import pytest

import cases
from axisem_amd.state import read_dims, read_state
from axisem_amd.solver import Batch, BlowUp, Solver
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

NAMES = ("disp", "velo", "chi", "dchi", "ddchi")


def _same_fields(a, b, what):
    for nm, x, y in zip(NAMES, a, b):
        np.testing.assert_array_equal(x, y, err_msg="%s %s" % (what, nm))


def _solo(path, n):
    with Solver.from_state_file(path) as s:
        seis = s.run(n)
        return seis, s.fields()


@pytest.mark.parametrize("batch", sorted(cases.BATCHES))
def test_batch_members_match_solo_and_oracle(batch, state_cache):
    names = cases.BATCHES[batch]
    print("batch vs solo and oracle:", batch, names, flush=True)
    paths = [state_cache(n)[1] for n in names]
    n = 300
    solo = [_solo(p, n) for p in paths]
    solvers = [Solver.from_state_file(p) for p in paths]
    try:
        with Batch(solvers) as b:
            first = b.run(117)
            second = b.run(n - 117)
        for k, nm in enumerate(names):
            seis = np.concatenate([first[k], second[k]])
            np.testing.assert_array_equal(seis, solo[k][0], err_msg=nm)
            _same_fields(solvers[k].fields(), solo[k][1], nm)
            # the sources differ: so must the members' seismograms
            if k:
                assert not np.array_equal(solo[k][0], solo[0][0])
    finally:
        for s in solvers:
            s.close()
    for k, p in enumerate(paths):
        d = read_dims(p)
        o = Oracle(p)
        oseis = o.run(1, n, d["num_rec"], d["seis_it"])
        np.testing.assert_array_equal(solo[k][0], oseis, err_msg=names[k])
        _same_fields(solo[k][1], o.fields(), names[k] + " oracle")


@pytest.mark.parametrize("batch,width", [("dipole_symplec4_anel", 2),
                                         ("dipole_symplec4_anel", 3),
                                         ("dipole_symplec4_anelfull", 2),
                                         ("dipole_symplec4_anelfull", 3)])
def test_group_widths_equal_solo(batch, width, state_cache):
    """Narrower groups (the kernels for 2 and 3 wavefields, and a remainder
    group) give every member its solo run, bit for bit; also with memory
    variables at all 25 points."""
    names = cases.BATCHES[batch]
    print("group width", width, names, flush=True)
    paths = [state_cache(n)[1] for n in names]
    n = 80
    solo = [_solo(p, n) for p in paths]
    solvers = [Solver.from_state_file(p) for p in paths]
    try:
        with Batch(solvers, group_width=width) as b:
            out = b.run(n)
        for k, nm in enumerate(names):
            np.testing.assert_array_equal(out[k], solo[k][0], err_msg=nm)
            _same_fields(solvers[k].fields(), solo[k][1], nm)
    finally:
        for s in solvers:
            s.close()


def test_batch_continues_solo_runs(state_cache):
    """Members that ran alone first (each on its own device mesh) join one
    batch and continue exactly."""
    names = cases.BATCHES["dipole_symplec4_anel"][:3]
    paths = [state_cache(n)[1] for n in names]
    print("solo then batch:", names, flush=True)
    ref = [_solo(p, 100) for p in paths]
    solvers = [Solver.from_state_file(p) for p in paths]
    try:
        head = [s.run(40) for s in solvers]
        with Batch(solvers) as b:
            mid = b.run(50)
        # after the batch is gone the members run alone again; members 1 and 2
        # keep stepping on member 0's device mesh after member 0 is closed
        tail = [solvers[0].run(10)]
        f0 = solvers[0].fields()
        solvers[0].close()
        tail += [s.run(10) for s in solvers[1:]]
        for k in range(len(solvers)):
            seis = np.concatenate([head[k], mid[k], tail[k]])
            np.testing.assert_array_equal(seis, ref[k][0], err_msg=names[k])
            _same_fields(f0 if k == 0 else solvers[k].fields(), ref[k][1], names[k])
    finally:
        for s in solvers:
            s.close()


def test_batch_of_one_equals_solo(state_cache):
    path = state_cache(cases.BATCHES["mono_newmark_anel"][0])[1]
    ref = _solo(path, 120)
    with Solver.from_state_file(path) as s:
        with Batch([s]) as b:
            seis = b.run(120)[0]
            assert b.last_run_ms() > 0
        np.testing.assert_array_equal(seis, ref[0])
        _same_fields(s.fields(), ref[1], "batch of one")


def test_batch_member_blowup_is_reported(state_cache):
    """The blow-up test of runtime_info (SOLVER/time_evol_wave.F90:1042) is per
    member: forced from the host in member 1 only."""
    names = cases.BATCHES["mono_newmark_anel"][:2]
    paths = [state_cache(n)[1] for n in names]
    st = read_state(paths[1])
    disp = np.zeros(75 * st.dims["nel_solid"], np.float32)
    g = st["igloc_solid"][6] - 1              # point (1,1) of element 1
    mask = (st["igloc_solid"] - 1) == g
    disp[:25 * st.dims["nel_solid"]][mask] = 1e22
    solvers = [Solver.from_state_file(p) for p in paths]
    try:
        solvers[1].set_fields(disp=disp)
        with Batch(solvers) as b:
            with pytest.raises(BlowUp, match="batch member 1"):
                b.run(3)
    finally:
        for s in solvers:
            s.close()


def test_batch_members_at_different_steps(state_cache):
    """Continuous batching: a member joins after others have advanced; each
    member still reproduces its solo run (axh_batch_step)."""
    names = cases.BATCHES["dipole_symplec4_anel"][:3]
    paths = [state_cache(n)[1] for n in names]
    print("staggered batch:", names, flush=True)
    ref = [_solo(p, 150) for p in paths]
    solvers = [Solver.from_state_file(p) for p in paths]
    try:
        got = [[solvers[0].run(30)], [solvers[1].run(70)], []]
        with Batch(solvers) as b:          # members at steps 30, 70 and 0
            for k, seis in enumerate(b.step(50)):
                got[k].append(seis)
        # finish alone: 150 steps each
        for k, s in enumerate(solvers):
            got[k].append(s.run(150 - s.iter))
        for k in range(3):
            np.testing.assert_array_equal(np.concatenate(got[k]), ref[k][0], err_msg=names[k])
            _same_fields(solvers[k].fields(), ref[k][1], names[k])
    finally:
        for s in solvers:
            s.close()


def test_mesh_handle_shares_the_device_mesh(state_cache):
    """A request bound to a resident mesh (axh_mesh_bind) runs as if alone and
    adds only its wavefield part to the device; the footprint estimate made
    without a GPU matches the device allocation."""
    import ctypes as C
    from axisem_amd import solver as S
    names = cases.BATCHES["mono_newmark_anel"][:2]
    paths = [state_cache(n)[1] for n in names]
    ref = _solo(paths[1], 80)
    L = S.library()
    a = Solver.from_state_file(paths[0])
    b = Solver.from_state_file(paths[1])
    try:
        mb, wb = C.c_size_t(), C.c_size_t()
        S._check(L.axh_footprint(a._h, C.byref(mb), C.byref(wb)))
        mesh = C.c_void_p()
        S._check(L.axh_mesh_get(a._h, C.byref(mesh)))
        S._check(L.axh_mesh_bind(b._h, mesh))
        dev_mesh = C.c_size_t()
        S._check(L.axh_mesh_bytes(mesh, C.byref(dev_mesh)))
        seis = b.run(80)
        np.testing.assert_array_equal(seis, ref[0])
        _same_fields(b.fields(), ref[1], "bound")
        # estimate vs allocation (the estimate ignores a few small tables)
        assert abs(dev_mesh.value - mb.value) <= 0.02 * dev_mesh.value
        assert abs(a.device_bytes() - dev_mesh.value - wb.value) <= 0.05 * wb.value
        S._check(L.axh_mesh_release(mesh))
    finally:
        a.close()
        b.close()
