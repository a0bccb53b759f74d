"""Kernel-wavefield snapshots on the GPU (dump_type displ_only; needs an MI355X).

The states are recorded by the drop-in host with AXH_KWF_STRAIN_IT set: they
carry the reference's own kernel-wavefield grid (build_kwf_grid) and the
fluid terms the reference computes for it.  Every snapshot the GPU gathers
(axh_run_kwf) must equal the oracle's displ_only sample of the same step bit
for bit (the oracle's sample is pinned against the reference program's xdmf
snapshots in tests/test_kwf.py):

* monopole, dipole and quadrupole sources, Newmark and symplectic, with
  attenuation, the source elements masked and not;
* runs split into several calls, launched directly and as replayed graphs,
  and a device ring of two snapshots (the run is cut and copied out many
  times);
* the 50 s benchmark mesh;
* the drop-in host writing the snapshots while it writes the reference's
  seismograms bit for bit.
"""
import os
import resource
import subprocess

import numpy as np
import pytest

import cases
from axisem_amd import inputs
from axisem_amd.solver import Solver
from axisem_amd.state import read_dims, read_state
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRAIN_IT = 7


def _unlimited_stack():
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


@pytest.fixture(scope="module")
def kwf_states(tmp_path_factory, host_exe):
    base = tmp_path_factory.mktemp("kwf_states")
    done = {}

    def get(name, mesh_dir=None, cfg=None):
        if name not in done:
            rundir = str(base / name)
            cfg = cfg or cases.ALL_CONFIGS[name]
            inputs.write_run_dir(rundir, mesh_dir or cases.mesh_of(name), cfg)
            done[name] = inputs.record_state(host_exe, rundir, os.path.join(rundir, "state.bin"),
                                             kwf_strain_it=STRAIN_IT)
        return done[name]
    return get


def _oracle_snapshots(stp, steps, strain_it, src_mask):
    """The oracle's displ_only samples at iter 0, strain_it, ... <= steps, on
    the state's (the reference's) grid."""
    st = read_state(stp)
    d = read_dims(stp)
    grid = (np.asarray(st["kwf_mapping"], np.int32), np.asarray(st["kwf_mask"], np.int32),
            *(int(x) for x in st["kwf_npoints"]))
    o = Oracle(stp)
    out = [o.kwf_sample(grid, src_mask)]
    it = 0
    while it + strain_it <= steps:
        o.run(it + 1, it + strain_it, d["num_rec"], d["seis_it"])
        it += strain_it
        out.append(o.kwf_sample(grid, src_mask))
    o.close()
    return np.stack(out)


def _gpu_snapshots(stp, pieces, strain_it, src_mask, graph=0, ring=0):
    with Solver.from_state_file(stp) as s:
        s.set_kwf(strain_it, src_mask, ring)
        if graph:
            s.set_graph_steps(graph)
        snaps, seis = [], []
        for n in pieces:
            a, k = s.run_kwf(n)
            seis.append(a)
            snaps.append(k)
    return np.concatenate(snaps), np.concatenate(seis)


CASES = ("dipole_thetaforce_newmark_anel", "mono_explosion_symplec4_anel", "quad_mtp_newmark",
         # regional mesh with absorbing boundaries; transversely isotropic PREM
         "reg_dipole_thetaforce_symplec4_anel", "ani_dipole_mtr_symplec4_anel")


@pytest.mark.parametrize("name", CASES)
def test_gpu_snapshots_equal_oracle(name, kwf_states):
    stp = kwf_states(name)
    steps = 60
    print("kernel wavefields, GPU vs oracle:", name, flush=True)
    ref = _oracle_snapshots(stp, steps, STRAIN_IT, True)
    got, _ = _gpu_snapshots(stp, [steps], STRAIN_IT, True)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))
    d = read_dims(stp)
    if d["nel_fluid"] > 0:
        nps = int(read_state(stp)["kwf_npoints"][0])
        assert np.abs(got[-1, :, :nps]).max() > 0


@pytest.mark.parametrize("mode", ["pieces_direct", "pieces_graph_ring2", "unmasked_graph"])
def test_gpu_snapshot_runs_split_and_replayed(mode, kwf_states):
    """The same snapshots whether the run is one call or several, launched
    directly or from replayed graphs (the gather in every captured step), with
    the ring cut to two snapshots; and with the source elements kept."""
    name = "dipole_thetaforce_newmark_anel"
    stp = kwf_states(name)
    steps = 75
    mask = mode != "unmasked_graph"
    print("kernel wavefields:", mode, flush=True)
    ref = _oracle_snapshots(stp, steps, STRAIN_IT, mask)
    plain, plain_seis = _gpu_snapshots(stp, [steps], STRAIN_IT, mask)
    if mode == "pieces_direct":
        got, seis = _gpu_snapshots(stp, [20, 1, 33, 21], STRAIN_IT, mask)
    elif mode == "pieces_graph_ring2":
        got, seis = _gpu_snapshots(stp, [40, 35], STRAIN_IT, mask, graph=8, ring=2)
    else:
        got, seis = _gpu_snapshots(stp, [steps], STRAIN_IT, mask, graph=16)
    np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))
    np.testing.assert_array_equal(plain.view(np.int32), ref.view(np.int32))
    np.testing.assert_array_equal(seis, plain_seis)
    with Solver.from_state_file(stp) as s:
        np.testing.assert_array_equal(s.run(steps), plain_seis)


def test_queued_kwf_runs_equal_one_run(kwf_states):
    """axh_run_kwf_begin / axh_run_end in chunks (the host free between them)
    give the snapshots and samples of one axh_run_kwf."""
    stp = kwf_states("dipole_thetaforce_newmark_anel")
    ref, ref_seis = _gpu_snapshots(stp, [75], STRAIN_IT, True)
    snaps, seis = [], []
    with Solver.from_state_file(stp) as s:
        s.set_kwf(STRAIN_IT)
        s.set_graph_steps(8)
        for n in (30, 45):
            a, k = s.run_kwf_begin(n)
            s.run_end()
            seis.append(a)
            snaps.append(k)
    np.testing.assert_array_equal(np.concatenate(snaps).view(np.int32), ref.view(np.int32))
    np.testing.assert_array_equal(np.concatenate(seis), ref_seis)


def test_gpu_snapshots_all_points_radius_limited(kwf_states):
    """The reference's grid over all 25 points of the elements between 2000 and
    5000 km radius (KERNEL_IBEG..JEND 0..4, KERNEL_RMIN/RMAX): points shared by
    several elements, solid and fluid, take the value of the element that
    numbers them."""
    name = "dipole_thetaforce_newmark_anel"
    cfg = inputs.RunConfig(**{**cases.ALL_CONFIGS[name].__dict__, "extra_advanced":
                              "KERNEL_IBEG 0\nKERNEL_IEND 4\nKERNEL_JBEG 0\nKERNEL_JEND 4\n"
                              "KERNEL_RMIN 2000\nKERNEL_RMAX 5000\n"})
    stp = kwf_states("all_points_r2000_5000", cases.mesh_of(name), cfg)
    st = read_state(stp)
    nps, npf = (int(x) for x in st["kwf_npoints"])
    assert nps > 0 and npf > 0
    steps = 50
    print("kernel wavefields, all points of a radius range", flush=True)
    ref = _oracle_snapshots(stp, steps, STRAIN_IT, True)
    got, _ = _gpu_snapshots(stp, [steps], STRAIN_IT, True, graph=8)
    np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))


def test_gpu_snapshots_50s_mesh(kwf_states):
    mesh = os.path.join(cases.HERE, "mesh_prem_iso_50s")
    cfg = inputs.RunConfig("thetaforce", 50.0, 1e20, "symplec4", True, 600.0,
                           [3.0 * k for k in range(61)])
    stp = kwf_states("p50_dipole_anel", mesh, cfg)
    steps = 30
    print("kernel wavefields on the 50 s mesh", flush=True)
    ref = _oracle_snapshots(stp, steps, STRAIN_IT, True)
    got, _ = _gpu_snapshots(stp, [steps], STRAIN_IT, True, graph=8)
    np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))


def test_dropin_host_writes_snapshots(host_exe, tmp_path):
    """AXH_KWF_STRAIN_IT: the drop-in writes Data/kwf_disp_<s|p|z>.bin and its
    seismograms stay the reference's bit for bit."""
    name = "dipole_thetaforce_newmark_anel"
    cfg = cases.CASES[name]
    rundir = str(tmp_path / name)
    inputs.write_run_dir(rundir, cases.mesh_of(name), cfg)
    kit = 97
    env = dict(os.environ, AXH_KWF_STRAIN_IT=str(kit),
               AXH_SAVE_STATE=os.path.join(rundir, "state.bin"))
    print("drop-in GPU run with kernel wavefields:", name, flush=True)
    with open(os.path.join(rundir, "OUTPUT"), "w") as log:
        r = subprocess.run([host_exe], cwd=rundir, stdout=log, stderr=subprocess.STDOUT,
                           timeout=300, preexec_fn=_unlimited_stack, env=env)
    assert r.returncode == 0, open(os.path.join(rundir, "OUTPUT"), errors="replace").read()[-2000:]
    nc = cases.NCOMP[cfg.source_type]
    seis = inputs.read_seismograms(os.path.join(rundir, "Data"), len(cfg.stations_deg), nc)
    g = np.load(cases.case_dir(name) + "/ref_seis.npz")
    np.testing.assert_array_equal(seis[1::int(g["stride"])], g["seis"])
    stp = os.path.join(rundir, "state.bin")
    d = read_dims(stp)
    st = read_state(stp)
    npt = int(sum(st["kwf_npoints"]))
    files = [np.fromfile(os.path.join(rundir, "Data", "kwf_disp_%s.bin" % c), np.float32)
             for c in ("s", "p", "z")]
    nsnap = d["niter"] // kit + 1
    got = np.stack([f.reshape(nsnap, npt) for f in files], axis=1)
    # the oracle for the first 400 steps (the rest of the run follows the same path)
    n = 400 // kit + 1
    ref = _oracle_snapshots(stp, (n - 1) * kit, kit, True)
    np.testing.assert_array_equal(got[:n].view(np.int32), ref.view(np.int32))
    assert np.abs(got[-1]).max() > 0


def _solo_snapshots(stp, pieces, strain_it, graph=0):
    with Solver.from_state_file(stp) as s:
        s.set_kwf(strain_it)
        if graph:
            s.set_graph_steps(graph)
        return np.concatenate([s.run_kwf(n)[1] for n in pieces])


@pytest.mark.parametrize("batch", ["mt4_symplec4_anel", "dipole_symplec4_anel", "mt16_newmark"])
def test_batch_snapshots_equal_solo_runs(batch, kwf_states):
    """axh_batch_step_kwf: the moment tensor's four runs (mixed source types),
    a group of four of one type, and the 16 runs at four depths, stepped as one
    batch in several calls with graphs: every member's snapshots equal its solo
    run's bit for bit (and, for the first member, the oracle's)."""
    from axisem_amd.solver import Batch
    names = cases.BATCHES[batch]
    paths = [kwf_states(n) for n in names]
    pieces = [23, 31]
    print("kernel wavefields of batch", batch, flush=True)
    members = [Solver.from_state_file(p) for p in paths]
    try:
        for m in members:
            m.set_kwf(STRAIN_IT)
        got = [[] for _ in members]
        with Batch(members, graph_steps=8) as b:
            for n in pieces:
                _, ks = b.step_kwf(n)
                for i, k in enumerate(ks):
                    got[i].append(k)
    finally:
        for m in members:
            m.close()
    for i, (nm, p) in enumerate(zip(names, paths)):
        g = np.concatenate(got[i])
        ref = _solo_snapshots(p, [sum(pieces)], STRAIN_IT)
        assert g.shape == ref.shape, nm
        np.testing.assert_array_equal(g.view(np.int32), ref.view(np.int32), err_msg=nm)
    ref0 = _oracle_snapshots(paths[0], sum(pieces), STRAIN_IT, True)
    np.testing.assert_array_equal(np.concatenate(got[0]).view(np.int32), ref0.view(np.int32))


def test_batch_snapshots_members_at_different_steps(kwf_states):
    """A member that joins the batch 10 steps ahead of the others: each
    member's snapshots continue its solo run's."""
    from axisem_amd.solver import Batch
    names = cases.BATCHES["dipole_symplec4_anel"]
    paths = [kwf_states(n) for n in names]
    members = [Solver.from_state_file(p) for p in paths]
    try:
        for m in members:
            m.set_kwf(STRAIN_IT)
        first = members[0].run_kwf(10)[1]
        with Batch(members) as b:
            _, ks = b.step_kwf(30)
    finally:
        for m in members:
            m.close()
    ref0 = _solo_snapshots(paths[0], [10, 30], STRAIN_IT)
    np.testing.assert_array_equal(np.concatenate([first, ks[0]]).view(np.int32),
                                  ref0.view(np.int32))
    for i in range(1, len(names)):
        ref = _solo_snapshots(paths[i], [30], STRAIN_IT)
        np.testing.assert_array_equal(ks[i].view(np.int32), ref.view(np.int32),
                                      err_msg=names[i])
