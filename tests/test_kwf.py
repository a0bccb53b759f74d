"""Kernel-wavefield output, dump_type displ_only (CPU checks).

The reference writes these dumps only through NetCDF, which it refuses without
the NetCDF library (SOLVER/parameters.F90:775), so there is no reference dump to
compare with.  The pieces are pinned separately against the reference's own code:

* the grid: the drop-in host runs the reference's build_kwf_grid
  (SOLVER/meshes_io.F90:489-778) and records its mapping_ijel_ikwf / kwf_mask
  in the state; the library's own builder (axh_kwf_grid) and the oracle's
  restatement must give the same numbering, with the default GLL ranges and
  with all 25 points of a radius-limited set of elements (shared points);
* the values: the reference's xdmf snapshots hold the same fields at the same
  GLL points (SOLVER/wavefields_io.f90:147-180: the solid displacement in
  (s, phi, z), the fluid displacement usz * inv_rho_fluid from
  axisym_gradient_fluid); the oracle's displ_only sample equals one of them
  (tests/golden/<case>/ref_xdmf.npz, written by the unmodified reference
  program) bit for bit at every kernel-wavefield point;
* the source mask (eradicate_src_elem_vec_values): zero on every entry of
  the ielsrc elements, unchanged elsewhere.

The GPU gather is compared with the oracle in tests/test_gpu_kwf.py.
"""
import math
import os

import numpy as np
import pytest

import cases
from axisem_amd import inputs
from axisem_amd.solver import Solver
from axisem_amd.state import read_dims, read_state
from oracle.oracle import Oracle


def _xdmf_index(st, gll_i, gll_j):
    """dump_xdmf_grid's numbering (SOLVER/meshes_io.F90:164-222), every element
    selected: fluid elements first, then solid; i over gll_i (outer), j over
    gll_j; a global point (fluid ids, then solid ids after nglob_fluid) is
    numbered where it first appears.  Returns {(domain, e, p): (index, owner)}
    for every visited entry, owner = the entry that numbered the point."""
    d = st.dims
    num = {}
    out = {}
    for dom, nel, ig, goff in ((1, d["nel_fluid"], st["igloc_fluid"], 0),
                               (0, d["nel_solid"], st["igloc_solid"], d["nglob_fluid"])):
        ig = np.asarray(ig).reshape(nel, 25)
        for e in range(nel):
            for i in gll_i:
                for j in gll_j:
                    p = i + 5 * j
                    g = int(ig[e, p]) - 1 + goff
                    if g not in num:
                        num[g] = (len(num), (dom, e, p))
                    out[(dom, e, p)] = num[g]
    return out


@pytest.mark.parametrize("name", sorted(cases.KWF_CASES))
def test_oracle_sample_equals_reference_xdmf_snapshot(name, state_cache):
    cfg = cases.KWF_CASES[name]
    _, stp = state_cache(name)
    st = read_state(stp)
    d = read_dims(stp)
    g = np.load(os.path.join(cases.case_dir(name), "ref_xdmf.npz"))
    snap_it = math.floor(float(g["snapshot_dt"]) / d["deltat"])
    niter = int(g["snap_index"]) * snap_it
    print(name, "oracle to step", niter, flush=True)
    o = Oracle(stp)
    o.run(1, niter, d["num_rec"], d["seis_it"])
    grid = o.kwf_grid()
    mapping, mask, nps, npf = grid
    assert npf > 0
    sample = o.kwf_sample(grid, src_mask=False)
    o.close()
    comps = ["s", "z"] if cases.NCOMP[cfg.source_type] == 2 else ["s", "p", "z"]
    gi = [int(x) for x in cfg.xdmf_gll_i.split()[1:]]
    gj = [int(x) for x in cfg.xdmf_gll_j.split()[1:]]
    xi = _xdmf_index(st, gi, gj)
    nel_s = d["nel_solid"]
    idx = np.full(nps + npf, -1)
    for x in np.nonzero(mask)[0]:
        e, p = divmod(int(x), 25)
        key = (0, e, p) if e < nel_s else (1, e - nel_s, p)
        plot, owner = xi[key]
        assert owner == key     # the snapshot's value comes from the same element
        idx[mapping[x] - 1] = plot
    assert (idx >= 0).all()
    for c, comp in enumerate(comps):
        ref = g["xdmf_" + comp][idx]
        got = sample[c]
        # the values are not trivially zero: the wavefield fills the mesh and
        # has reached the fluid (where phi is 0)
        assert np.count_nonzero(got[:nps]) > 0.5 * nps
        if comp != "p":
            assert np.count_nonzero(got[nps:]) > 0.5 * npf
        np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32),
                                      err_msg="%s component %s" % (name, comp))
    if len(comps) == 3:
        assert not np.any(sample[1][nps:])     # phi in the fluid is 0


# grids recorded by the reference's build_kwf_grid in the drop-in host
GRID_CASES = {
    "default": ("dipole_thetaforce_newmark_anel", ""),
    "all_points_radius_limited": ("mono_vertforce_newmark",
                                  "KERNEL_IBEG 0\nKERNEL_IEND 4\nKERNEL_JBEG 0\nKERNEL_JEND 4\n"
                                  "KERNEL_RMIN 2000\nKERNEL_RMAX 5000\n"),
}


@pytest.fixture(scope="module")
def reference_grids(tmp_path_factory, host_exe):
    out = {}
    base = tmp_path_factory.mktemp("kwf_grids")
    for key, (case, extra) in GRID_CASES.items():
        cfg = cases.ALL_CONFIGS[case]
        cfg = inputs.RunConfig(**{**cfg.__dict__, "extra_advanced": extra})
        rundir = str(base / key)
        inputs.write_run_dir(rundir, cases.mesh_of(case), cfg)
        out[key] = inputs.record_state(host_exe, rundir, os.path.join(rundir, "state.bin"),
                                       kwf_strain_it=7)
    return out


@pytest.mark.parametrize("key", sorted(GRID_CASES))
def test_grid_builders_equal_reference(key, reference_grids):
    stp = reference_grids[key]
    st = read_state(stp)
    rmap = np.asarray(st["kwf_mapping"])
    rmask = np.asarray(st["kwf_mask"])
    nps, npf = (int(x) for x in st["kwf_npoints"])
    assert rmask.sum() == nps + npf > 0
    ext = dict(default=(1, 3, 1, 3), all_points_radius_limited=(0, 4, 0, 4))[key]
    # the element selection of the reference (every selected element numbers
    # at least its interior points)
    emask = None
    if key != "default":
        emask = rmask.reshape(-1, 25).any(axis=1).astype(np.int32)
        assert 0 < emask.sum() < emask.size
    # the library's builder (host code: no GPU)
    s = Solver.from_state_file(stp)
    try:
        lnps, lnpf = s.kwf_grid(*ext, elem_mask=emask)
        lmap, lmask, _, _ = s.kwf_get_grid()
    finally:
        s.close()
    assert (lnps, lnpf) == (nps, npf)
    np.testing.assert_array_equal(lmask, rmask)
    np.testing.assert_array_equal(lmap[rmask != 0], rmap[rmask != 0])
    # the oracle's restatement
    o = Oracle(stp)
    omap, omask, onps, onpf = o.kwf_grid(*ext, elem_mask=emask)
    o.close()
    assert (onps, onpf) == (nps, npf)
    np.testing.assert_array_equal(omask, rmask)
    np.testing.assert_array_equal(omap[rmask != 0], rmap[rmask != 0])
    if key != "default":
        # shared points: the later copies carry the number of the first
        assert (omap != 0).sum() > rmask.sum()


def test_source_mask_zeroes_the_source_elements(reference_grids):
    """src_dump_type 'mask': the dipole force's source elements read zero,
    every other solid point is the unmasked sample."""
    stp = reference_grids["default"]
    st = read_state(stp)
    d = read_dims(stp)
    o = Oracle(stp)
    o.run(1, 40, d["num_rec"], d["seis_it"])
    grid = o.kwf_grid()
    a = o.kwf_sample(grid, src_mask=False)
    b = o.kwf_sample(grid, src_mask=True)
    o.close()
    mapping, mask, nps, npf = grid
    src = set(int(e) - 1 for e in np.asarray(st["ielsrc"])[:d["nelsrc"]])
    in_src = np.zeros(nps + npf, bool)
    for x in np.nonzero(mask)[0]:
        e = int(x) // 25
        if e in src:
            in_src[mapping[x] - 1] = True
    assert in_src.any()
    assert np.abs(a[:, in_src]).max() > 0           # the source elements moved
    assert not np.any(b[:, in_src])
    np.testing.assert_array_equal(a[:, ~in_src], b[:, ~in_src])


def test_kwf_entry_points_check_their_arguments(reference_grids):
    from axisem_amd.solver import AxhError
    s = Solver.from_state_file(reference_grids["default"])
    try:
        with pytest.raises(AxhError):
            s.kwf_grid(2, 1, 1, 3)        # empty range
        with pytest.raises(AxhError):
            s.kwf_grid(0, 5, 1, 3)        # beyond npol
        with pytest.raises(AxhError):
            s.set_kwf(0)
        with pytest.raises(AxhError):
            s.set_kwf(5, ring_snapshots=1)
        mapping, mask, nps, npf = s.kwf_get_grid()
        bad = mapping.copy()
        k = np.nonzero(mask)[0][:2]
        bad[k[1]] = bad[k[0]]             # one number twice
        L = s._L
        import ctypes as C
        rc = L.axh_upload_kwf_grid(s._h, nps, npf, bad.ctypes.data_as(C.c_void_p),
                                   mask.ctypes.data_as(C.c_void_p))
        assert rc != 0
        s.set_kwf(5)
        assert L.axh_kwf_snapshots(s._h, 1, 20) == 5     # iter 0, 5, 10, 15, 20
        assert L.axh_kwf_snapshots(s._h, 6, 20) == 3     # 10, 15, 20
        assert L.axh_kwf_snapshots(s._h, 6, 9) == 0
    finally:
        s.close()
