"""The C-ABI library without a GPU: symbols, argument checking, state files.

Nothing here launches device work (the library touches the GPU only from the
first axh_run / axh_download_fields / axh_device_bytes call).
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from axisem_amd import solver as S
from axisem_amd.state import read_state, write_state

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(REPO, "include", "axihip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(axh_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = S.library()
    names = _declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_bad_dims_are_rejected():
    lib = S.library()
    d = S.Dims()
    d.abi_version, d.npol, d.nel_solid, d.nglob_solid = 1, 5, 10, 100
    d.seis_it, d.deltat = 1, 0.1
    h = C.c_void_p()
    rc = lib.axh_create(C.byref(d), C.byref(h))
    assert rc == -5 and b"npol" in lib.axh_last_error()
    d.npol = 4
    d.abi_version = 99
    assert lib.axh_create(C.byref(d), C.byref(h)) == -1
    d.abi_version = 1
    d.deltat = -1.0
    assert lib.axh_create(C.byref(d), C.byref(h)) == -1
    d.deltat, d.nelsrc = 0.1, 9
    assert lib.axh_create(C.byref(d), C.byref(h)) == -1
    assert lib.axh_run(None, 1, 1, None) == -1


def test_state_file_roundtrip(state_cache, tmp_path):
    """Uploads through the C-ABI, written by axh_save_state, equal the file
    the Fortran host wrote; the python reader/writer round-trips it too."""
    _, path = state_cache("dipole_thetaforce_newmark_anel")
    st = read_state(path)
    s = S.Solver.from_state(st)
    out = str(tmp_path / "again.bin")
    S._check(S.library().axh_save_state(s._h, out.encode()))
    st2 = read_state(out)
    assert st.dims == st2.dims
    assert set(st.arrays) == set(st2.arrays)
    for k in st.arrays:
        np.testing.assert_array_equal(st[k], st2[k], err_msg=k)
        assert st.shape(k) == st2.shape(k)
    out2 = str(tmp_path / "py.bin")
    write_state(st, out2)
    st3 = read_state(out2)
    for k in st.arrays:
        np.testing.assert_array_equal(st[k], st3[k])
    s.close()


def test_upload_size_checks(state_cache):
    _, path = state_cache("mono_vertforce_newmark")
    st = read_state(path)
    lib = S.library()
    d = S.Dims()
    for k in ("abi_version", "npol", "nel_solid", "nel_fluid", "nglob_solid", "nglob_fluid",
              "nel_bdry", "src_type", "scheme", "niter", "seis_it", "num_rec", "nelsrc"):
        setattr(d, k, st.dims[k])
    d.deltat, d.magnitude = st.dims["deltat"], st.dims["magnitude"]
    h = C.c_void_p()
    assert lib.axh_create(C.byref(d), C.byref(h)) == 0
    stf = np.zeros(5, np.float32)
    src = np.zeros(8 * 75, np.float32)
    iel = np.ones(8, np.int32)
    rc = lib.axh_upload_source(h, iel.ctypes.data_as(C.c_void_p), src.ctypes.data_as(C.c_void_p),
                               stf.ctypes.data_as(C.c_void_p), 5)
    assert rc == -1 and b"stf table" in lib.axh_last_error()
    assert lib.axh_upload_solid_term(h, 33, src.ctypes.data_as(C.c_void_p)) == -1
    assert lib.axh_destroy(h) == 0


def test_attenuation_uploads_follow_the_memory_variable_mode(state_cache):
    """axh_upload_attenuation_full serves COARSE_GRAINED false only, and the
    coarse-grained upload the default mode: each refuses the other's context,
    an elastic one and a NULL argument, naming the reason."""
    lib = S.library()
    st = read_state(state_cache("mono_explosion_newmark_anelfull")[1])
    assert st.dims["anel"] == 1 and st.dims["coarse_grained"] == 0
    assert "Y" in st.arrays and "Y_cg4" not in st.arrays

    def ctx(anel, cg):
        d = S.Dims()
        for k in S.Dims._fields_:
            if k[0] in st.dims:
                setattr(d, k[0], st.dims[k[0]])
        d.anel, d.coarse_grained = anel, cg
        h = C.c_void_p()
        assert lib.axh_create(C.byref(d), C.byref(h)) == 0, lib.axh_last_error()
        return h

    full, cg4 = S.AttFull(), S.Att()
    h = ctx(1, 0)
    assert lib.axh_upload_attenuation_full(h, None) == -1
    assert lib.axh_upload_attenuation_cg4(h, C.byref(cg4)) == -2
    assert b"all points" in lib.axh_last_error()
    assert lib.axh_destroy(h) == 0
    h = ctx(1, 1)
    assert lib.axh_upload_attenuation_full(h, C.byref(full)) == -2
    assert b"coarse-grained" in lib.axh_last_error()
    assert lib.axh_destroy(h) == 0
    h = ctx(0, 1)
    assert lib.axh_upload_attenuation_full(h, C.byref(full)) == -2
    assert b"elastic" in lib.axh_last_error()
    assert lib.axh_destroy(h) == 0


def test_symbols_have_reference_citations():
    """Every declaration in include/axihip.h documents what it replaces."""
    text = open(os.path.join(REPO, "include", "axihip.h")).read()
    assert text.count("SOLVER/") >= 15


def test_inputs_pass_layout_checks_and_numbering_is_igloc(state_cache):
    """The library's gather maps are a relabelling of the reference's igloc:
    two element-local points share a library index iff they share igloc."""
    import cases
    for name in sorted(cases.CASES):
        _, path = state_cache(name)
        st = read_state(path)
        s = S.Solver.from_state(st)
        S._check(S.library().axh_check_inputs(s._h))
        for dom, key, nk in ((0, "igloc_solid", "nglob_solid"), (1, "igloc_fluid", "nglob_fluid")):
            ig = st[key]
            gid = np.zeros(ig.size, np.int32)
            n = C.c_int32()
            S._check(S.library().axh_global_numbering(s._h, dom, gid.ctypes.data_as(C.c_void_p),
                                                      C.byref(n)))
            assert n.value == len(np.unique(ig)) == st.dims[nk]
            # bijection between the two labellings
            pairs = np.unique(np.stack([ig, gid]), axis=1)
            assert pairs.shape[1] == n.value
            full_memvars = st.dims["anel"] and not st.dims["coarse_grained"]
            if dom == 1 or "solid_absorbing_gamma" in st.arrays or full_memvars:
                # fluid, and the solid of a mesh with absorbing boundaries or
                # memory variables at all points (every point assembled): by
                # first appearance in element order
                first = np.unique(gid, return_index=True)[1]
                assert np.all(np.diff(first) > 0)
            else:
                # solid: the points with several copies come first; only
                # single-copy points on the axis or the fluid boundary may
                # share their range
                ncopy = np.bincount(gid)
                multi = ncopy[gid] > 1
                el = np.arange(ig.size) // 25
                special = np.isin(el, np.nonzero(st["axis_solid"])[0]) | \
                    np.isin(el, st["bdry_solid_el"] - 1)
                nsingle_special = len(np.unique(gid[~multi & special]))
                assert gid[multi].max() < len(np.unique(gid[multi])) + nsingle_special
        s.close()


def _ctx_array(solvers):
    return (C.c_void_p * len(solvers))(*[s._h for s in solvers])


def test_batch_members_share_the_mesh(state_cache):
    """Sources of one type at different depths on one mesh form a batch: every
    upload except the per-source ones is byte-equal (checked by
    axh_batch_create, which touches no GPU)."""
    import cases
    lib = S.library()
    for batch, names in sorted(cases.BATCHES.items()):
        solvers = [S.Solver.from_state_file(state_cache(n)[1]) for n in names]
        h = C.c_void_p()
        S._check(lib.axh_batch_create(_ctx_array(solvers), len(solvers), C.byref(h)))
        # members are held: no destroy, no uploads
        assert lib.axh_destroy(solvers[0]._h) == -2
        assert b"batch" in lib.axh_last_error()
        rec = np.ones(3 * solvers[1].dims["num_rec"], np.int32)
        assert lib.axh_upload_receivers(solvers[1]._h, rec.ctypes.data_as(C.c_void_p)) == -2
        S._check(lib.axh_batch_destroy(h))
        for s in solvers:
            s.close()


def test_batch_rejects_other_meshes_and_bad_sizes(state_cache):
    import cases
    lib = S.library()
    mono = S.Solver.from_state_file(state_cache(cases.BATCHES["mono_newmark_anel"][0])[1])
    di = [S.Solver.from_state_file(state_cache(n)[1])
          for n in cases.BATCHES["dipole_symplec4_anel"][:2]]
    h = C.c_void_p()
    assert lib.axh_batch_create(_ctx_array([di[0], mono]), 2, C.byref(h)) == -1
    assert b"differs" in lib.axh_last_error()
    assert lib.axh_batch_create(_ctx_array([di[0], di[0]]), 2, C.byref(h)) == -1
    assert lib.axh_batch_create(_ctx_array(di), 0, C.byref(h)) == -1
    nine = [di[0]] * 9
    assert lib.axh_batch_create(_ctx_array(nine), 9, C.byref(h)) == -1
    assert lib.axh_batch_run(None, 1, 1, None) == -1
    # a different elastic/anelastic setting is a different mesh part
    el = S.Solver.from_state_file(state_cache("dipole_mtr_symplec4")[1])
    assert lib.axh_batch_create(_ctx_array([di[0], el]), 2, C.byref(h)) == -1
    for s in (mono, el, *di):
        s.close()
