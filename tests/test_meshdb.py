"""The library's reader of the reference MESHER's mesh database (no GPU).

Every array the reader returns must equal what the reference SOLVER itself
reads from the same file (SOLVER/get_mesh.f90:47-409), as the drop-in host
hands it to the library (the recorded uploads of the golden cases): the
global numbering, the axis flags, the solid-fluid boundary elements and the
spectral operators.  A context given its mesh-only uploads by the reader
belongs to the same batches as the recorded one.  The benchmark meshes are
checked against the counts the MESHER writes to mesh_params.h.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cases
from axisem_amd import solver as S
from axisem_amd.state import read_state

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = ["mono_vertforce_newmark", "reg_dipole_thetaforce_symplec4_anel",
          "p50_mono_vertforce_newmark", "ani_dipole_thetaforce_newmark_anel"]


def _key(h):
    k = (C.c_uint64 * 2)()
    S._check(S.library().axh_batch_key(h, k))
    return tuple(k)


@pytest.mark.parametrize("name", GOLDEN)
def test_reader_equals_the_solvers_own_reading(name, state_cache):
    _, path = state_cache(name)
    st = read_state(path)
    d = st.dims
    db = S.MeshDB(os.path.join(cases.mesh_of(name), "meshdb.dat0000"))
    I = db.info
    for k in ("nel_solid", "nel_fluid", "nglob_solid", "nglob_fluid", "nel_bdry"):
        assert I[k] == d[k], k
    assert I["npol"] == 4 and I["nproc"] == 1 and I["realkind_bytes"] == 4
    for k in ("igloc_solid", "igloc_fluid", "axis_solid", "axis_fluid"):
        np.testing.assert_array_equal(db.array(k), st[k], err_msg=k)
    if d["nel_bdry"]:
        for k in ("bdry_solid_el", "bdry_fluid_el", "bdry_jpol_solid", "bdry_jpol_fluid"):
            np.testing.assert_array_equal(db.array(k), st[k], err_msg=k)
    for k in ("G0", "G1", "G1T", "G2", "G2T"):
        np.testing.assert_array_equal(db.array(k), st[k], err_msg=k)
    assert I["bkgrdmodel"] == ("prem_ani" if name.startswith("ani") else "prem_iso")
    # a context with the reader's mesh uploads batches with the recorded one
    rec = S.Solver.from_state_file(path)
    dims = S.Dims()
    for k in rec.dims:
        if hasattr(dims, k) and k not in ("deltat", "magnitude"):
            setattr(dims, k, rec.dims[k])
    dims.deltat, dims.magnitude = rec.dims["deltat"], rec.dims["magnitude"]
    h = C.c_void_p()
    S._check(S.library().axh_create(C.byref(dims), C.byref(h)))
    fresh = S.Solver(h, dict(rec.dims))
    db.upload(fresh)
    assert _key(fresh._h) == _key(rec._h)
    fresh.close()
    rec.close()
    db.close()


@pytest.mark.parametrize("mesh", ["prem_iso_50s", "prem_iso_20s", "prem_iso_10s"])
def test_benchmark_meshes_match_mesh_params(mesh):
    mdir = os.path.join(REPO, "data", "meshes", mesh)
    if not os.path.exists(os.path.join(mdir, "meshdb.dat0000")):
        pytest.skip("benchmark mesh not built here")
    params = {}
    for line in open(os.path.join(mdir, "mesh_params.h")):
        m = re.match(r"\s*integer, parameter ::\s*(\w+)\s*=\s*(\d+)", line)
        if m:
            params[m.group(1)] = int(m.group(2))
    db = S.MeshDB(os.path.join(mdir, "meshdb.dat0000"))
    I = db.info
    assert I["nelem"] == params["nelem"]
    assert I["nel_solid"] == params["nel_solid"] and I["nel_fluid"] == params["nel_fluid"]
    ig = db.array("igloc_solid")
    assert ig.min() == 1 and ig.max() == I["nglob_solid"]
    assert len(np.unique(ig)) == I["nglob_solid"]
    assert db.array("axis_solid").sum() == I["naxel_solid"]
    db.close()


def test_reader_rejects_other_files(tmp_path):
    p = tmp_path / "junk"
    p.write_bytes(b"\x07\x00\x00\x00abcdefg\x07\x00\x00\x00")
    h = C.c_void_p()
    assert S.library().axh_meshdb_open(str(p).encode(), C.byref(h)) == -4
    assert S.library().axh_meshdb_open(str(tmp_path / "none").encode(), C.byref(h)) == -4
