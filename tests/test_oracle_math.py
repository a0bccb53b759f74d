"""Oracle checks fixed by mathematics rather than by the reference's outputs.

* 5x5 products (SOLVER/unrolled_loops.f90) against a textbook matmul.
* The assembled stiffness operator K (stiffness + direct stiffness summation)
  is symmetric on assembled vectors: <v, K u> = <u, K v>.
* Null spaces: a rigid vertical translation of the solid (monopole) and a
  constant fluid potential produce no force.
* Time schemes: consistency of the symplectic coefficients and 4th/6th order
  convergence on the harmonic oscillator; ML_SO4m5 and ML_SO6m7 cannot be run
  by the reference program itself (it lower-cases the scheme name before its
  select case), so these order checks are what pins them.
* Source injection in the fluid (add_source_fl, which the reference program
  cannot place on its meshes): with every operator zeroed, one Newmark step
  from rest is a point-mass update, ddchi = inv_mass_fluid * source_term_fl *
  stf(iter) at the source elements' points and zero everywhere else.
"""
import numpy as np
import pytest

from axisem_amd.state import read_state, write_state
from fluid_source import fluid_source_state, pick_fluid_elements, zero_operators
from oracle.oracle import Oracle, mxm, symplectic_coefficients


def test_mxm_matches_matmul():
    rng = np.random.default_rng(1)
    for _ in range(200):
        a = rng.standard_normal((5, 5)).astype(np.float32)
        b = rng.standard_normal((5, 5)).astype(np.float32)
        c = mxm(a, b)
        ref = a.astype(np.float64) @ b.astype(np.float64)
        np.testing.assert_allclose(c, ref, rtol=0, atol=2e-6 * np.abs(ref).max() + 1e-7)
    # exact on small integers
    a = rng.integers(-8, 8, (5, 5)).astype(np.float32)
    b = rng.integers(-8, 8, (5, 5)).astype(np.float32)
    np.testing.assert_array_equal(mxm(a, b), a @ b)


def _assembled_random(st, rng, comps):
    ig = st["igloc_solid"] - 1
    g = rng.standard_normal((3, st.dims["nglob_solid"]))
    u = np.concatenate([g[c][ig] if c in comps else np.zeros(ig.size) for c in range(3)])
    return u.astype(np.float32)


@pytest.mark.parametrize("name", ["mono_vertforce_newmark", "dipole_mtr_symplec4",
                                  "quad_mtp_newmark"])
def test_assembled_stiffness_is_symmetric(name, state_cache):
    _, path = state_cache(name)
    st = read_state(path)
    o = Oracle(path)
    rng = np.random.default_rng(7)
    comps = (0, 2) if st.dims["src_type"] == 0 else (0, 1, 2)
    u = _assembled_random(st, rng, comps)
    v = _assembled_random(st, rng, comps)
    # mask the axis like the time loop does (non-zero components vanish there)
    ku = o.assemble_solid(o.solid_stiffness(u))
    kv = o.assemble_solid(o.solid_stiffness(v))
    # inner product over global points: weight each element-local copy by
    # 1/multiplicity so every global point counts once
    ig = st["igloc_solid"] - 1
    mult = np.bincount(ig, minlength=st.dims["nglob_solid"])[ig].astype(np.float64)
    w = np.tile(1.0 / mult, 3)
    sel = np.zeros(3 * ig.size, bool)
    for c in comps:
        sel[c * ig.size:(c + 1) * ig.size] = True
    a = np.sum((w * v.astype(np.float64) * ku)[sel])
    b = np.sum((w * u.astype(np.float64) * kv)[sel])
    scale = np.sqrt(np.sum((w * ku.astype(np.float64) ** 2)[sel]) *
                    np.sum((w * v.astype(np.float64) ** 2)[sel]))
    assert abs(a - b) <= 1e-4 * scale, (a, b, scale)


def test_monopole_rigid_translation_has_no_force(state_cache):
    _, path = state_cache("mono_vertforce_newmark")
    st = read_state(path)
    o = Oracle(path)
    ns = 25 * st.dims["nel_solid"]
    u = np.zeros(3 * ns, np.float32)
    u[2 * ns:] = 1.0                       # u_z = const
    ku = o.assemble_solid(o.solid_stiffness(u))
    ref = o.assemble_solid(o.solid_stiffness(_assembled_random(st, np.random.default_rng(3), (0, 2))))
    assert np.abs(ku).max() < 1e-4 * np.abs(ref).max()


def test_fluid_constant_potential_has_no_force(state_cache):
    _, path = state_cache("mono_vertforce_newmark")
    st = read_state(path)
    o = Oracle(path)
    chi = np.ones(25 * st.dims["nel_fluid"], np.float32)
    k = o.assemble_fluid(o.fluid_stiffness(chi))
    rng = np.random.default_rng(5)
    chi_r = rng.standard_normal(st.dims["nglob_fluid"])[st["igloc_fluid"] - 1].astype(np.float32)
    kr = o.assemble_fluid(o.fluid_stiffness(chi_r))
    assert np.abs(k).max() < 1e-4 * np.abs(kr).max()


def _one_step_error(scheme, dt):
    """Local error of one step on x'' = -x from (1, 0), with the drift/kick
    sequence of the symplectic loop (SOLVER/time_evol_wave.F90:595-725)."""
    cd, cv, _ = symplectic_coefficients(scheme, dt)
    x, v = 1.0, 0.0
    for i in range(len(cv)):
        x = x + v * cd[i]
        v = v + (-x) * cv[i]
    x = x + v * cd[len(cv)]
    return abs(x - np.cos(dt)) + abs(v + np.sin(dt))


@pytest.mark.parametrize("scheme,order", [(1, 4), (2, 4), (3, 6)])
def test_symplectic_scheme_order(scheme, order):
    cd, cv, cf = symplectic_coefficients(scheme, 1.0)
    # single-precision literals in the reference limit the sums to ~1e-7
    assert abs(cd.sum() - 1.0) < 1e-6 and abs(cv.sum() - 1.0) < 1e-6
    np.testing.assert_allclose(cf, np.cumsum(cd)[:len(cv)])
    # symmetric compositions: coefd and coefv are palindromes
    np.testing.assert_array_equal(cd, cd[::-1])
    np.testing.assert_array_equal(cv, cv[::-1])
    # a p-th order scheme has local error O(dt^(p+1)); measured between
    # dt = 1 and 1/2 (above the floor of the single-precision constants)
    rate = np.log2(_one_step_error(scheme, 1.0) / _one_step_error(scheme, 0.5))
    assert order + 0.5 < rate < order + 1.5, rate


def test_library_and_oracle_coefficients_agree():
    from axisem_amd.solver import symplectic_coefficients as lib_coeffs, library
    try:
        library()
    except (ImportError, OSError):
        pytest.skip("libaxihip.so not built")
    for scheme in (1, 2, 3):
        for dt in (0.1432156680021883, 0.09380123750737684):
            a = lib_coeffs(scheme, dt)
            b = symplectic_coefficients(scheme, dt)
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x, y)


def test_add_source_fl_formula(state_cache, tmp_path):
    """add_source_fl (SOLVER/time_evol_wave.F90:1062-1076) with the stiffness
    and coupling operators zeroed: from rest, step n of newmark2
    (SOLVER/time_evol_wave.F90:357-492) leaves chi = 0 and
    ddchi = -inv_mass_fluid * (0 - source_term_fl(:,:,k) * stf(n)) on the
    points of fluid element ielsrc(k); stf(n) = 0 injects nothing.  The source
    term is set on the interior points only (one element copy each, so the
    assembly adds nothing to it) and is asymmetric in (i, j)."""
    _, path = state_cache("mono_vertforce_newmark")
    st = read_state(path)
    d = st.dims
    elems = pick_fluid_elements(st, 3)
    stf = np.zeros(d["niter"], np.float32)
    stf[1], stf[2] = 0.75, -1.25          # nothing at step 1
    fs = zero_operators(fluid_source_state(st, elems, seed=3, interior_only=True, stf=stf))
    p = str(tmp_path / "fluid_src.bin")
    write_state(fs, p)
    o = Oracle(p)
    nf = 25 * d["nel_fluid"]
    minv = np.asarray(st["inv_mass_fluid"], np.float32)
    term = fs["source_term"][: 25 * len(elems)].reshape(len(elems), 25)
    hdt = 0.5 * d["deltat"]

    def expected(s):
        dd = np.zeros(nf, np.float32)
        for k, e in enumerate(elems):
            x = term[k] * np.float32(s)               # source_term_fl * stf, FP32
            dd[25 * e: 25 * e + 25] = minv[25 * e: 25 * e + 25] * x
        return dd

    o.run(1, 1, d["num_rec"], d["seis_it"])
    disp, velo, chi, dchi, ddchi = o.fields()
    for a in (disp, velo, chi, dchi, ddchi):
        assert not a.any()
    o.run(2, 2, d["num_rec"], d["seis_it"])
    disp, velo, chi, dchi, ddchi = o.fields()
    dd2 = expected(stf[1])
    assert np.count_nonzero(dd2) == 9 * len(elems)
    np.testing.assert_array_equal(ddchi, dd2)
    np.testing.assert_array_equal(dchi, (hdt * dd2.astype(np.float64)).astype(np.float32))
    assert not chi.any() and not disp.any() and not velo.any()
    # the next step: chi from the predictor, the new source on top of it
    o.run(3, 3, d["num_rec"], d["seis_it"])
    _, _, chi3, dchi3, ddchi3 = o.fields()
    dd3 = expected(stf[2])
    np.testing.assert_array_equal(ddchi3, dd3)
    dt, hdt2 = d["deltat"], 0.5 * (d["deltat"] * d["deltat"])
    chi_pred = (0.0 + dt * dchi.astype(np.float64) + hdt2 * dd2.astype(np.float64))
    np.testing.assert_array_equal(chi3, chi_pred.astype(np.float32))
    np.testing.assert_array_equal(
        dchi3, (dchi.astype(np.float64) + hdt * (dd2 + dd3).astype(np.float64)).astype(np.float32))
