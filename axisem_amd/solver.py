"""Python host of the GPU time loop, over the C-ABI of include/axihip.h.

Mirrors the reference's two-phase interface of module time_evol_wave
(SOLVER/time_evol_wave.F90:26-41): a solver is *prepared* from the arrays the
reference's setup produces (here: an axh state file written by the drop-in
Fortran host, or the uploads done one by one) and the *time loop* is then run,
in one call or in chunks, returning the receiver seismograms the reference
writes (SOLVER/seismograms.f90:742-780).

The library must be built (``axisem_amd/libaxihip.so``); there is no CPU
fallback: without the library or without a GPU every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import state as _state

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libaxihip.so")
# the build whose kernels contract a*b+c into fused multiply-adds: within the
# reference's seismogram misfit, not bit for bit
LIB_FMA_PATH = os.path.join(PKG, "libaxihip_fma.so")

AXH_OK, AXH_ERR_ARG, AXH_ERR_STATE, AXH_ERR_BLOWUP = 0, -1, -2, -6
MAX_BATCH = 16

SOLID_TERMS = ["M11s", "M21s", "M41s", "M12s", "M22s", "M42s", "M11z", "M21z",
               "M41z", "M32s", "M13s", "M33s", "M43s", "M1phi", "M2phi", "M4phi",
               "M_1", "M_2", "M_3", "M_4", "M_5", "M_6", "M_7", "M_8",
               "M_w1", "M_w2", "M_w3", "M_w4", "M_w5"]
FLUID_TERMS = ["M1chi_fl", "M2chi_fl", "M4chi_fl", "M_w_fl"]
# arguments of axh_upload_fluid_gradient, in order
FLUID_GRADIENT = ["inv_rho_fluid", "DsDeta_over_J_flu", "DzDeta_over_J_flu",
                  "DsDxi_over_J_flu", "DzDxi_over_J_flu"]


class AxhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("axh error %d: %s" % (code, msg))
        self.code = code


class BlowUp(AxhError):
    pass


class Dims(C.Structure):
    _fields_ = [(n, C.c_int32) for n in _state.DIMS_FIELDS] + \
               [("reserved", C.c_int32 * 5), ("deltat", C.c_double), ("magnitude", C.c_double)]


class Att(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "Y", "V_s_eta", "V_s_xi", "V_z_eta", "V_z_xi", "DsDeta_over_J", "DzDeta_over_J",
        "DsDxi_over_J", "DzDxi_over_J", "inv_s", "delta_mu", "delta_kappa",
        "exp_w_j_deltat", "ts_fac_t", "ts_fac_tm1", "a_j_mu", "a_j_kappa")]


class AttFull(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "Y", "V_s_eta", "V_s_xi", "V_z_eta", "V_z_xi", "Y0", "V0_s_eta", "V0_s_xi",
        "V0_z_eta", "V0_z_xi", "DsDeta_over_J", "DzDeta_over_J", "DsDxi_over_J",
        "DzDxi_over_J", "inv_s", "delta_mu", "delta_kappa",
        "exp_w_j_deltat", "ts_fac_t", "ts_fac_tm1", "a_j_mu", "a_j_kappa")]


class MeshDBInfo(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "nproc", "npol", "nelem", "npoint", "nel_solid", "nel_fluid", "nglob_solid",
        "nglob_fluid", "nel_bdry", "ndisc", "npoin", "nglob", "naxel", "naxel_solid",
        "naxel_fluid", "have_bdry", "have_fluid", "realkind_bytes")] + \
        [(n, C.c_double) for n in ("pts_wavelngth", "period", "courant", "deltat", "router")] + \
        [("bkgrdmodel", C.c_char * 64)]


_libs = {}


def library(fused: bool = False):
    """Load libaxihip.so, or with fused=True libaxihip_fma.so (raises if it
    has not been built).  Contexts of one build are used with that build."""
    L = _libs.get(fused)
    if L is None:
        path = LIB_FMA_PATH if fused else LIB_PATH
        if not os.path.exists(path):
            raise ImportError("%s is not built; run __graft_entry__.build()"
                              % os.path.basename(path))
        L = C.CDLL(path)
        P, I = C.c_void_p, C.c_int
        L.axh_last_error.restype = C.c_char_p
        L.axh_create.argtypes = [C.POINTER(Dims), C.POINTER(P)]
        L.axh_destroy.argtypes = [P]
        L.axh_load_state.argtypes = [C.c_char_p, C.POINTER(P)]
        L.axh_save_state.argtypes = [P, C.c_char_p]
        L.axh_upload_spectral.argtypes = [P] * 6
        L.axh_upload_solid_term.argtypes = [P, I, P]
        L.axh_upload_fluid_term.argtypes = [P, I, P]
        L.axh_upload_mass.argtypes = [P, P, P]
        L.axh_upload_mesh_maps.argtypes = [P, P, P, P, P]
        L.axh_upload_boundary.argtypes = [P, P, P, P, P, P]
        L.axh_upload_fluid_surface_mask.argtypes = [P, P]
        L.axh_upload_absorbing.argtypes = [P, P, P]
        L.axh_upload_source.argtypes = [P, P, P, P, C.c_int32]
        L.axh_upload_attenuation_cg4.argtypes = [P, C.POINTER(Att)]
        L.axh_upload_attenuation_full.argtypes = [P, C.POINTER(AttFull)]
        L.axh_upload_receivers.argtypes = [P, P]
        L.axh_run.argtypes = [P, C.c_int32, C.c_int32, P]
        L.axh_run_begin.argtypes = [P, C.c_int32, C.c_int32, P]
        L.axh_run_end.argtypes = [P]
        L.axh_seis_samples.argtypes = [P, C.c_int32, C.c_int32]
        L.axh_download_fields.argtypes = [P, P, P, P, P, P]
        L.axh_upload_fields.argtypes = [P, P, P, P, P, P]
        L.axh_last_run_ms.argtypes = [P, C.POINTER(C.c_double)]
        L.axh_device_bytes.argtypes = [P, C.POINTER(C.c_size_t)]
        L.axh_kernel_stats.argtypes = [P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.axh_kernel_name.restype = C.c_char_p
        L.axh_kernel_name.argtypes = [C.c_int]
        L.axh_set_profiling.argtypes = [P, C.c_int]
        L.axh_set_graph_steps.argtypes = [P, C.c_int32]
        L.axh_batch_set_graph_steps.argtypes = [P, C.c_int32]
        L.axh_run_noalloc.argtypes = [P, C.c_int32, C.c_int32, P]
        L.axh_convolve_traces.argtypes = [C.c_int32, C.c_int32, P, C.c_int32, P, C.c_double,
                                          C.c_double, P]
        L.axh_check_inputs.argtypes = [P]
        L.axh_global_numbering.argtypes = [P, C.c_int32, P, C.POINTER(C.c_int32)]
        L.axh_symplectic_coefficients.argtypes = [C.c_int32, C.c_double, P, P, P,
                                                  C.POINTER(C.c_int32)]
        L.axh_batch_create.argtypes = [C.POINTER(P), C.c_int32, C.POINTER(P)]
        L.axh_batch_destroy.argtypes = [P]
        L.axh_batch_run.argtypes = [P, C.c_int32, C.c_int32, C.POINTER(P)]
        L.axh_batch_last_run_ms.argtypes = [P, C.POINTER(C.c_double)]
        L.axh_batch_set_profiling.argtypes = [P, C.c_int]
        L.axh_batch_set_group_width.argtypes = [P, C.c_int32]
        L.axh_batch_step.argtypes = [P, C.c_int32, C.POINTER(P)]
        L.axh_mesh_fingerprint.argtypes = [P, P]
        L.axh_footprint.argtypes = [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.axh_mesh_get.argtypes = [P, C.POINTER(P)]
        L.axh_mesh_bind.argtypes = [P, P]
        L.axh_mesh_bytes.argtypes = [P, C.POINTER(C.c_size_t)]
        L.axh_mesh_release.argtypes = [P]
        L.axh_batch_kernel_stats.argtypes = [P, C.c_int, C.POINTER(C.c_double),
                                             C.POINTER(C.c_int64)]
        L.axh_batch_key.argtypes = [P, P]
        L.axh_meshdb_open.argtypes = [C.c_char_p, C.POINTER(P)]
        L.axh_meshdb_close.argtypes = [P]
        L.axh_meshdb_get_info.argtypes = [P, C.POINTER(MeshDBInfo)]
        L.axh_meshdb_array.argtypes = [P, C.c_char_p, P, C.c_size_t]
        L.axh_meshdb_upload.argtypes = [P, P]
        L.axh_upload_fluid_gradient.argtypes = [P] * 6
        I32, PI32 = C.c_int32, C.POINTER(C.c_int32)
        L.axh_kwf_grid.argtypes = [P, I32, I32, I32, I32, P, PI32, PI32]
        L.axh_upload_kwf_grid.argtypes = [P, I32, I32, P, P]
        L.axh_kwf_get_grid.argtypes = [P, P, P, PI32, PI32]
        L.axh_set_kwf.argtypes = [P, I32, I32, I32]
        L.axh_kwf_snapshots.argtypes = [P, I32, I32]
        L.axh_run_kwf.argtypes = [P, I32, I32, P, P]
        L.axh_batch_step_kwf.argtypes = [P, I32, C.POINTER(P), C.POINTER(P)]
        L.axh_run_kwf_begin.argtypes = [P, I32, I32, P, P]
        _libs[fused] = L
    return L


def _check(rc, L=None):
    if rc != AXH_OK:
        msg = (L or library()).axh_last_error().decode(errors="replace")
        raise (BlowUp if rc == AXH_ERR_BLOWUP else AxhError)(rc, msg)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def symplectic_coefficients(scheme: int, dt: float):
    cd, cv, cf = (np.zeros(16) for _ in range(3))
    n = C.c_int32()
    _check(library().axh_symplectic_coefficients(scheme, dt, _p(cd), _p(cv), _p(cf),
                                                 C.byref(n)))
    return cd[: n.value + 1], cv[: n.value], cf[: n.value]


class Solver:
    """One wavefield on one GPU: prepared once, then stepped by ``run``."""

    def __init__(self, handle, dims: dict, fused: bool = False):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        self.dims = dims
        self.iter = 0
        self.fused = fused
        self._L = library(fused)

    def _ok(self, rc):
        _check(rc, self._L)

    # --- construction -----------------------------------------------------
    @classmethod
    def from_state_file(cls, path: str, fused: bool = False) -> "Solver":
        L = library(fused)
        h = C.c_void_p()
        _check(L.axh_load_state(path.encode(), C.byref(h)), L)
        return cls(h, _state.read_dims(path), fused)

    @classmethod
    def from_state(cls, st: _state.State) -> "Solver":
        """Upload a State through the individual C-ABI upload calls, exactly as
        the Fortran host does."""
        L = library()
        d = st.dims
        dims = Dims()
        for k in _state.DIMS_FIELDS:
            setattr(dims, k, d[k])
        dims.deltat, dims.magnitude = d["deltat"], d["magnitude"]
        h = C.c_void_p()
        _check(L.axh_create(C.byref(dims), C.byref(h)))
        self = cls(h, dict(d))
        f32 = lambda n: np.ascontiguousarray(st[n], np.float32)
        i32 = lambda n: np.ascontiguousarray(st[n], np.int32)
        keep = []

        def K(a):
            keep.append(a)
            return _p(a)
        _check(L.axh_upload_spectral(h, *[K(f32(n)) for n in ("G0", "G1", "G1T", "G2", "G2T")]))
        for t, n in enumerate(SOLID_TERMS):
            if n in st.arrays:
                _check(L.axh_upload_solid_term(h, t, K(f32(n))))
        for w in range(1, 11):
            n = "M0_w%d" % w
            if n in st.arrays:
                _check(L.axh_upload_solid_term(h, 64 + w - 1, K(f32(n))))
        for t, n in enumerate(FLUID_TERMS):
            if n in st.arrays:
                _check(L.axh_upload_fluid_term(h, t, K(f32(n))))
        if "M0_w_fl" in st.arrays:
            _check(L.axh_upload_fluid_term(h, 64, K(f32("M0_w_fl"))))
        _check(L.axh_upload_mass(h, K(f32("inv_mass_rho")), K(f32("inv_mass_fluid"))))
        _check(L.axh_upload_mesh_maps(h, K(i32("igloc_solid")), K(i32("igloc_fluid")),
                                      K(i32("axis_solid")), K(i32("axis_fluid"))))
        if d["nel_bdry"] > 0:
            _check(L.axh_upload_boundary(h, *[K(i32(n)) for n in (
                "bdry_solid_el", "bdry_fluid_el", "bdry_jpol_solid", "bdry_jpol_fluid")],
                K(f32("bdry_matr"))))
        if "fluid_free_surface_mask" in st.arrays:
            _check(L.axh_upload_fluid_surface_mask(h, K(f32("fluid_free_surface_mask"))))
        if "solid_absorbing_gamma" in st.arrays:
            gf = (K(f32("fluid_absorbing_gamma")) if "fluid_absorbing_gamma" in st.arrays
                  else None)
            _check(L.axh_upload_absorbing(h, K(f32("solid_absorbing_gamma")), gf))
        stf = f32("stf")
        _check(L.axh_upload_source(h, K(i32("ielsrc")), K(f32("source_term")), K(stf),
                                   int(stf.size)))
        if d["anel"] and not d["coarse_grained"]:
            a = AttFull()
            for fld, n in (("Y", "Y"), ("V_s_eta", "V_s_eta"), ("V_s_xi", "V_s_xi"),
                           ("V_z_eta", "V_z_eta"), ("V_z_xi", "V_z_xi"), ("Y0", "Y0"),
                           ("V0_s_eta", "V0_s_eta"), ("V0_s_xi", "V0_s_xi"),
                           ("V0_z_eta", "V0_z_eta"), ("V0_z_xi", "V0_z_xi"),
                           ("DsDeta_over_J", "DsDeta_over_J_sol"),
                           ("DzDeta_over_J", "DzDeta_over_J_sol"),
                           ("DsDxi_over_J", "DsDxi_over_J_sol"),
                           ("DzDxi_over_J", "DzDxi_over_J_sol"),
                           ("inv_s", "inv_s_solid"), ("delta_mu", "delta_mu"),
                           ("delta_kappa", "delta_kappa")):
                setattr(a, fld, K(f32(n)))
            for fld in ("exp_w_j_deltat", "ts_fac_t", "ts_fac_tm1", "a_j_mu", "a_j_kappa"):
                setattr(a, fld, K(np.ascontiguousarray(st[fld], np.float64)))
            _check(L.axh_upload_attenuation_full(h, C.byref(a)))
        elif d["anel"]:
            a = Att()
            for fld, n in (("Y", "Y_cg4"), ("V_s_eta", "V_s_eta_cg4"), ("V_s_xi", "V_s_xi_cg4"),
                           ("V_z_eta", "V_z_eta_cg4"), ("V_z_xi", "V_z_xi_cg4"),
                           ("DsDeta_over_J", "DsDeta_over_J_sol_cg4"),
                           ("DzDeta_over_J", "DzDeta_over_J_sol_cg4"),
                           ("DsDxi_over_J", "DsDxi_over_J_sol_cg4"),
                           ("DzDxi_over_J", "DzDxi_over_J_sol_cg4"),
                           ("inv_s", "inv_s_solid_cg4"), ("delta_mu", "delta_mu_cg4"),
                           ("delta_kappa", "delta_kappa_cg4")):
                setattr(a, fld, K(f32(n)))
            for fld in ("exp_w_j_deltat", "ts_fac_t", "ts_fac_tm1", "a_j_mu", "a_j_kappa"):
                setattr(a, fld, K(np.ascontiguousarray(st[fld], np.float64)))
            _check(L.axh_upload_attenuation_cg4(h, C.byref(a)))
        if d["num_rec"] > 0:
            _check(L.axh_upload_receivers(h, K(i32("recfile_el"))))
        if "inv_rho_fluid" in st.arrays:
            _check(L.axh_upload_fluid_gradient(h, *[K(f32(n)) for n in FLUID_GRADIENT]))
        if "kwf_npoints" in st.arrays:
            nps, npf = (int(x) for x in st["kwf_npoints"])
            _check(L.axh_upload_kwf_grid(h, nps, npf, K(i32("kwf_mapping")),
                                         K(i32("kwf_mask"))))
        return self

    def close(self):
        if self._h:
            self._L.axh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- running ------------------------------------------------------------
    def run(self, nsteps: int | None = None) -> np.ndarray:
        """Advance nsteps (default: to niter); returns the receiver samples of
        the sampled steps, shape (nsamples, num_rec, 3)."""
        it0 = self.iter + 1
        it1 = self.dims["niter"] if nsteps is None else min(self.dims["niter"], self.iter + nsteps)
        n = self._L.axh_seis_samples(self._h, it0, it1)
        seis = np.zeros((max(n, 1), max(self.dims["num_rec"], 1), 3), np.float32)
        rc = self._L.axh_run(self._h, it0, it1, _p(seis))
        self.iter = it1 if rc in (AXH_OK, AXH_ERR_BLOWUP) else self.iter
        self._ok(rc)
        return seis[:n]

    def run_begin(self, nsteps: int) -> np.ndarray:
        """Queue nsteps on the GPU and return at once; the returned array holds
        the samples after run_end()."""
        it0 = self.iter + 1
        it1 = min(self.dims["niter"], self.iter + nsteps)
        n = self._L.axh_seis_samples(self._h, it0, it1)
        self._pending = (np.zeros((max(n, 1), max(self.dims["num_rec"], 1), 3), np.float32),
                         n, it1)
        self._ok(self._L.axh_run_begin(self._h, it0, it1, _p(self._pending[0])))
        return self._pending[0][:n]

    def run_end(self):
        it1 = self._pending[2]
        self._pending = None
        rc = self._L.axh_run_end(self._h)
        if rc in (AXH_OK, AXH_ERR_BLOWUP):
            self.iter = it1
        self._ok(rc)

    # --- kernel wavefields (displ_only) ---------------------------------------
    def kwf_grid(self, ibeg=1, iend=3, jbeg=1, jend=3, elem_mask=None):
        """Build the kernel-wavefield grid (build_kwf_grid; the reference's
        defaults select GLL points 1..3 x 1..3 of every element plus the axis
        points); returns (npoint_solid, npoint_fluid).  Before the first run."""
        em = None if elem_mask is None else np.ascontiguousarray(elem_mask, np.int32)
        nps, npf = C.c_int32(0), C.c_int32(0)
        self._ok(self._L.axh_kwf_grid(self._h, ibeg, iend, jbeg, jend, _p(em), C.byref(nps),
                                      C.byref(npf)))
        return nps.value, npf.value

    def kwf_get_grid(self):
        """(mapping_ijel_ikwf, kwf_mask, npoint_solid, npoint_fluid); the arrays
        [nel_solid + nel_fluid][25], mapping 1-based at the masked entries."""
        n = 25 * (self.dims["nel_solid"] + self.dims["nel_fluid"])
        mapping = np.zeros(n, np.int32)
        mask = np.zeros(n, np.int32)
        nps, npf = C.c_int32(0), C.c_int32(0)
        self._ok(self._L.axh_kwf_get_grid(self._h, _p(mapping), _p(mask), C.byref(nps),
                                          C.byref(npf)))
        return mapping, mask, nps.value, npf.value

    def set_kwf(self, strain_it: int, src_mask: bool = True, ring_snapshots: int = 0):
        """Snapshots every strain_it steps (and of the initial fields);
        src_mask zeroes the source elements (src_dump_type 'mask');
        ring_snapshots bounds the device ring (0: automatic)."""
        self._ok(self._L.axh_set_kwf(self._h, int(strain_it), int(bool(src_mask)),
                                     int(ring_snapshots)))

    def run_kwf_begin(self, nsteps: int):
        """run_kwf queued: returns (seis, kwf) at once, filled after run_end()."""
        it0 = self.iter + 1
        it1 = min(self.dims["niter"], self.iter + nsteps)
        n = self._L.axh_seis_samples(self._h, it0, it1)
        seis = np.zeros((max(n, 1), max(self.dims["num_rec"], 1), 3), np.float32)
        nsnap = self._L.axh_kwf_snapshots(self._h, it0, it1)
        nps, npf = C.c_int32(0), C.c_int32(0)
        self._ok(self._L.axh_kwf_get_grid(self._h, None, None, C.byref(nps), C.byref(npf)))
        ncomp = 2 if self.dims["src_type"] == 0 else 3
        npt = nps.value + npf.value
        kwf = np.zeros((max(nsnap, 1), ncomp, max(npt, 1)), np.float32)
        self._pending = (seis, n, it1, kwf)
        self._ok(self._L.axh_run_kwf_begin(self._h, it0, it1, _p(seis), _p(kwf)))
        return seis[:n], kwf[:nsnap, :, :npt]

    def run_kwf(self, nsteps: int | None = None):
        """run(), plus the kernel-wavefield snapshots of the steps it covers:
        (seis, kwf) with kwf of shape (nsnap, ncomp, npoint), components s, z
        (monopole) or s, phi, z."""
        it0 = self.iter + 1
        it1 = self.dims["niter"] if nsteps is None else min(self.dims["niter"], self.iter + nsteps)
        n = self._L.axh_seis_samples(self._h, it0, it1)
        seis = np.zeros((max(n, 1), max(self.dims["num_rec"], 1), 3), np.float32)
        nsnap = self._L.axh_kwf_snapshots(self._h, it0, it1)
        nps, npf = C.c_int32(0), C.c_int32(0)
        self._ok(self._L.axh_kwf_get_grid(self._h, None, None, C.byref(nps), C.byref(npf)))
        ncomp = 2 if self.dims["src_type"] == 0 else 3
        kwf = np.zeros((max(nsnap, 1), ncomp, max(nps.value + npf.value, 1)), np.float32)
        rc = self._L.axh_run_kwf(self._h, it0, it1, _p(seis), _p(kwf))
        self.iter = it1 if rc in (AXH_OK, AXH_ERR_BLOWUP) else self.iter
        self._ok(rc)
        return seis[:n], kwf[:nsnap, :, :nps.value + npf.value]

    def fields(self):
        """(disp, velo, chi, dchi, ddchi) in the reference's element-local layout."""
        d = self.dims
        ns, nf = 25 * d["nel_solid"], 25 * d["nel_fluid"]
        disp = np.zeros(3 * ns, np.float32)
        velo = np.zeros(3 * ns, np.float32)
        chi = np.zeros(max(nf, 1), np.float32)
        dchi = np.zeros(max(nf, 1), np.float32)
        ddchi = np.zeros(max(nf, 1), np.float32)
        self._ok(self._L.axh_download_fields(self._h, _p(disp), _p(velo), _p(chi), _p(dchi),
                                             _p(ddchi)))
        return disp, velo, chi[:nf], dchi[:nf], ddchi[:nf]

    def set_fields(self, disp=None, velo=None, chi=None, dchi=None, ddchi=None):
        f = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        arrs = [f(disp), f(velo), f(chi), f(dchi), f(ddchi)]
        self._ok(self._L.axh_upload_fields(self._h, *[_p(a) for a in arrs]))

    def last_run_ms(self) -> float:
        v = C.c_double()
        self._ok(self._L.axh_last_run_ms(self._h, C.byref(v)))
        return v.value

    def device_bytes(self) -> int:
        v = C.c_size_t()
        self._ok(self._L.axh_device_bytes(self._h, C.byref(v)))
        return v.value

    def set_profiling(self, on: bool):
        self._ok(self._L.axh_set_profiling(self._h, 1 if on else 0))

    def set_graph_steps(self, steps: int):
        """Replay a captured HIP graph of `steps` time steps (0: launch every
        kernel directly); results are identical either way."""
        self._ok(self._L.axh_set_graph_steps(self._h, steps))

    def kernel_stats(self) -> dict:
        """Per kernel class: (total ms, launches) since profiling was enabled."""
        return _kernel_stats(self._L.axh_kernel_stats, self._h)



def _kernel_stats(fn, handle) -> dict:
    out = {}
    for k in range(16):
        name = library().axh_kernel_name(k)
        if not name:
            break
        ms, n = C.c_double(), C.c_int64()
        _check(fn(handle, k, C.byref(ms), C.byref(n)))
        out[name.decode()] = (ms.value, n.value)
    return out


class Batch:
    """Several wavefields on one mesh (one Solver each, differing in source and
    receivers), stepped together: every kernel launch advances all of them and
    reads the mesh terms once.  Each member's seismograms and fields are the
    ones it gets when run alone."""

    def __init__(self, solvers, group_width: int | None = None, graph_steps: int = 0):
        self.solvers = list(solvers)
        n = len(self.solvers)
        if len({s.fused for s in self.solvers}) > 1:
            raise ValueError("batch members must come from one build of the library")
        self._L = self.solvers[0]._L
        self.fused = self.solvers[0].fused
        arr = (C.c_void_p * n)(*[s._h for s in self.solvers])
        self._h = C.c_void_p()
        self._ok(self._L.axh_batch_create(arr, n, C.byref(self._h)))
        if group_width is not None:
            self._ok(self._L.axh_batch_set_group_width(self._h, group_width))
        if graph_steps:
            self.set_graph_steps(graph_steps)

    def _ok(self, rc):
        _check(rc, self._L)

    def close(self):
        if self._h:
            self._L.axh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def iter(self) -> int:
        return self.solvers[0].iter

    def run(self, nsteps: int | None = None) -> list:
        """Advance every member nsteps (default: to the smallest niter);
        returns one (nsamples, num_rec, 3) array per member."""
        it0 = self.iter + 1
        top = min(s.dims["niter"] for s in self.solvers)
        it1 = top if nsteps is None else min(top, self.iter + nsteps)
        outs, ptrs = [], []
        for s in self.solvers:
            n = s._L.axh_seis_samples(s._h, it0, it1)
            a = np.zeros((max(n, 1), max(s.dims["num_rec"], 1), 3), np.float32)
            outs.append((a, n))
            ptrs.append(a.ctypes.data_as(C.c_void_p))
        arr = (C.c_void_p * len(ptrs))(*ptrs)
        rc = self._L.axh_batch_run(self._h, it0, it1, arr)
        if rc in (AXH_OK, AXH_ERR_BLOWUP):
            for s in self.solvers:
                s.iter = it1
        self._ok(rc)
        return [a[:n] for a, n in outs]

    def step(self, nsteps: int) -> list:
        """Advance every member nsteps from its own current step (members may
        be at different steps); returns one (nsamples, num_rec, 3) array per
        member."""
        outs, ptrs = [], []
        for s in self.solvers:
            n = s._L.axh_seis_samples(s._h, s.iter + 1, s.iter + nsteps)
            a = np.zeros((max(n, 1), max(s.dims["num_rec"], 1), 3), np.float32)
            outs.append((a, n))
            ptrs.append(a.ctypes.data_as(C.c_void_p))
        arr = (C.c_void_p * len(ptrs))(*ptrs)
        rc = self._L.axh_batch_step(self._h, nsteps, arr)
        if rc in (AXH_OK, AXH_ERR_BLOWUP):
            for s in self.solvers:
                s.iter += nsteps
        self._ok(rc)
        return [a[:n] for a, n in outs]

    def step_kwf(self, nsteps: int):
        """step(), plus every member's kernel-wavefield snapshots of its steps
        (each member's set_kwf and grid): (seis list, kwf list), kwf of shape
        (nsnap, ncomp, npoint) per member."""
        outs, ptrs, kouts, kptrs = [], [], [], []
        for s in self.solvers:
            n = s._L.axh_seis_samples(s._h, s.iter + 1, s.iter + nsteps)
            a = np.zeros((max(n, 1), max(s.dims["num_rec"], 1), 3), np.float32)
            outs.append((a, n))
            ptrs.append(a.ctypes.data_as(C.c_void_p))
            nsnap = s._L.axh_kwf_snapshots(s._h, s.iter + 1, s.iter + nsteps)
            nps, npf = C.c_int32(0), C.c_int32(0)
            s._ok(s._L.axh_kwf_get_grid(s._h, None, None, C.byref(nps), C.byref(npf)))
            npt = nps.value + npf.value
            ncomp = 2 if s.dims["src_type"] == 0 else 3
            k = np.zeros((max(nsnap, 1), ncomp, max(npt, 1)), np.float32)
            kouts.append((k, nsnap, npt))
            kptrs.append(k.ctypes.data_as(C.c_void_p))
        arr = (C.c_void_p * len(ptrs))(*ptrs)
        karr = (C.c_void_p * len(kptrs))(*kptrs)
        rc = self._L.axh_batch_step_kwf(self._h, nsteps, arr, karr)
        if rc in (AXH_OK, AXH_ERR_BLOWUP):
            for s in self.solvers:
                s.iter += nsteps
        self._ok(rc)
        return [a[:n] for a, n in outs], [k[:m, :, :npt] for k, m, npt in kouts]

    def last_run_ms(self) -> float:
        v = C.c_double()
        self._ok(self._L.axh_batch_last_run_ms(self._h, C.byref(v)))
        return v.value

    def set_profiling(self, on: bool):
        self._ok(self._L.axh_batch_set_profiling(self._h, 1 if on else 0))

    def set_graph_steps(self, steps: int):
        """Solver.set_graph_steps for every group of the batch."""
        self._ok(self._L.axh_batch_set_graph_steps(self._h, steps))

    def kernel_stats(self) -> dict:
        """Per kernel class: (total ms, launches) since profiling was enabled."""
        return _kernel_stats(self._L.axh_batch_kernel_stats, self._h)


# element counts of the arrays axh_meshdb_array returns: name -> (dtype, size)
def _meshdb_shapes(I):
    ns, nf, nb, ne = I["nel_solid"], I["nel_fluid"], I["nel_bdry"], I["nelem"]
    i4, f4, f8 = np.int32, np.float32, np.float64
    return {
        "igloc_solid": (i4, 25 * ns), "igloc_fluid": (i4, 25 * nf),
        "ielsolid": (i4, ns), "ielfluid": (i4, nf),
        "axis_solid": (i4, ns), "axis_fluid": (i4, nf),
        "ax_el_solid": (i4, I["naxel_solid"]), "ax_el_fluid": (i4, I["naxel_fluid"]),
        "bdry_solid_el": (i4, nb if I["have_bdry"] else 0),
        "bdry_fluid_el": (i4, nb if I["have_bdry"] else 0),
        "bdry_jpol_solid": (i4, nb if I["have_bdry"] else 0),
        "bdry_jpol_fluid": (i4, nb if I["have_bdry"] else 0),
        "G0": (f4, 5), "G1": (f4, 25), "G1T": (f4, 25), "G2": (f4, 25), "G2T": (f4, 25),
        "crd_s": (f8, I["npoin"]), "crd_z": (f8, I["npoin"]),
        "lnods": (i4, 8 * ne), "coarsing": (i4, ne), "eltype": (np.uint8, 6 * ne),
    }


class MeshDB:
    """The reference MESHER's mesh database (Mesh/meshdb.dat0000) read by the
    library (axh_meshdb_*): the counts of its header and its arrays in the
    layout the SOLVER reads them (SOLVER/get_mesh.f90:47-409)."""

    def __init__(self, path: str):
        self._h = C.c_void_p()
        _check(library().axh_meshdb_open(path.encode(), C.byref(self._h)))
        raw = MeshDBInfo()
        _check(library().axh_meshdb_get_info(self._h, C.byref(raw)))
        self.info = {k: getattr(raw, k) for k, _ in MeshDBInfo._fields_}
        self.info["bkgrdmodel"] = raw.bkgrdmodel.decode(errors="replace").strip()

    def array(self, name: str) -> np.ndarray:
        dt, n = _meshdb_shapes(self.info)[name]
        a = np.zeros(max(n, 1), dt)
        _check(library().axh_meshdb_array(self._h, name.encode(), _p(a), n * a.itemsize))
        return a[:n]

    def upload(self, solver: "Solver"):
        """Give a context its mesh-only uploads from the database."""
        _check(library().axh_meshdb_upload(self._h, solver._h))

    def close(self):
        if self._h:
            library().axh_meshdb_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
