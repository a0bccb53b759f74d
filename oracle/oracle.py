"""ctypes wrapper of the CPU oracle (oracle/axisem_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Never used by the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build_ref  # noqa: F401  (build on demand in-tree)
            build_ref.build_oracle()
        L = C.CDLL(LIB_PATH)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        L.orc_load.restype = P
        L.orc_load.argtypes = [C.c_char_p]
        L.orc_free.argtypes = [P]
        L.orc_run.restype = I
        L.orc_run.argtypes = [P, I, I, P]
        for f in ("orc_nel_solid", "orc_nel_fluid", "orc_iter", "orc_src_type"):
            getattr(L, f).restype = I
            getattr(L, f).argtypes = [P]
        L.orc_get_fields.argtypes = [P, P, P, P, P, P]
        L.orc_set_fields.argtypes = [P, P, P, P, P, P]
        L.orc_solid_stiffness.argtypes = [P, P, P]
        L.orc_fluid_stiffness.argtypes = [P, P, P]
        L.orc_anel_stiffness.argtypes = [P, P, P]
        L.orc_memvar_update.argtypes = [P, P]
        L.orc_get_memvars.argtypes = [P, P, P, P]
        L.orc_assemble_solid.argtypes = [P, P, I]
        L.orc_assemble_fluid.argtypes = [P, P]
        L.orc_symplectic_coefficients.restype = I
        L.orc_symplectic_coefficients.argtypes = [I, D, P, P, P]
        L.orc_mxm.argtypes = [P, P, P]
        L.orc_strain_cg4.argtypes = [P, P, I, P]
        L.orc_kwf_grid.argtypes = [P, I, I, I, I, P, P, P, P, P]
        L.orc_kwf_sample.restype = I
        L.orc_kwf_sample.argtypes = [P, P, P, I, I, I, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """CPU restatement of the time loop on one state file."""

    def __init__(self, state_path: str):
        self.h = lib().orc_load(state_path.encode())
        if not self.h:
            raise RuntimeError("oracle could not load " + state_path)
        self.nel_s = lib().orc_nel_solid(self.h)
        self.nel_f = lib().orc_nel_fluid(self.h)

    def close(self):
        if self.h:
            lib().orc_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def iter(self) -> int:
        return lib().orc_iter(self.h)

    def run(self, it0: int, it1: int, num_rec: int, seis_it: int):
        nsamp = it1 // seis_it - (it0 - 1) // seis_it if num_rec else 0
        out = np.zeros((max(nsamp, 1), max(num_rec, 1), 3), np.float32)
        n = lib().orc_run(self.h, it0, it1, _p(out))
        if n < 0:
            raise FloatingPointError("oracle blow-up at step %d" % -n)
        return out[:n]

    def fields(self):
        ns, nf = self.nel_s * 25, self.nel_f * 25
        d = np.zeros(3 * ns, np.float32)
        v = np.zeros(3 * ns, np.float32)
        c = np.zeros(max(nf, 1), np.float32)
        dc = np.zeros(max(nf, 1), np.float32)
        ddc = np.zeros(max(nf, 1), np.float32)
        lib().orc_get_fields(self.h, _p(d), _p(v), _p(c), _p(dc), _p(ddc))
        return d, v, c[:nf], dc[:nf], ddc[:nf]

    def set_fields(self, disp=None, velo=None, chi=None, dchi=None, ddchi=None):
        f = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        arrs = [f(disp), f(velo), f(chi), f(dchi), f(ddchi)]
        lib().orc_set_fields(self.h, *[_p(a) for a in arrs])

    def solid_stiffness(self, u):
        u = np.ascontiguousarray(u, np.float32)
        out = np.zeros_like(u)
        lib().orc_solid_stiffness(self.h, _p(u), _p(out))
        return out

    def fluid_stiffness(self, chi):
        chi = np.ascontiguousarray(chi, np.float32)
        out = np.zeros_like(chi)
        lib().orc_fluid_stiffness(self.h, _p(chi), _p(out))
        return out

    def kwf_grid(self, ibeg=1, iend=3, jbeg=1, jend=3, elem_mask=None):
        """build_kwf_grid: (mapping, mask, npoint_solid, npoint_fluid), the
        first two [nel_solid + nel_fluid][25] (1-based numbers, 0: none)."""
        n = (self.nel_s + self.nel_f) * 25
        mapping = np.zeros(max(n, 1), np.int32)
        mask = np.zeros(max(n, 1), np.int32)
        em = None if elem_mask is None else np.ascontiguousarray(elem_mask, np.int32)
        nps, npf = C.c_int32(0), C.c_int32(0)
        lib().orc_kwf_grid(self.h, ibeg, iend, jbeg, jend, _p(em), _p(mapping), _p(mask),
                           C.byref(nps), C.byref(npf))
        return mapping[:n], mask[:n], nps.value, npf.value

    def kwf_sample(self, grid, src_mask=True):
        """The displ_only kernel-wavefield sample of the current fields:
        [ncomp][npoint] (monopole: s, z; else s, phi, z)."""
        mapping, mask, nps, npf = grid
        nc = 2 if lib_src_is_mono(self) else 3
        out = np.zeros((nc, max(nps + npf, 1)), np.float32)
        if lib().orc_kwf_sample(self.h, _p(mapping), _p(mask), nps, npf, int(src_mask),
                                _p(out)) != 0:
            raise ValueError("the state holds no fluid gradient arrays")
        return out[:, :nps + npf]

    def assemble_solid(self, vec, nc=3):
        vec = np.array(vec, np.float32, copy=True)
        lib().orc_assemble_solid(self.h, _p(vec), nc)
        return vec

    def assemble_fluid(self, vec):
        vec = np.array(vec, np.float32, copy=True)
        lib().orc_assemble_fluid(self.h, _p(vec))
        return vec


def lib_src_is_mono(o: Oracle) -> bool:
    return lib().orc_src_type(o.h) == 0


def symplectic_coefficients(scheme: int, dt: float):
    cd, cv, cf = (np.zeros(40) for _ in range(3))
    n = lib().orc_symplectic_coefficients(scheme, dt, _p(cd), _p(cv), _p(cf))
    return cd[: n + 1], cv[:n], cf[:n]


def mxm(a, b):
    """5x5 product in the reference's summation order (Fortran layout)."""
    a = np.asfortranarray(a, np.float32)
    b = np.asfortranarray(b, np.float32)
    c = np.zeros((5, 5), np.float32, order="F")
    lib().orc_mxm(a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                  c.ctypes.data_as(C.c_void_p))
    return c
