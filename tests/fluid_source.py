"""States with the source in the fluid (add_source_fl).

The reference's source location cannot place a source in the fluid on these
meshes (SOLVER/source.f90:531 indexes ielsolid with a fluid element), so a
fluid-source state is made from a recorded one: fluid_src = 1, ielsrc names
fluid elements (1-based in the fluid numbering, as add_source_fl indexes
chi(:,:,ielsrc(iel)), SOLVER/time_evol_wave.F90:1062-1076) and source_term
holds source_term_fl(0:4,0:4,8) in its first 200 entries
(SOLVER/data_source.f90:47-49).
"""
from __future__ import annotations

import numpy as np

from axisem_amd.state import State


def fluid_source_state(st: State, elems, seed: int = 0, interior_only: bool = False,
                       stf=None) -> State:
    """A copy of st with its source moved into the fluid elements `elems`
    (0-based fluid element indices): random source_term_fl on their points
    (only on the 9 interior points with interior_only, which have one element
    copy each), and the source time function table replaced by stf if
    given."""
    out = State(dict(st.dims), {k: (np.array(a, copy=True), fd) for k, (a, fd) in
                                st.arrays.items()})
    d = out.dims
    if d["nel_fluid"] == 0:
        raise ValueError("the mesh has no fluid")
    rng = np.random.default_rng(seed)
    n = len(elems)
    assert 1 <= n <= 8
    term = np.zeros(3 * 8 * 25, np.float32)
    vals = (rng.standard_normal((n, 5, 5)) * 1e3).astype(np.float32)   # [k][j][i]
    if interior_only:
        mask = np.zeros((5, 5), bool)
        mask[1:4, 1:4] = True
        vals = vals * mask
    term[: n * 25] = vals.reshape(-1)
    out.arrays["source_term"] = (term, st.shape("source_term"))
    isrc = np.asarray(elems, np.int32) + 1
    out.arrays["ielsrc"] = (isrc, (n,))
    d["nelsrc"] = n
    d["fluid_src"] = 1
    if stf is not None:
        a = np.asarray(stf, np.float32)
        out.arrays["stf"] = (a, (a.size,))
    return out


def zero_operators(st: State) -> State:
    """A copy of st with every stiffness term and the solid-fluid coupling
    matrix zeroed: what the loop then does to the fluid is the source
    injection, the assembly, the inverse mass and the update alone."""
    out = State(dict(st.dims), {k: (np.array(a, copy=True), fd) for k, (a, fd) in
                                st.arrays.items()})
    names = [k for k in out.arrays if k.startswith(("M", "bdry_matr"))]
    for k in names:
        a, fd = out.arrays[k]
        out.arrays[k] = (np.zeros_like(a), fd)
    return out


def pick_fluid_elements(st: State, n: int = 3):
    """n fluid elements spread over the fluid: an axial one (if any) and
    others from the middle and the end of the numbering."""
    nf = st.dims["nel_fluid"]
    ax = np.nonzero(np.asarray(st["axis_fluid"]) != 0)[0]
    picks = []
    if len(ax):
        picks.append(int(ax[len(ax) // 2]))
    for f in (0.5, 0.83, 0.31, 0.97):
        e = int(f * (nf - 1))
        if e not in picks:
            picks.append(e)
    return picks[:n]
