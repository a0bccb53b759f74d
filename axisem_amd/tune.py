"""Autotuner of the launch schedule.

Picks, by measuring on the actual mesh, how a run is launched:

* ``graph_steps`` -- time steps per replayed HIP graph (0: every kernel
  launched directly).  Small meshes are bound by launch latency and gain from
  longer graphs; large ones are indifferent.
* ``group_width`` -- wavefields per kernel instance of a batch (1-4): wider
  groups read the mesh terms fewer times, narrower ones need fewer registers
  and keep more of their fields in the caches.

Every candidate runs on throwaway contexts loaded from the same state files
(the production contexts are never stepped), a warm-up call first, then the
device time of a few timed steps (HIP events on the library's stream).  The
schedules give bit-identical results; only their speed differs.  The
reference has no counterpart (its loop is fixed host code).
"""
from __future__ import annotations

from .solver import Batch, Solver

GRAPH_CANDIDATES = (0, 4, 8, 16, 32)
WIDTH_CANDIDATES = (2, 3, 4)


def _device_ms_per_step(runner, graph_steps: int, steps: int) -> float:
    runner.set_graph_steps(graph_steps)
    runner.run(max(graph_steps, 1) + 1)          # capture the graph outside the timing
    runner.run(steps)
    return runner.last_run_ms() / steps


def tune_solo(state_path: str, steps: int = 64, candidates=GRAPH_CANDIDATES) -> dict:
    """Best graph_steps for one wavefield of this state; returns
    {"graph_steps": S, "ms_per_step": {S: device ms per step}}."""
    with Solver.from_state_file(state_path) as s:
        need = sum(steps + max(c, 1) + 1 for c in candidates) + 2
        if need > s.dims["niter"]:
            return {"graph_steps": 0, "ms_per_step": {}}
        s.run(2)
        t = {c: _device_ms_per_step(s, c, steps) for c in candidates}
    return {"graph_steps": min(t, key=t.get), "ms_per_step": t}


def tune_batch(state_paths: list, steps: int = 32, widths=WIDTH_CANDIDATES,
               candidates=GRAPH_CANDIDATES) -> dict:
    """Best (group_width, graph_steps) for a batch of these states; returns
    {"group_width": w, "graph_steps": S, "ms_per_step": {"w/S": ms}}."""
    t = {}
    for w in widths:
        if w > len(state_paths) and w != widths[0]:
            continue
        members = [Solver.from_state_file(p) for p in state_paths]
        try:
            need = sum(steps + max(c, 1) + 1 for c in candidates) + 2
            if need > min(m.dims["niter"] for m in members):
                return {"group_width": 4, "graph_steps": 0, "ms_per_step": {}}
            with Batch(members, group_width=w) as b:
                b.run(2)
                for c in candidates:
                    t[(w, c)] = _device_ms_per_step(b, c, steps)
        finally:
            for m in members:
                m.close()
    best = min(t, key=t.get)
    return {"group_width": best[0], "graph_steps": best[1],
            "ms_per_step": {"%d/%d" % k: v for k, v in t.items()}}
