"""Request loop over the GPU time loop: continuous batching of solves.

Solves (one source on one mesh) arrive over time, either as a state file
already prepared by the reference's setup or as a run configuration that the
loop prepares when the request arrives: one drop-in host per mesh, model,
time scheme and source type stays running in record mode
(axisem_amd.inputs.SourceRecorder), so only the first request of such a key
pays the full setup and later ones only the reference's source routines.
The loop

* keeps meshes resident on the device (one device mesh part per distinct mesh
  fingerprint, shared by every request on it, kept while memory allows);
* admits arrived requests in arrival order as device memory allows (the
  footprint of a request is its wavefield part, plus its mesh part when that
  mesh is not resident; idle meshes are evicted, least recently used first);
* co-batches the admitted requests that share a mesh numbering and time
  scheme (axh_batch_key), whatever their source types (the four runs of a
  moment tensor share one batch, one group per type): they advance together,
  one kernel launch per step and stage for each group, each member from its
  own step (axh_batch_step);
* re-forms a batch between chunks of steps when requests join or finish, so a
  request that arrives while others run starts at the next chunk boundary;
* runs the batches of different meshes concurrently (one host thread and one
  HIP stream each), so their kernels overlap on the device;
* records per request the time from arrival to its complete seismograms
  (including its preparation when it arrived as a configuration).

Each request's seismograms are the ones it gets alone (the batched kernels
are bit-identical per wavefield).  The reference has no counterpart: it runs
one source per SOLVER job (submit.py:276-376).
"""
from __future__ import annotations

import collections
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    state_path: str | None
    arrival_s: float = 0.0          # seconds after the loop starts
    mesh_dir: str | None = None     # prepared on arrival from (mesh_dir, cfg)
    cfg: object = None
    ready_s: float | None = None    # state prepared
    admitted_s: float | None = None
    done_s: float | None = None
    niter: int = 0
    steps: int | None = None        # steps to run (None: the state's niter)
    gll_per_step: int = 0
    key: tuple | None = None        # batch key
    fp: tuple | None = None         # mesh fingerprint
    seis: np.ndarray | None = None  # (nsamples, num_rec, 3)
    error: str | None = None
    _sol: object = field(default=None, repr=False)   # loaded, not yet admitted

    @property
    def latency_s(self):
        return None if self.done_s is None else self.done_s - self.arrival_s

    @property
    def setup_s(self):
        return None if self.ready_s is None else self.ready_s - self.arrival_s


class GpuBackend:
    """The library through axisem_amd.solver (ctypes)."""

    def __init__(self):
        from . import solver as S
        self.S = S

    def load(self, path):
        return self.S.Solver.from_state_file(path)

    def key(self, sol):
        import ctypes as C
        fp = (C.c_uint64 * 2)()
        self.S._check(self.S.library().axh_mesh_fingerprint(sol._h, fp))
        return (int(fp[0]), int(fp[1]))

    def batch_key(self, sol):
        import ctypes as C
        k = (C.c_uint64 * 2)()
        self.S._check(self.S.library().axh_batch_key(sol._h, k))
        return (int(k[0]), int(k[1]))

    def footprint(self, sol):
        import ctypes as C
        m, w = C.c_size_t(), C.c_size_t()
        self.S._check(self.S.library().axh_footprint(sol._h, C.byref(m), C.byref(w)))
        return m.value, w.value

    def mesh_get(self, sol):
        import ctypes as C
        h = C.c_void_p()
        self.S._check(self.S.library().axh_mesh_get(sol._h, C.byref(h)))
        return h

    def mesh_bind(self, sol, mesh):
        self.S._check(self.S.library().axh_mesh_bind(sol._h, mesh))

    def mesh_release(self, mesh):
        self.S._check(self.S.library().axh_mesh_release(mesh))

    def batch(self, sols, group_width):
        return self.S.Batch(sols, group_width=group_width)


@dataclass
class _Active:
    req: Request
    sol: object
    wave_bytes: int
    chunks: list = field(default_factory=list)


class Preparer:
    """States of requests that arrive as (mesh, configuration): one recorder
    (a drop-in host in record mode fed through a pipe) per mesh, model, time
    scheme and source type, started by the first request of that key.

    A recorder records every source for the longest duration it serves
    (max_length_s); a request runs only the first niter steps its own
    duration needs, ceiling((length + smallval_dble) / deltat) as the
    reference counts them (SOLVER/parameters.F90:901).  Nothing in a step
    depends on the run's length (the source time function and the receiver
    sampling are functions of the step), so that prefix is the shorter run
    bit for bit."""

    def __init__(self, host_exe: str, workdir: str, max_length_s: float = 3600.0):
        import os
        self.host = host_exe
        self.max_length = max_length_s
        self.workdir = workdir
        os.makedirs(workdir, exist_ok=True)
        self._rec = {}
        self._lock = threading.Lock()
        self._n = 0

    def _key(self, mesh_dir, cfg):
        from .inputs import SOURCE_KIND
        base = {k: v for k, v in cfg.__dict__.items()
                if k not in ("source_type", "source_depth_km", "source_amplitude",
                             "seismogram_length_s")}
        return (mesh_dir, repr(sorted(base.items())), SOURCE_KIND[cfg.source_type.lower()])

    @staticmethod
    def steps_for(length_s: float, deltat: float) -> int:
        """niter of SOLVER/parameters.F90:901 (smallval_dble is the REAL(4)
        literal 1e-11 widened to REAL(8))."""
        import math
        return int(math.ceil((length_s + float(np.float32(1e-11))) / deltat))

    def prepare(self, mesh_dir, cfg):
        """(state path, steps of this request)."""
        from .inputs import RunConfig
        from .state import read_dims
        if cfg.seismogram_length_s > self.max_length:
            raise ValueError("longer than the recorders' %.0f s" % self.max_length)
        rec_cfg = RunConfig(**{**cfg.__dict__, "seismogram_length_s": self.max_length})
        path = self._record(mesh_dir, rec_cfg)
        d = read_dims(path)
        return path, min(d["niter"], self.steps_for(cfg.seismogram_length_s, d["deltat"]))

    def _record(self, mesh_dir, cfg) -> str:
        import os
        from .inputs import SourceRecorder
        key = self._key(mesh_dir, cfg)
        with self._lock:
            self._n += 1
            n = self._n
            entry = self._rec.get(key)
            if entry is None:
                entry = self._rec[key] = [threading.Lock(), None]
        path = os.path.join(self.workdir, "req%04d.bin" % n)
        with entry[0]:
            if entry[1] is None:
                entry[1] = SourceRecorder(self.host, os.path.join(self.workdir, "host%04d" % n),
                                          mesh_dir, cfg, path)
                return entry[1].first_state
            return entry[1].record(cfg, path)

    def close(self):
        for lock, rec in self._rec.values():
            if rec is not None:
                rec.close()
        self._rec.clear()


class RequestLoop:
    def __init__(self, hbm_budget_bytes: float = 200e9, max_members: int = 16,
                 group_width: int = 4, chunk_steps: int = 100, backend=None,
                 clock=time.perf_counter, log=None, preparer: Preparer | None = None):
        self.budget = float(hbm_budget_bytes)
        self.max_members = max_members
        self.group_width = group_width
        self.chunk = chunk_steps
        self.be = backend or GpuBackend()
        self.clock = clock
        self.log = log or (lambda msg: None)
        self.requests: list[Request] = []
        self.used = 0.0
        self.peak_used = 0.0
        self._meshes = collections.OrderedDict()  # key -> [handle, bytes, active count]
        self._active: dict = collections.defaultdict(list)
        self._batches: dict = {}                  # key -> (batch, member ids)
        self.chunks_run = 0
        self.max_cobatched = 0
        self.preparer = preparer

    def submit(self, state_path: str, arrival_s: float = 0.0, steps: int | None = None) -> Request:
        """A request with its state prepared; steps (default: the state's niter)
        is how many steps it runs."""
        r = Request(rid=len(self.requests), state_path=state_path, arrival_s=arrival_s,
                    steps=steps)
        self.requests.append(r)
        return r

    def submit_config(self, mesh_dir: str, cfg, arrival_s: float = 0.0) -> Request:
        """A request prepared by the loop's Preparer when it arrives."""
        if self.preparer is None:
            raise ValueError("the loop has no preparer")
        r = Request(rid=len(self.requests), state_path=None, arrival_s=arrival_s,
                    mesh_dir=mesh_dir, cfg=cfg)
        self.requests.append(r)
        return r

    # -- admission -----------------------------------------------------------
    def _evict_idle(self, need: float) -> None:
        for key in list(self._meshes):
            if self.used + need <= self.budget:
                return
            h, nbytes, nact = self._meshes[key]
            if nact == 0:
                self.be.mesh_release(h)
                del self._meshes[key]
                self.used -= nbytes
                self.log("evicted mesh %x (%.1f MB)" % (key[0] & 0xffff, nbytes / 1e6))

    def _admit(self, req: Request, now: float) -> bool:
        if req._sol is None:
            req._sol = self.be.load(req.state_path)
        sol = req._sol
        fp = self.be.key(sol)
        key = self.be.batch_key(sol)
        mesh_b, wave_b = self.be.footprint(sol)
        resident = fp in self._meshes
        if len(self._active.get(key, ())) >= self.max_members:
            return False
        need = wave_b + (0 if resident else mesh_b)
        if self.used + need > self.budget:
            self._evict_idle(need)
            resident = fp in self._meshes
            need = wave_b + (0 if resident else mesh_b)
            if self.used + need > self.budget:
                return False
        req._sol = None
        if resident:
            self.be.mesh_bind(sol, self._meshes[fp][0])
            self._meshes.move_to_end(fp)
        else:
            h = self.be.mesh_get(sol)
            self._meshes[fp] = [h, mesh_b, 0]
        self._meshes[fp][2] += 1
        self.used += need
        self.peak_used = max(self.peak_used, self.used)
        req.key, req.fp, req.admitted_s = key, fp, now
        req.niter = sol.dims["niter"] if req.steps is None else min(req.steps, sol.dims["niter"])
        req.gll_per_step = 25 * (sol.dims["nel_solid"] + sol.dims["nel_fluid"])
        self._active[key].append(_Active(req, sol, wave_b))
        self.log("admitted request %d at %.3f s (key %x, %d on it)"
                 % (req.rid, now, key[0] & 0xffff, len(self._active[key])))
        return True

    # -- stepping --------------------------------------------------------------
    def _batch_for(self, key):
        members = self._active[key]
        ids = [a.req.rid for a in members]
        cur = self._batches.get(key)
        if cur is not None and cur[1] == ids:
            return cur[0]
        if cur is not None:
            cur[0].close()          # state moves back to the members
        b = self.be.batch([a.sol for a in members], self.group_width)
        self._batches[key] = (b, ids)
        self.max_cobatched = max(self.max_cobatched, len(ids))
        return b

    def _step_key(self, key):
        members = self._active[key]
        nsteps = min([self.chunk] + [a.req.niter - a.sol.iter for a in members])
        b = self._batch_for(key)
        out = b.step(nsteps)
        for a, s in zip(members, out):
            a.chunks.append(s)

    def _retire(self, now: float) -> None:
        for key in list(self._active):
            done = [a for a in self._active[key] if a.sol.iter >= a.req.niter]
            if not done:
                continue
            cur = self._batches.pop(key, None)
            if cur is not None:
                cur[0].close()
            for a in done:
                a.req.seis = (np.concatenate(a.chunks) if a.chunks else
                              np.zeros((0, max(a.sol.dims["num_rec"], 1), 3), np.float32))
                a.req.done_s = now
                a.sol.close()
                self.used -= a.wave_bytes
                self._meshes[a.req.fp][2] -= 1
                self.log("request %d done at %.3f s" % (a.req.rid, now))
            self._active[key] = [a for a in self._active[key] if a not in done]
            if not self._active[key]:
                del self._active[key]

    def run(self) -> dict:
        """Serve every submitted request; returns summary statistics."""
        order = sorted(self.requests, key=lambda r: (r.arrival_s, r.rid))
        pending = collections.deque(order)
        preparing: list = []
        waiting: collections.deque = collections.deque()
        t0 = self.clock()
        lock = threading.Lock()
        with ThreadPoolExecutor(max_workers=8) as pool, \
                ThreadPoolExecutor(max_workers=8) as prep:
            while pending or preparing or waiting or self._active:
                now = self.clock() - t0
                while pending and pending[0].arrival_s <= now:
                    req = pending.popleft()
                    if req.state_path is None:
                        preparing.append((req, prep.submit(self.preparer.prepare, req.mesh_dir,
                                                           req.cfg)))
                    else:
                        req.ready_s = req.arrival_s
                        waiting.append(req)
                for item in list(preparing):
                    req, fut = item
                    if not fut.done():
                        continue
                    preparing.remove(item)
                    try:
                        req.state_path, req.steps = fut.result()
                        req.ready_s = self.clock() - t0
                        waiting.append(req)
                        self.log("request %d prepared in %.2f s" % (req.rid, req.setup_s))
                    except Exception as e:      # reported per request
                        req.error = "preparation failed: %s" % e
                # admission in arrival order; a request that does not fit waits
                # without blocking smaller ones behind it
                for req in list(waiting):
                    if self._admit(req, now):
                        waiting.remove(req)
                if not self._active:
                    if waiting and not pending and not preparing:
                        req = waiting.popleft()
                        req.error = "does not fit the device memory budget"
                        if req._sol is not None:
                            req._sol.close()
                            req._sol = None
                        continue
                    if preparing:
                        time.sleep(0.005)
                    elif pending:
                        time.sleep(max(0.0, pending[0].arrival_s - (self.clock() - t0)))
                    continue
                keys = list(self._active)
                futs = [pool.submit(self._step_key, k) for k in keys]
                for f in futs:
                    f.result()
                with lock:
                    self.chunks_run += 1
                self._retire(self.clock() - t0)
        for b, _ in self._batches.values():
            b.close()
        self._batches.clear()
        for h, nbytes, _ in self._meshes.values():
            self.be.mesh_release(h)
            self.used -= nbytes
        self._meshes.clear()
        return self.summary(self.clock() - t0)

    def summary(self, makespan: float) -> dict:
        done = [r for r in self.requests if r.done_s is not None]
        lat = np.array([r.latency_s for r in done]) if done else np.zeros(0)
        upd = sum(r.gll_per_step * r.niter for r in done)
        setup = np.array([r.setup_s for r in done if r.setup_s is not None])
        return {
            "requests": len(self.requests), "completed": len(done),
            "setup_p50_s": float(np.percentile(setup, 50)) if setup.size else None,
            "setup_max_s": float(setup.max()) if setup.size else None,
            "rejected": sum(1 for r in self.requests if r.error),
            "makespan_s": makespan,
            "time_to_seismogram_p50_s": float(np.percentile(lat, 50)) if done else None,
            "time_to_seismogram_p99_s": float(np.percentile(lat, 99)) if done else None,
            "gll_point_updates_per_s": upd / makespan if makespan > 0 else None,
            "peak_device_bytes": self.peak_used, "max_cobatched": self.max_cobatched,
            "chunks": self.chunks_run,
        }
