"""Benchmark of the GPU time loop on its headline workload.

Workload (one MI355X, whole mesh resident): prem_iso, DOMINANT_PERIOD 10 s,
NPOL 4, anelastic with 5 standard linear solids (coarse-grained memory
variables), symplectic 4th-order PEFRL, dipole source (horizontal force) at
50 km, Gaussian STF, FP32, 3600 s seismograms at 181 receivers at 1 degree
spacing.  The mesh is the reference MESHER's; the precomputed terms are those
of the reference's own setup, recorded by the drop-in host.

A "step" is one full time step (4 PEFRL stages) of the whole mesh; the metric
is element-local GLL-point updates per second, 25 (nel_solid + nel_fluid) per
step.  ``value`` is the plain path (one wavefield).  The batched path steps 4
wavefields on the same mesh (the workload's source at 50, 150, 300 and 600 km)
in one launch per kernel and is reported beside it, per wavefield and in
total, with the step latency of both.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import resource
import shutil
import subprocess
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from axisem_amd import inputs  # noqa: E402
from axisem_amd.state import read_dims, read_state  # noqa: E402

HOST = os.path.join(REPO, "axisem_amd", "fortran", "_build", "axisem_gpu")
MESH = os.path.join(REPO, "data", "meshes", "prem_iso_10s")
HBM_PEAK_GBS = 8000.0

WORKLOAD = inputs.RunConfig(
    source_type="thetaforce", source_depth_km=50.0, source_amplitude=1e20,
    time_scheme="symplec4", attenuation=True, seismogram_length_s=3600.0,
    stations_deg=[float(k) for k in range(181)])
WORKLOAD_NAME = ("prem_iso 10 s, NPOL 4, anelastic 5 SLS (cg4), symplec4, dipole "
                 "thetaforce 50 km, 3600 s, 181 receivers")
BATCH_DEPTHS_KM = (50.0, 150.0, 300.0, 600.0)   # member 0 is the plain workload


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def _stack():
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def member_config(depth_km: float) -> inputs.RunConfig:
    return inputs.RunConfig(**{**WORKLOAD.__dict__, "source_depth_km": depth_km})


def prepare_state(workdir: str, cfg: inputs.RunConfig = WORKLOAD, mesh: str = MESH) -> str:
    """Record (once) the uploads of the drop-in host for one source on a
    mesh."""
    if not os.path.exists(os.path.join(mesh, "meshdb.dat0000")):
        raise SystemExit("missing %s: run __graft_entry__.build() where the reference "
                         "tree exists" % mesh)
    if not os.path.exists(HOST):
        raise SystemExit("missing drop-in host %s" % HOST)
    os.makedirs(workdir, exist_ok=True)
    rundir = os.path.join(workdir, "run_gpu")
    state = os.path.join(workdir, "state.bin")
    stamp = os.path.join(workdir, "stamp.json")
    key = json.dumps({"cfg": repr(cfg), "mesh": mesh,
                      "meshsize": os.path.getsize(os.path.join(mesh, "meshdb.dat0000")),
                      "host": os.path.getmtime(HOST)})
    if os.path.exists(state) and os.path.exists(stamp) and open(stamp).read() == key:
        return state
    inputs.write_run_dir(rundir, mesh, cfg)
    t = time.time()
    inputs.record_state(HOST, rundir, state)
    log("recorded the state of the %.0f km source in %.1fs" % (cfg.source_depth_km,
                                                                 time.time() - t))
    with open(stamp, "w") as f:
        f.write(key)
    return state


def prepare_states(workdir: str, nbatch: int) -> list:
    """States of the plain workload (in workdir) and of the other batch
    members (workdir/member<k>), recorded in parallel."""
    jobs = [(workdir, WORKLOAD)] + [
        (os.path.join(workdir, "member%d" % k), member_config(BATCH_DEPTHS_KM[k]))
        for k in range(1, nbatch)]
    with ThreadPoolExecutor(max_workers=4) as ex:
        return list(ex.map(lambda j: prepare_state(*j), jobs))


def direct_points(st) -> int:
    """Solid points the single-wavefield stiffness kernel kicks itself: one
    element copy, not on the axis, not on the solid-fluid boundary, no
    source term (the library's kFlagDirect)."""
    import numpy as np
    d = st.dims
    ig = np.asarray(st["igloc_solid"], np.int64) - 1
    nglob = int(ig.max()) + 1
    single = np.bincount(ig, minlength=nglob) == 1
    off = np.zeros(nglob, bool)
    ax = np.asarray(st["axis_solid"]) != 0
    k = np.arange(25)
    off[ig.reshape(-1, 25)[ax][:, k % 5 == 0].ravel()] = True
    for b in range(d["nel_bdry"]):
        e, j = int(st["bdry_solid_el"][b]) - 1, int(st["bdry_jpol_solid"][b])
        off[ig[e * 25 + 5 * j: e * 25 + 5 * j + 5]] = True
    if not d["fluid_src"]:
        term = np.asarray(st["source_term"], np.float32).reshape(3, 8, 25)
        for kk, e in enumerate(np.asarray(st["ielsrc"])[: d["nelsrc"]]):
            nz = np.any(term[:, kk, :] != 0, axis=0)
            off[ig[(int(e) - 1) * 25 + np.nonzero(nz)[0]]] = True
    return int(np.count_nonzero(single & ~off))


# serving stream: meshes, source kinds, depths and durations of the requests
SERVE_MESHES = {"prem_iso_50s": 50.0, "prem_iso_20s": 20.0, "prem_iso_10s": 10.0}
SERVE_KINDS = ("vertforce", "thetaforce", "explosion", "mt")
MT_RUNS = ("mrr", "mtt_p_mpp", "mtr", "mtp")


def serving_stream(nsolves: int, rate_per_s: float, seed: int = 0) -> list:
    """Solves of the serving workload with Poisson arrivals (seed 0): mesh
    uniform over SERVE_MESHES, source uniform over SERVE_KINDS, depth U(0,700)
    km, duration U(600,3600) s, elastic FP32 newmark2, 181 receivers.  A
    moment tensor with entries |N(0,1)| x 1e20 N m becomes its 4 runs (one
    request each, same arrival).  Returns (arrival_s, mesh name, RunConfig)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    for _ in range(nsolves):
        t += rng.exponential(1.0 / rate_per_s)
        mesh = list(SERVE_MESHES)[rng.integers(len(SERVE_MESHES))]
        kind = SERVE_KINDS[rng.integers(len(SERVE_KINDS))]
        depth = float(rng.uniform(0.0, 700.0))
        dur = float(rng.uniform(600.0, 3600.0))
        runs = ([(r, float(abs(rng.standard_normal())) * 1e20) for r in MT_RUNS]
                if kind == "mt" else [(kind, 1e20)])
        for src, amp in runs:
            out.append((t, mesh, inputs.RunConfig(
                source_type=src, source_depth_km=round(depth, 1), source_amplitude=amp,
                time_scheme="newmark2", attenuation=False, seismogram_length_s=round(dur, 1),
                stations_deg=[float(k) for k in range(181)])))
    return out


def serving_leg(workdir: str, nsolves: int, rate: float, chunk: int) -> dict:
    """The request loop on the serving stream.  Requests arrive as run
    configurations and are prepared by the loop on arrival (one drop-in host
    kept running per mesh and source type), so time to seismogram includes
    the setup; the first request of a mesh and source type pays the full
    setup, later ones the source routines only."""
    from axisem_amd.serve import Preparer, RequestLoop
    stream = serving_stream(nsolves, rate)
    if os.path.isdir(workdir):
        shutil.rmtree(workdir)
    prep = Preparer(HOST, os.path.join(workdir, "prep"))
    loop = RequestLoop(hbm_budget_bytes=200e9, chunk_steps=chunk, preparer=prep)
    try:
        for arr, m, cfg in stream:
            loop.submit_config(os.path.join(REPO, "data", "meshes", m), cfg, arrival_s=arr)
        st = loop.run()
    finally:
        prep.close()
    st["solves"] = nsolves
    st["arrival_rate_per_s"] = rate
    st["chunk_steps"] = chunk
    st["mix"] = ["%s %s %.0fkm %.0fs" % (m, c.source_type, c.source_depth_km, c.seismogram_length_s)
                 for _, m, c in stream]
    st["setup_s"] = [round(r.setup_s, 2) if r.setup_s is not None else None
                     for r in loop.requests]
    return st


# the stiffness terms the library packs per element point, per source type
PACKED_TERMS = {
    0: ("M_1", "M_2", "M_3", "M_4", "M_w1", "M11s", "M21s", "M41s", "M12s", "M22s", "M32s",
        "M42s", "M11z", "M21z", "M41z"),
    1: ("M_1", "M_2", "M_3", "M_4", "M_5", "M_6", "M_7", "M_8", "M_w1", "M_w2", "M_w3",
        "M11s", "M21s", "M41s", "M12s", "M22s", "M42s", "M13s", "M32s", "M33s", "M43s",
        "M11z", "M21z", "M41z"),
    2: ("M_1", "M_2", "M_3", "M_4", "M_5", "M_6", "M_7", "M_8", "M_w1", "M_w2", "M_w3",
        "M_w4", "M_w5", "M11s", "M21s", "M41s", "M12s", "M22s", "M32s", "M42s", "M11z",
        "M21z", "M41z", "M1phi", "M2phi", "M4phi"),
}
ZERO_TERM = {1: "M_w2", 2: "M_w3"}


def read_terms(st, nwave: int = 1) -> int:
    """Packed stiffness terms the stiffness kernel reads: with one wavefield,
    the term an isotropic model leaves +0 at every point (M_w2 of a dipole,
    M_w3 of a quadrupole) is not read; a group reads them all."""
    import numpy as np
    src = st.dims["src_type"]
    n = len(PACKED_TERMS[src])
    z = ZERO_TERM.get(src)
    if nwave == 1 and z and not np.asarray(st[z]).view(np.uint32).any():
        n -= 1
    return n


def solid_kernel_bytes(st, nwave: int = 1, ndirect: int = 0) -> int:
    """Algorithmic HBM bytes of one launch of the stiffness kernel (solid and
    fluid tiles) in a symplectic stage advancing nwave wavefields: every
    precomputed term the kernel reads (read_terms) at
    every element-local point once, the element->global index, the axial
    index, the axial terms and the stage cg4 terms once (shared by the
    wavefields); per wavefield the displacement of every global point once,
    the element-local force written once and the SLS sums.  A single
    wavefield (nwave 1) kicks its ndirect single-copy points itself: no force
    for them, but their velocity read and written, their displacement
    written and their inverse mass read (they are one index range; only the
    few tiles holding a source point read the point flags)."""
    d = st.dims
    nel, nglob = d["nel_solid"], d["nglob_solid"]
    ncomp = 2 if d["src_type"] == 0 else 3
    nterms = read_terms(st, nwave)
    naxel = int(st["axis_solid"].sum())
    mesh = 4 * 25 * nel * nterms       # stiffness terms
    mesh += 4 * 25 * nel               # element -> global index
    mesh += 4 * nel                    # axial flag / index
    mesh += 4 * 50 * naxel             # M0_w terms of axial elements
    wave = 4 * ncomp * nglob           # displacement gather
    wave += 4 * ncomp * 25 * nel       # force
    if d["anel"]:
        mesh += nel * 5 * 4 * 4        # Y, V_{s,z}_{eta,xi} at the cg points
        wave += nel * 24 * 4           # SLS sums
    if nwave == 1 and ndirect:
        wave -= 4 * ncomp * ndirect                    # no force for them
        wave += ndirect * (4 + 3 * 4 * ncomp)          # minv, velo r/w, disp w
    # the fluid tiles of the same launch: terms (3, multipoles 4), index,
    # axial index once; chi gather and force per wavefield
    nelf, nglobf = d["nel_fluid"], d["nglob_fluid"]
    if nelf:
        mesh += 4 * 25 * nelf * (3 if d["src_type"] == 0 else 4) + 4 * 25 * nelf + 4 * nelf
        wave += 4 * nglobf + 4 * 25 * nelf
    return mesh + nwave * wave


def kernel_table(stats: dict) -> dict:
    return {k: {"avg_us": round(v[0] / v[1] * 1e3, 2), "launches": v[1]}
            for k, v in stats.items() if v[1]}


def step_latency(runner, nsteps: int) -> dict:
    """Wall time of single-step calls (launch, run and synchronize)."""
    import numpy as np
    t = []
    for _ in range(nsteps):
        t0 = time.perf_counter()
        runner.run(1)
        t.append((time.perf_counter() - t0) * 1e3)
    return {"p50_ms": float(np.percentile(t, 50)), "p99_ms": float(np.percentile(t, 99)),
            "samples": nsteps}


def _reference_run(exe: str, steps: int, dt: float):
    """One run of the reference program on the workload's mesh and inputs,
    shortened to `steps` time steps; returns (loop seconds by its own 'Time
    loop routine' clock, steps run) or None."""
    with tempfile.TemporaryDirectory() as tmp:
        cfg = inputs.RunConfig(**{**WORKLOAD.__dict__})
        cfg.seismogram_length_s = steps * dt
        inputs.write_run_dir(tmp, MESH, cfg)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([exe], cwd=tmp, capture_output=True, text=True, preexec_fn=_stack,
                           timeout=1800, env=env)
        if r.returncode != 0:
            log("reference run failed: " + r.stdout[-500:])
            return None
        m = re.search(r"Time loop routine\s+\d+\s+([\d.]+)\s+([\d.]+)", r.stdout)
        n = re.search(r"# iterations\s*:\s*(\d+)", r.stdout)
        if not (m and n):
            return None
        return float(m.group(2)), int(n.group(1))


def cpu_baseline(seconds: float, dt: float, cores: int):
    """The reference program (oracle/_ref/axisem, serial, realkind 4) on a
    bounded sample of the same workload (the same mesh and inputs with the
    seismogram shortened to a few steps), timed by the reference's own 'Time
    loop routine' clock: once alone on one core, then as `cores` independent
    serial runs at the same time (the throughput of the host's cores, the way
    a multi-source study runs the reference: one process per source)."""
    exe = os.path.join(REPO, "oracle", "_ref", "axisem")
    if not os.path.exists(exe):
        return None
    # the reference needs at least 100 steps (runtime_info takes
    # mod(iter, floor(niter/100.)), SOLVER/time_evol_wave.F90:1017-1024)
    steps = max(100, int(seconds / 0.35))
    ne = re.search(r"nelem\s*=\s*(\d+)", open(os.path.join(MESH, "mesh_params.h")).read())
    if not ne:
        return None
    nelem = int(ne.group(1))
    one = _reference_run(exe, steps, dt)
    if one is None:
        return None
    t_loop, niter = one
    out = {"value": 25.0 * nelem * niter / t_loop, "unit": "GLL-point updates/s",
           "cores": 1, "kind": "reference",
           "sample": "reference SOLVER (amdflang -O2, serial, FP32) on the same mesh and "
                     "inputs, %d time steps, 'Time loop routine' clock %.1f s" % (niter, t_loop)}
    if cores > 1:
        with ThreadPoolExecutor(max_workers=cores) as ex:
            runs = list(ex.map(lambda _: _reference_run(exe, steps, dt), range(cores)))
        if all(r is not None for r in runs):
            agg = sum(25.0 * nelem * n / t for t, n in runs)
            out = {"value": agg, "unit": "GLL-point updates/s", "cores": cores,
                   "kind": "reference",
                   "sample": "%d independent serial runs of the reference SOLVER (amdflang -O2, "
                             "FP32) at once, each on the same mesh and inputs for %d time "
                             "steps; sum of their updates/s by each run's 'Time loop routine' "
                             "clock (%.1f-%.1f s)" % (cores, runs[0][1], min(t for t, _ in runs),
                                                      max(t for t, _ in runs)),
                   "one_core": out}
    return out


def copy_calibration(nbytes: int, reps: int = 5) -> None:
    """Device-to-device copies of nbytes (the HIP runtime's plain copy
    kernel): the reference point for scaling the FETCH_SIZE / WRITE_SIZE
    counters of a profiled run to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}))


# The plain CPU-runnable configuration on the GPU: prem_iso 50 s, elastic,
# newmark2, vertical surface force, 1800 s at 61 receivers 3 degrees apart.
P50_MESH = os.path.join(REPO, "data", "meshes", "prem_iso_50s")
P50_WORKLOAD = inputs.RunConfig(
    source_type="vertforce", source_depth_km=0.0, source_amplitude=1e20,
    time_scheme="newmark2", attenuation=False, seismogram_length_s=1800.0,
    stations_deg=[3.0 * k for k in range(61)])


def p50_leg(workdir: str, steps: int, warmup: int) -> dict:
    """Step time of the 50 s configuration (6 050 elements: a few tiles per
    CU, launch-latency bound), every kernel launched directly and with the
    autotuned replayed graph of steps."""
    from axisem_amd.solver import Solver
    from axisem_amd.tune import tune_solo
    path = prepare_state(workdir, P50_WORKLOAD, P50_MESH)
    st = read_dims(path)
    gll = 25 * (st["nel_solid"] + st["nel_fluid"])
    tuned = tune_solo(path)
    out = {"workload": "prem_iso 50 s, NPOL 4, elastic, newmark2, monopole vertforce 0 km, "
                       "1800 s, 61 receivers",
           "nel": st["nel_solid"] + st["nel_fluid"], "steps": steps,
           "tuned_graph_steps": tuned["graph_steps"],
           "tuner_device_ms_per_step": tuned["ms_per_step"]}
    for name, gs in (("direct", 0), ("graph", tuned["graph_steps"])):
        with Solver.from_state_file(path) as s:
            s.set_graph_steps(gs)
            s.run(warmup)
            t0 = time.perf_counter()
            s.run(steps)
            wall = time.perf_counter() - t0
            dev = s.last_run_ms()
        out[name] = {"graph_steps": gs, "ms_per_step": wall / steps * 1e3,
                     "device_ms_per_step": dev / steps, "value": gll * steps / wall,
                     "unit": "GLL-point updates/s"}
    return out


def kwf_leg(workdir: str, steps: int, warmup: int) -> dict:
    """Kernel-wavefield dumps (displ_only) on the plain workload: the step time
    with and without the snapshots (gathered on the GPU every strain_it steps
    and copied to host memory), and the gather kernel's bytes and time.
    strain_it follows the reference's rule for a 10 s source period with the
    default 8 samples per period, floor(period / 8 / deltat)
    (SOLVER/parameters.F90:1016-1022)."""
    import math
    from axisem_amd.solver import Solver
    from axisem_amd.solver import Batch

    def record(sub, cfg):
        wd = os.path.join(workdir, sub)
        os.makedirs(wd, exist_ok=True)
        rundir, state = os.path.join(wd, "run_gpu"), os.path.join(wd, "state.bin")
        stamp = os.path.join(wd, "stamp.json")
        key = json.dumps({"cfg": repr(cfg), "host": os.path.getmtime(HOST), "kwf": 1})
        if not (os.path.exists(state) and os.path.exists(stamp) and open(stamp).read() == key):
            inputs.write_run_dir(rundir, MESH, cfg)
            inputs.record_state(HOST, rundir, state, kwf_strain_it=1)
            with open(stamp, "w") as f:
                f.write(key)
        return state
    t = time.time()
    with ThreadPoolExecutor(max_workers=4) as ex:
        states = list(ex.map(lambda k: record("d%g" % BATCH_DEPTHS_KM[k], member_config(
            BATCH_DEPTHS_KM[k])), range(len(BATCH_DEPTHS_KM))))
    log("kernel-wavefield states in %.1fs" % (time.time() - t))
    state = states[0]
    st = read_state(state)
    d = st.dims
    gll = 25 * (d["nel_solid"] + d["nel_fluid"])
    strain_it = max(1, math.floor(10.0 / 8 / d["deltat"]))
    nps, npf = (int(x) for x in st["kwf_npoints"])
    from axisem_amd.tune import GRAPH_CANDIDATES, tune_solo
    # graphs no longer than the warm-up, which captures them
    gs = tune_solo(state, candidates=[c for c in GRAPH_CANDIDATES if c <= warmup])["graph_steps"]
    out = {"workload": WORKLOAD_NAME, "strain_it": strain_it, "graph_steps": gs,
           "npoint_solid": nps, "npoint_fluid": npf, "steps": steps}
    with Solver.from_state_file(state) as s:
        s.set_graph_steps(gs)
        s.run(warmup)
        t0 = time.perf_counter()
        s.run(steps)
        plain = time.perf_counter() - t0
    with Solver.from_state_file(state) as s:
        s.set_graph_steps(gs)
        s.set_kwf(strain_it)
        s.run_kwf(warmup)
        t0 = time.perf_counter()
        _, kwf = s.run_kwf(steps)
        wall = time.perf_counter() - t0
        s.set_profiling(True)
        s.run_kwf(strain_it * 8)
        ks = s.kernel_stats()
    kms, kn = ks["receivers"]      # the gather's launches (the receivers ride in k_memvar)
    snap_bytes = kwf[0].nbytes
    # algorithmic bytes of one gather: solid point 4 B index + 12 B displacement +
    # 12 B written; fluid point 40 B row/column ids + 4 B (i, j) + 20 B terms + 12 B
    # written; chi once
    nbytes = nps * 28 + npf * 76 + 4 * d["nglob_fluid"]
    avg_s = kms / max(kn, 1) * 1e-3
    out.update({
        "snapshots": int(kwf.shape[0]), "snapshot_bytes": snap_bytes,
        "ms_per_step_plain": plain / steps * 1e3, "ms_per_step_kwf": wall / steps * 1e3,
        "overhead": wall / plain - 1.0, "value_kwf": gll * steps / wall,
        "kernel": {"name": "k_kwf", "launches": kn, "avg_us": avg_s * 1e6,
                   "algorithmic_bytes": nbytes, "achieved_GBps": nbytes / avg_s / 1e9,
                   "frac": nbytes / avg_s / 1e9 / HBM_PEAK_GBS},
        "snapshot_copy_GBps": snap_bytes * kwf.shape[0] / max(wall - plain, 1e-9) / 1e9})
    # the same workload at its four depths as one batch, with and without dumps
    from axisem_amd.tune import tune_batch
    bt = tune_batch(states, candidates=[c for c in GRAPH_CANDIDATES if c <= warmup])
    times = {}
    for dumps in (False, True):
        members = [Solver.from_state_file(p) for p in states]
        try:
            for m in members:
                m.set_kwf(strain_it)
            with Batch(members, group_width=bt["group_width"], graph_steps=bt["graph_steps"]) as b:
                (b.step_kwf if dumps else b.step)(warmup)
                t0 = time.perf_counter()
                r = (b.step_kwf if dumps else b.step)(steps)
                times[dumps] = time.perf_counter() - t0
                if dumps:
                    nsnap4 = sum(k.shape[0] for k in r[1])
        finally:
            for m in members:
                m.close()
    out["batch4"] = {"schedule": [bt["group_width"], bt["graph_steps"]],
                     "ms_per_step": times[False] / steps * 1e3,
                     "ms_per_step_kwf": times[True] / steps * 1e3,
                     "overhead": times[True] / times[False] - 1.0, "snapshots": nsnap4,
                     "value_kwf": 4 * gll * steps / times[True]}
    return out


# Batched wavefields on one mesh: prem_iso 5 s, elastic, newmark2 (the time
# scheme is not stated for this configuration; newmark2 is the reference's
# default), the four moment-tensor runs at 100 km, and the same four at 10,
# 100, 300 and 600 km (16 wavefields); 1800 s at 181 receivers.
MT_MESH = os.path.join(REPO, "data", "meshes", "prem_iso_5s")
MT_DEPTHS_KM = (10.0, 100.0, 300.0, 600.0)


def mt_config(run: str, depth_km: float) -> inputs.RunConfig:
    return inputs.RunConfig(source_type=run, source_depth_km=depth_km, source_amplitude=1e20,
                            time_scheme="newmark2", attenuation=False, seismogram_length_s=1800.0,
                            stations_deg=[float(k) for k in range(181)])


def prepare_mt_states(workdir: str) -> dict:
    """States of the 16 moment-tensor wavefields, (run, depth) -> path: one
    record run of the drop-in host per source type (monopole: Mrr and
    Mtt+Mpp; dipole: Mtr; quadrupole: Mtp), the three in parallel."""
    if not os.path.exists(os.path.join(MT_MESH, "meshdb.dat0000")):
        raise SystemExit("missing %s: run __graft_entry__.build() where the reference "
                         "tree exists" % MT_MESH)
    kinds = {"monopole": ("mrr", "mtt_p_mpp"), "dipole": ("mtr",), "quadpole": ("mtp",)}
    jobs = []
    for kind, runs in kinds.items():
        keys = [(r, d) for r in runs for d in MT_DEPTHS_KM]
        cfgs = [mt_config(*k) for k in keys]
        paths = [os.path.join(workdir, "%s_%g.bin" % k) for k in keys]
        jobs.append((kind, keys, cfgs, paths))
    os.makedirs(workdir, exist_ok=True)
    stamp = os.path.join(workdir, "stamp.json")
    key = json.dumps({"jobs": [repr(j[2]) for j in jobs],
                      "meshsize": os.path.getsize(os.path.join(MT_MESH, "meshdb.dat0000")),
                      "host": os.path.getmtime(HOST)})
    out = {k: p for _, keys, _, paths in jobs for k, p in zip(keys, paths)}
    if (os.path.exists(stamp) and open(stamp).read() == key
            and all(os.path.exists(p) for p in out.values())):
        return out
    t = time.time()

    def rec(job):
        kind, _, cfgs, paths = job
        inputs.record_states(HOST, os.path.join(workdir, "run_" + kind), MT_MESH, cfgs, paths)
    with ThreadPoolExecutor(max_workers=3) as ex:
        list(ex.map(rec, jobs))
    log("recorded the 16 moment-tensor states in %.1fs" % (time.time() - t))
    with open(stamp, "w") as f:
        f.write(key)
    return out


def mt_leg(workdir: str, steps: int, warmup: int) -> dict:
    """Config of batched wavefields: each moment-tensor run alone, the four
    at 100 km as one batch of mixed source types, and all 16 as one batch;
    vs_plain = per-wavefield gain = sum of the members' plain step times over
    the batch's step time."""
    from axisem_amd.solver import Batch, Solver
    paths = prepare_mt_states(workdir)
    d = read_dims(paths[("mrr", 100.0)])
    gll = 25 * (d["nel_solid"] + d["nel_fluid"])

    # graphs of 4 steps (captured in the warm-up); the tuner's run over 16
    # states of this mesh would cost more than the leg itself
    gs = 4 if warmup >= 4 else 0

    def timed(runner):
        runner.set_graph_steps(gs)
        runner.run(warmup)
        t0 = time.perf_counter()
        runner.run(steps)
        return (time.perf_counter() - t0) / steps * 1e3, runner.last_run_ms() / steps

    plain = {}
    for run in MT_RUNS:
        with Solver.from_state_file(paths[(run, 100.0)]) as s:
            plain[run] = timed(s)[0]
    log("moment tensor, plain ms per step: " + json.dumps(plain))
    out = {"workload": "prem_iso 5 s, NPOL 4, elastic, newmark2, moment-tensor runs mrr, "
                       "mtt_p_mpp, mtr, mtp at 100 km (x4: at 10/100/300/600 km), 1800 s, "
                       "181 receivers",
           "nel_solid": d["nel_solid"], "nel_fluid": d["nel_fluid"],
           "niter_full_run": d["niter"], "deltat": d["deltat"], "steps": steps,
           "graph_steps": gs, "plain_ms_per_step": plain}
    for name, keys in (("batch4", [(r, 100.0) for r in MT_RUNS]),
                       ("batch16", [(r, dd) for dd in MT_DEPTHS_KM for r in MT_RUNS])):
        members = [Solver.from_state_file(paths[k]) for k in keys]
        try:
            with Batch(members) as b:
                ms, dev = timed(b)
                mem = sum(m.device_bytes() for m in members)
        finally:
            for m in members:
                m.close()
        ref = sum(plain[r] for r, _ in keys)
        out[name] = {"wavefields": len(keys), "ms_per_step": ms, "device_ms_per_step": dev,
                     "value": len(keys) * gll / (ms * 1e-3), "unit": "GLL-point updates/s",
                     "vs_plain": ref / ms}
        log("moment tensor %s: %s" % (name, json.dumps(out[name])))
    return out


# The large single-GPU configuration: prem_ani (transversely isotropic
# PREM) at 2 s, anelastic 5 SLS, symplec4, the four moment-tensor runs at
# 100 km batched, 1800 s at 181 receivers.  Opt-in (--config5): the mesh
# (2.5 M elements) and the three setups (one per source type) take the host
# about ten minutes; --config5-prepare does only that host part (no GPU), so
# that it can run beside other work, and --config5-only then times the GPU
# part on the states it left.
C5_MESH = os.path.join(REPO, "data", "meshes", "prem_ani_2s")
C5_KINDS = {"monopole": ("mrr", "mtt_p_mpp"), "dipole": ("mtr",), "quadpole": ("mtp",)}
C5_RUNS = ("mrr", "mtt_p_mpp", "mtr", "mtp")


def config5_states(workdir: str) -> dict:
    """The mesh and the four runs' states (host only).  Returns the state
    paths and the host times; states a previous call completed are reused."""
    done = os.path.join(workdir, "states.json")
    if os.path.exists(done):
        with open(done) as f:
            info = json.load(f)
        if all(os.path.exists(p) for p in info["paths"].values()):
            return info
    t0 = time.time()
    if not os.path.exists(os.path.join(C5_MESH, "meshdb.dat0000")):
        inputs.run_mesher(os.path.join(REPO, "oracle", "_ref", "xmesh"), C5_MESH, 2.0,
                          "prem_ani")
    tmesh = time.time() - t0
    log("config 5: mesh in %.0f s" % tmesh)

    def cfg(run):
        return inputs.RunConfig(source_type=run, source_depth_km=100.0, source_amplitude=1e20,
                                time_scheme="symplec4", attenuation=True,
                                seismogram_length_s=1800.0,
                                stations_deg=[float(k) for k in range(181)])
    os.makedirs(workdir, exist_ok=True)
    paths = {r: os.path.join(workdir, r + ".bin") for rs in C5_KINDS.values() for r in rs}
    t1 = time.time()

    def rec(item):
        kind, runs = item
        t = time.time()
        inputs.record_states(HOST, os.path.join(workdir, "run_" + kind), C5_MESH,
                             [cfg(r) for r in runs], [paths[r] for r in runs])
        log("config 5: %s state(s) %s in %.0f s" % (kind, ",".join(runs), time.time() - t))
    with ThreadPoolExecutor(max_workers=len(C5_KINDS)) as ex:
        list(ex.map(rec, C5_KINDS.items()))
    tsetup = time.time() - t1
    log("config 5: states in %.0f s" % tsetup)
    info = {"paths": paths, "host_mesh_s": tmesh, "host_setup_s": tsetup,
            "state_bytes": {r: os.path.getsize(p) for r, p in paths.items()}}
    with open(done, "w") as f:
        json.dump(info, f)
    return info


def config5_leg(workdir: str, steps: int, warmup: int) -> dict:
    from axisem_amd.solver import Batch, Solver
    info = config5_states(workdir)
    paths = info["paths"]
    d = read_dims(paths["mtr"])
    gll = 25 * (d["nel_solid"] + d["nel_fluid"])
    out = {"workload": "prem_ani 2 s, NPOL 4, anelastic 5 SLS (cg4), symplec4, moment-tensor "
                       "runs mrr, mtt_p_mpp, mtr, mtp at 100 km, 1800 s, 181 receivers",
           "nel_solid": d["nel_solid"], "nel_fluid": d["nel_fluid"],
           "niter_full_run": d["niter"], "deltat": d["deltat"], "steps": steps,
           "host_mesh_s": info["host_mesh_s"], "host_setup_s": info["host_setup_s"],
           "state_bytes": info["state_bytes"]}
    import ctypes as C
    from axisem_amd import solver as S
    plain = {}
    for r in C5_RUNS:
        t = time.time()
        with Solver.from_state_file(paths[r]) as s:
            log("config 5: %s uploaded in %.0f s" % (r, time.time() - t))
            mb, wb = C.c_size_t(), C.c_size_t()
            S._check(S.library().axh_footprint(s._h, C.byref(mb), C.byref(wb)))
            out.setdefault("footprint_bytes", {})[r] = {"mesh": mb.value, "wavefield": wb.value}
            s.set_graph_steps(4)
            s.run(warmup)
            t = time.perf_counter()
            s.run(steps)
            plain[r] = (time.perf_counter() - t) / steps * 1e3
            out.setdefault("device_bytes", {})[r] = s.device_bytes()
        log("config 5: %s alone %.3f ms per step" % (r, plain[r]))
    out["plain_ms_per_step"] = plain
    members = [Solver.from_state_file(paths[r]) for r in C5_RUNS]
    try:
        with Batch(members, graph_steps=4) as b:
            b.run(warmup)
            t = time.perf_counter()
            b.run(steps)
            ms = (time.perf_counter() - t) / steps * 1e3
            out["batch4_device_bytes"] = sum(m.device_bytes() for m in members)
    finally:
        for m in members:
            m.close()
    out["batch4"] = {"ms_per_step": ms, "value": 4 * gll / (ms * 1e-3),
                     "unit": "GLL-point updates/s", "vs_plain": sum(plain.values()) / ms}
    out["full_run_estimate_s"] = d["niter"] * ms * 1e-3
    return out


# Other source types and time schemes on the 10 s mesh (not the headline
# workload): plain and 4-wavefield step times, stiffness launch per kernel.
# name, source, scheme, attenuation, coarse-grained memory variables
SWEEP = (("mono_vertforce_newmark", "vertforce", "newmark2", False, True),
         ("mono_vertforce_newmark_anel", "vertforce", "newmark2", True, True),
         ("mono_explosion_symplec4_anel", "explosion", "symplec4", True, True),
         ("dipole_thetaforce_newmark_anel", "thetaforce", "newmark2", True, True),
         ("quad_mtp_symplec4_anel", "mtp", "symplec4", True, True),
         # memory variables at all 25 points (COARSE_GRAINED false)
         ("dipole_thetaforce_symplec4_anelfull", "thetaforce", "symplec4", True, False))


def sweep(workdir: str, steps: int = 300, only=None) -> dict:
    from axisem_amd.solver import Batch, Solver
    out = {}
    for name, src, scheme, anel, cg in SWEEP:
        if only and name not in only:
            continue
        paths = []
        for depth in BATCH_DEPTHS_KM:
            cfg = inputs.RunConfig(source_type=src, source_depth_km=depth, source_amplitude=1e20,
                                   time_scheme=scheme, attenuation=anel, coarse_grained=cg,
                                   seismogram_length_s=1800.0,
                                   stations_deg=[float(k) for k in range(181)])
            paths.append(prepare_state(os.path.join(workdir, "%s_%g" % (name, depth)), cfg))
        from axisem_amd.tune import GRAPH_CANDIDATES, tune_batch, tune_solo
        cands = [c for c in GRAPH_CANDIDATES if c <= 50]
        tuned = tune_solo(paths[0], candidates=cands)
        btuned = tune_batch(paths, candidates=cands)
        s = Solver.from_state_file(paths[0])
        s.set_graph_steps(tuned["graph_steps"])
        s.run(50)
        t0 = time.perf_counter()
        s.run(steps)
        plain = (time.perf_counter() - t0) / steps * 1e3
        s.set_profiling(True)
        s.run(20)
        ks = kernel_table(s.kernel_stats())
        s.close()
        members = [Solver.from_state_file(p) for p in paths]
        with Batch(members, group_width=btuned["group_width"],
                   graph_steps=btuned["graph_steps"]) as b:
            b.run(50)
            t0 = time.perf_counter()
            b.run(steps)
            bat = (time.perf_counter() - t0) / steps * 1e3
            b.set_profiling(True)
            b.run(20)
            kb = kernel_table(b.kernel_stats())
        for m in members:
            m.close()
        out[name] = {"plain_ms_per_step": plain, "batch4_ms_per_step": bat,
                     "batch4_vs_plain": 4 * plain / bat,
                     "plain_graph_steps": tuned["graph_steps"],
                     "batch4_schedule": [btuned["group_width"], btuned["graph_steps"]],
                     "kernels_batch4": kb,
                     "stiffness_us": [ks["stiffness"]["avg_us"], kb["stiffness"]["avg_us"]],
                     "kernels_plain": ks}
        log("sweep %s: %s" % (name, json.dumps(out[name])))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--latency-steps", type=int, default=100)
    ap.add_argument("--batch", type=int, default=len(BATCH_DEPTHS_KM),
                    help="wavefields of the batched leg (1 = plain only)")
    ap.add_argument("--serve-solves", type=int, default=6,
                    help="solves of the serving leg (0: no serving leg)")
    ap.add_argument("--serve-rate", type=float, default=0.5, help="arrivals per second")
    ap.add_argument("--serve-chunk", type=int, default=200, help="steps per batch chunk")
    ap.add_argument("--group-width", type=int, default=None,
                    help="wavefields per kernel instance in the batched leg (default: tuned)")
    ap.add_argument("--graph-steps", type=int, default=None,
                    help="steps per replayed HIP graph (default: tuned, at most --warmup)")
    ap.add_argument("--no-p50", action="store_true", help="no 50 s configuration leg")
    ap.add_argument("--no-fused", action="store_true",
                    help="no legs with the fused multiply-add build")
    ap.add_argument("--p50-steps", type=int, default=1000, help="timed steps of the 50 s leg")
    ap.add_argument("--no-kwf", action="store_true",
                    help="no kernel-wavefield (displ_only) leg")
    ap.add_argument("--no-mt", action="store_true",
                    help="no moment-tensor batch leg (5 s mesh, 4 and 16 wavefields)")
    ap.add_argument("--config5", action="store_true",
                    help="also the large single-GPU configuration (prem_ani 2 s; ~10 min "
                         "of host setup)")
    ap.add_argument("--config5-prepare", action="store_true",
                    help="only the host part of the large configuration: mesh and states "
                         "(no GPU; one JSON line)")
    ap.add_argument("--config5-only", action="store_true",
                    help="only the large configuration (one JSON line)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-cores", type=int, default=min(16, os.cpu_count() or 1),
                    help="independent reference runs of the all-cores CPU leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workdir", default=os.path.join(REPO, "data", "bench_10s"))
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device copies of BYTES (counter calibration)")
    ap.add_argument("--sweep", action="store_true",
                    help="only time other source types and schemes (one JSON line)")
    ap.add_argument("--sweep-only", default="",
                    help="comma-separated sweep workloads (default: all)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.sweep:
        only = [x for x in a.sweep_only.split(",") if x]
        print(json.dumps({"sweep": sweep(os.path.join(REPO, "data", "bench_sweep"),
                                         only=only)}))
        return
    if a.config5_prepare:
        print(json.dumps({"config5_states": config5_states(os.path.join(REPO, "data",
                                                                        "bench_c5"))}))
        return
    if a.config5_only:
        print(json.dumps({"config5": config5_leg(os.path.join(REPO, "data", "bench_c5"),
                                                 a.steps, a.warmup)}))
        return
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the time loop keeps the whole mesh resident on "
              "one MI355X; --gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)

    from axisem_amd.solver import Batch, Solver

    nbatch = max(1, min(a.batch, len(BATCH_DEPTHS_KM)))
    states = prepare_states(a.workdir, nbatch if a.batch > 1 else 1)
    state = states[0]
    st = read_state(state)
    d = st.dims
    gll = 25 * (d["nel_solid"] + d["nel_fluid"])
    need = a.warmup + a.steps + a.profile_steps + a.latency_steps
    if need > d["niter"]:
        raise SystemExit("warmup + steps exceed the %d steps of the workload" % d["niter"])

    # ---- launch schedule: autotuned on throwaway contexts; graphs only as
    # long as the warm-up (which captures them) ----
    from axisem_amd.tune import GRAPH_CANDIDATES, tune_batch, tune_solo
    cands = [c for c in GRAPH_CANDIDATES if c <= a.warmup]
    tuned = tune_solo(state, candidates=cands) if a.graph_steps is None else \
        {"graph_steps": a.graph_steps, "ms_per_step": {}}
    log("plain: launch schedule %s" % json.dumps(tuned))

    # ---- plain path: one wavefield ----
    solver = Solver.from_state_file(state)
    solver.set_graph_steps(tuned["graph_steps"])
    log("plain: warmup %d steps" % a.warmup)
    solver.run(a.warmup)
    log("plain: timed %d steps" % a.steps)
    t0 = time.perf_counter()
    solver.run(a.steps)           # returns after the library stream is synchronized
    t1 = time.perf_counter()
    wall = t1 - t0
    dev_ms = solver.last_run_ms()
    value = gll * a.steps / wall
    latency = step_latency(solver, a.latency_steps)

    # per-kernel durations (HIP events around every launch, same stream)
    solver.set_profiling(True)
    solver.run(a.profile_steps)
    stats = solver.kernel_stats()
    solver.set_profiling(False)
    sk_ms, sk_n = stats["stiffness"]
    sk_avg_s = sk_ms / sk_n * 1e-3
    ndirect = direct_points(st)
    sk_bytes = solid_kernel_bytes(st, 1, ndirect)
    achieved = sk_bytes / sk_avg_s / 1e9
    kernels = kernel_table(stats)
    log("plain kernels: " + json.dumps(kernels))
    solver.close()

    # ---- batched path: nbatch wavefields on the same mesh ----
    batched = None
    if a.batch > 1:
        if a.group_width is None or a.graph_steps is None:
            btuned = tune_batch(states, candidates=cands)
        else:
            btuned = {"group_width": a.group_width, "graph_steps": a.graph_steps,
                      "ms_per_step": {}}
        log("batched: launch schedule %s" % json.dumps(btuned))
        members = [Solver.from_state_file(p) for p in states]
        with Batch(members, group_width=btuned["group_width"],
                   graph_steps=btuned["graph_steps"]) as b:
            log("batched x%d: warmup %d steps" % (nbatch, a.warmup))
            b.run(a.warmup)
            log("batched x%d: timed %d steps" % (nbatch, a.steps))
            t0 = time.perf_counter()
            b.run(a.steps)
            bwall = time.perf_counter() - t0
            blat = step_latency(b, a.latency_steps)
            b.set_profiling(True)
            b.run(a.profile_steps)
            bstats = b.kernel_stats()
            b.set_profiling(False)
        for m in members:
            m.close()
        bsk_ms, bsk_n = bstats["stiffness"]
        bsk_avg_s = bsk_ms / bsk_n * 1e-3
        bsk_bytes = solid_kernel_bytes(st, nbatch)
        btotal = nbatch * gll * a.steps / bwall
        batched = {
            "wavefields": nbatch,
            "group_width": btuned["group_width"],
            "graph_steps": btuned["graph_steps"],
            "tuner_device_ms_per_step": btuned["ms_per_step"],
            "sources": "dipole thetaforce at %s km" % "/".join("%g" % x for x in
                                                               BATCH_DEPTHS_KM[:nbatch]),
            # GLL-point updates of all member wavefields per second, against
            # the plain path's updates per second
            "value": btotal,
            "vs_plain": btotal / value,
            "ms_per_step": bwall / a.steps * 1e3,
            "latency": blat,
            "roofline": {"kernel": "stiffness, %d wavefields per launch" % nbatch,
                         "achieved": bsk_bytes / bsk_avg_s / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": bsk_bytes / bsk_avg_s / 1e9 / HBM_PEAK_GBS,
                         "algorithmic_bytes_per_launch": bsk_bytes,
                         "avg_launch_us": bsk_avg_s * 1e6},
            "kernels_us": kernel_table(bstats),
        }
        log("batched kernels: " + json.dumps(batched["kernels_us"]))

    # ---- the same legs with the build that fuses multiply-adds (within the
    # misfit budget, not bit for bit) ----
    fused = None
    if not a.no_fused and os.path.exists(os.path.join(REPO, "axisem_amd", "libaxihip_fma.so")):
        import numpy as np
        fused = {}
        with Solver.from_state_file(state, fused=True) as s:
            s.set_graph_steps(tuned["graph_steps"])
            s.run(a.warmup)
            t0 = time.perf_counter()
            s.run(a.steps)
            fwall = time.perf_counter() - t0
            s.set_profiling(True)
            s.run(a.profile_steps)
            fk = kernel_table(s.kernel_stats())
        fused["plain"] = {"value": gll * a.steps / fwall, "ms_per_step": fwall / a.steps * 1e3,
                          "kernels_us": fk}
        if batched is not None:
            members = [Solver.from_state_file(p, fused=True) for p in states]
            with Batch(members, group_width=btuned["group_width"],
                       graph_steps=btuned["graph_steps"]) as b:
                b.run(a.warmup)
                t0 = time.perf_counter()
                b.run(a.steps)
                fbwall = time.perf_counter() - t0
                b.set_profiling(True)
                b.run(a.profile_steps)
                fbk = kernel_table(b.kernel_stats())
            for m in members:
                m.close()
            fbsk = fbk["stiffness"]["avg_us"] * 1e-6
            fbtotal = nbatch * gll * a.steps / fbwall
            fused["batched"] = {
                "value": fbtotal, "vs_plain": fbtotal / value,
                "ms_per_step": fbwall / a.steps * 1e3,
                "roofline": {"kernel": "stiffness, %d wavefields per launch" % nbatch,
                             "achieved": bsk_bytes / fbsk / 1e9, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": bsk_bytes / fbsk / 1e9 / HBM_PEAK_GBS,
                             "avg_launch_us": fbsk * 1e6},
                "kernels_us": fbk}
        # misfit of the fields after 100 steps against the exact build
        disp = []
        for fl in (False, True):
            with Solver.from_state_file(state, fused=fl) as s:
                s.run(100)
                disp.append(s.fields()[0].astype(np.float64))
        fused["disp_rel_l2_vs_exact_100_steps"] = float(
            np.linalg.norm(disp[1] - disp[0]) / np.linalg.norm(disp[0]))
        log("fused build: " + json.dumps({k: v for k, v in fused.items()}))

    p50 = None
    if not a.no_p50:
        log("50 s configuration: %d steps" % a.p50_steps)
        p50 = p50_leg(os.path.join(REPO, "data", "bench_50s"), a.p50_steps, 50)
        log("50 s: " + json.dumps(p50))

    kwf = None
    if not a.no_kwf:
        log("kernel-wavefield dumps on the plain workload")
        # at least 40 snapshots, whatever --steps is: the copies overlap the
        # steps, which a run of a few snapshots cannot show
        kwf = kwf_leg(os.path.join(a.workdir, "kwf"), max(a.steps, 520), max(a.warmup, 20))
        log("kernel wavefields: " + json.dumps(kwf))

    mt = None
    if not a.no_mt:
        log("moment-tensor batches on the 5 s mesh")
        mt = mt_leg(os.path.join(REPO, "data", "bench_5s"), max(a.steps, 20), max(a.warmup, 5))

    c5 = None
    if a.config5:
        log("config 5: prem_ani 2 s")
        c5 = config5_leg(os.path.join(REPO, "data", "bench_c5"), max(a.steps, 10),
                         max(a.warmup, 5))
        log("config 5: " + json.dumps(c5))

    serving = None
    if a.serve_solves > 0:
        log("serving: %d solves, %.2f arrivals/s" % (a.serve_solves, a.serve_rate))
        serving = serving_leg(os.path.join(REPO, "data", "bench_serve"), a.serve_solves,
                              a.serve_rate, a.serve_chunk)
        log("serving: " + json.dumps({k: v for k, v in serving.items() if k != "mix"}))

    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline (reference program, ~%.0f s loop)" % a.cpu_seconds)
        cpu = cpu_baseline(a.cpu_seconds, d["deltat"], a.cpu_cores)

    traffic = traffic4 = None
    tfile = os.path.join(REPO, "profiles", "solid_stiffness_prem10s_traffic.json")
    if os.path.exists(tfile):
        try:
            tj = json.load(open(tfile))
            traffic = tj.get("bytes_per_launch")
            traffic4 = tj.get("batched_4", {}).get("bytes_per_launch")
        except ValueError:
            traffic = None
    if batched is not None and batched["wavefields"] == 4:
        batched["roofline"]["traffic"] = traffic4

    out = {
        "metric": "GLL-point updates/s per GPU, batched vs plain at \u22641e-4 misfit; p50/p99 latency",
        "value": value,
        "unit": "GLL-point updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "device_ms_per_step": dev_ms / a.steps,
        "graph_steps": tuned["graph_steps"],
        "tuner_device_ms_per_step": tuned["ms_per_step"],
        "time_steps_per_s": a.steps / wall,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "reference MESHER mesh and reference setup terms (prem_iso 10 s), deterministic",
        "config": {"workload": WORKLOAD_NAME, "nel_solid": d["nel_solid"],
                   "nel_fluid": d["nel_fluid"], "gll_points_per_step": gll,
                   "niter_full_run": d["niter"], "deltat": d["deltat"],
                   "receivers": len(WORKLOAD.stations_deg)},
        "roofline": {"bound": "hbm", "kernel": "stiffness (k_solid with the fluid tiles, symplectic stage)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": sk_bytes,
                     "avg_launch_us": sk_avg_s * 1e6},
        "latency": latency,
        "kernels_us": kernels,
        "batched": batched,
        "fused_multiply_add_build": fused,
        "plain_50s": p50,
        "kernel_wavefields": kwf,
        "moment_tensor_5s": mt,
        "prem_ani_2s": c5,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
