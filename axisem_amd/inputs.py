"""Run directories for the reference's own input files.

Writes ``inparam_basic``, ``inparam_advanced``, ``inparam_source`` and
``STATIONS`` in the keyword format the SOLVER parses
(SOLVER/parameters.F90:228-660, SOLVER/source.f90:40-140,
SOLVER/seismograms.f90:305-377) and ``inparam_mesh`` for the MESHER
(MESHER/input.f90:38).  The drop-in host and the reference read the same
directory, so both see exactly the same inputs.
"""
from __future__ import annotations

import os
import shutil
from dataclasses import dataclass, field


@dataclass
class RunConfig:
    source_type: str = "vertforce"      # mrr, explosion, vertforce, thetaforce, mtr, mtp, ...
    source_depth_km: float = 0.0
    source_amplitude: float = 1e20
    time_scheme: str = "newmark2"       # newmark2, symplec4, ML_SO4m5, ML_SO6m7
    attenuation: bool = False
    seismogram_length_s: float = 1800.0
    stations_deg: list = field(default_factory=lambda: [3.0 * k for k in range(61)])
    source_function: str = "gauss_0"
    n_sls: int = 5
    coarse_grained: bool = True
    snapshots: str = ""                 # "", "xdmf", "vtk" or "both" (SAVE_SNAPSHOTS)
    snapshot_dt: float = 20.0           # SNAPSHOT_DT, seconds
    energy: bool = False                # SAVE_ENERGY
    simulation_type: str = "single"     # SIMULATION_TYPE: single, or moment (one MT run)
    xdmf_gll_i: str = "3 0 2 4"         # XDMF_GLL_I: count, then the ipol indices
    xdmf_gll_j: str = "3 0 2 4"         # XDMF_GLL_J
    extra_advanced: str = ""            # more inparam_advanced lines (e.g. KERNEL_IBEG 0)


def write_mesh_inparam(path: str, period_s: float, model: str = "prem_iso",
                       coarsening_layers: int = 3, max_colat: float | None = None) -> None:
    """inparam_mesh of the reference MESHER; max_colat (LOCAL_MAX_COLAT, degrees)
    gives a regional mesh with absorbing boundaries."""
    with open(path, "w") as f:
        f.write("BACKGROUND_MODEL    %s\n" % model)
        f.write("DOMINANT_PERIOD     %.1f\n" % period_s)
        f.write("NTHETA_SLICES       1\nNRADIAL_SLICES      1\n")
        f.write("ONLY_SUGGEST_NTHETA false\nWRITE_VTK           false\n")
        f.write("WRITE_1DMODEL       false\n")
        f.write("COARSENING_LAYERS   %d\n" % coarsening_layers)
        if max_colat is not None:
            f.write("LOCAL_MAX_COLAT     %.1f\n" % max_colat)


def run_mesher(xmesh: str, outdir: str, period_s: float, model: str = "prem_iso",
               max_colat: float | None = None) -> str:
    """Run the reference MESHER binary into outdir; returns the meshdb path."""
    import subprocess
    import resource

    os.makedirs(os.path.join(outdir, "Diags"), exist_ok=True)
    write_mesh_inparam(os.path.join(outdir, "inparam_mesh"), period_s, model,
                       max_colat=max_colat)

    def unlimited_stack():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    env = dict(os.environ, OMP_STACKSIZE="4G")
    env.setdefault("OMP_NUM_THREADS", "8")
    with open(os.path.join(outdir, "OUTPUT_MESH"), "w") as log:
        subprocess.run([os.path.abspath(xmesh)], cwd=outdir, stdout=log, stderr=subprocess.STDOUT,
                       check=True, preexec_fn=unlimited_stack, env=env)
    return os.path.join(outdir, "meshdb.dat0000")


def write_run_dir(rundir: str, mesh_dir: str, cfg: RunConfig, copy_mesh: bool = False) -> str:
    """Create a SOLVER run directory using mesh_dir (meshdb.dat0000)."""
    os.makedirs(os.path.join(rundir, "Data"), exist_ok=True)
    os.makedirs(os.path.join(rundir, "Info"), exist_ok=True)
    mesh_link = os.path.join(rundir, "Mesh")
    if os.path.lexists(mesh_link):
        if os.path.islink(mesh_link):
            os.remove(mesh_link)
        else:
            shutil.rmtree(mesh_link)
    if copy_mesh:
        os.makedirs(mesh_link)
        for fn in ("meshdb.dat0000", "mesh_params.h"):
            src = os.path.join(mesh_dir, fn)
            if os.path.exists(src):
                shutil.copy(src, mesh_link)
    else:
        os.symlink(os.path.abspath(mesh_dir), mesh_link)
    with open(os.path.join(rundir, "inparam_basic"), "w") as f:
        f.write("SIMULATION_TYPE   %s\n" % cfg.simulation_type)
        f.write("SEISMOGRAM_LENGTH  %.1f\n" % cfg.seismogram_length_s)
        f.write("RECFILE_TYPE      stations\nMESHNAME           Mesh\n")
        f.write("LAT_HETEROGENEITY  false\n")
        f.write("ATTENUATION         %s\n" % ("true" if cfg.attenuation else "false"))
        f.write("SAVE_SNAPSHOTS         %s\n" % ("true" if cfg.snapshots else "false"))
    with open(os.path.join(rundir, "inparam_advanced"), "w") as f:
        f.write("SAMPLING_PERIOD     0.0\nTIME_STEP           0.0\nSOURCE_PERIOD       0.0\n")
        f.write("SOURCE_FUNCTION     %s\n" % cfg.source_function)
        f.write("TIME_SCHEME     %s\n" % cfg.time_scheme)
        f.write('DATA_DIR        "./Data"\nINFO_DIR        "./Info"\n')
        f.write("USE_NETCDF          false\nCHECKPOINTING       false\n")
        f.write("VERBOSITY               1\nDIAGNOSTIC_FILE_OUTPUT false\n")
        f.write("MESH_TEST           false\nKERNEL_WAVEFIELDS   false\n")
        f.write("SAVE_ENERGY         %s\n" % (".true." if cfg.energy else ".false."))
        f.write("NR_LIN_SOLIDS       %d\n" % cfg.n_sls)
        f.write("COARSE_GRAINED      %s\n" % ("true" if cfg.coarse_grained else "false"))
        if cfg.snapshots:
            f.write("SNAPSHOT_DT         %.3f\n" % cfg.snapshot_dt)
            f.write("SNAPSHOTS_FORMAT    %s\n" % cfg.snapshots)
            f.write("XDMF_GLL_I       '%s'\nXDMF_GLL_J       '%s'\n"
                    % (cfg.xdmf_gll_i, cfg.xdmf_gll_j))
        if cfg.extra_advanced:
            f.write(cfg.extra_advanced.rstrip("\n") + "\n")
    write_source(os.path.join(rundir, "inparam_source"), cfg)
    with open(os.path.join(rundir, "STATIONS"), "w") as f:
        for k, dist in enumerate(cfg.stations_deg):
            f.write("R%03d  AX  %9.4f  0.0000  0.0  0.0\n" % (k, 90.0 - dist))
    return rundir


def write_source(path: str, cfg: RunConfig) -> None:
    """inparam_source of a source on the axis (SOLVER/source.f90:40-140)."""
    with open(path, "w") as f:
        f.write("SOURCE_TYPE       %s\n" % cfg.source_type)
        f.write("SOURCE_DEPTH      %.3f\n" % cfg.source_depth_km)
        f.write("SOURCE_LAT         90.0\nSOURCE_LON          0.0\n")
        f.write("SOURCE_AMPLITUDE   %.6E\n" % cfg.source_amplitude)


# source type (src_type(1)) of each SOURCE_TYPE, SOLVER/source.f90:66-74
SOURCE_KIND = {"mrr": 0, "explosion": 0, "mtt_p_mpp": 0, "mpp_p_mtt": 0, "vertforce": 0,
               "mtr": 1, "mpr": 1, "mrt": 1, "mrp": 1, "thetaforce": 1, "phiforce": 1,
               "mtp": 2, "mpt": 2, "mtt_m_mpp": 2}


def record_states(host_exe: str, rundir: str, mesh_dir: str, cfgs: list, state_paths: list,
                  timeout: float = 3600) -> list:
    """States of several sources of one source type on one mesh from one run of
    the drop-in host (AXH_EXTRA_SOURCES): the reference setup of the mesh and
    model runs once, then the source routines once per source.  The configs
    may differ only in the source (type of one kind, depth, amplitude)."""
    import subprocess
    import resource

    if len(cfgs) != len(state_paths) or not cfgs:
        raise ValueError("one state path per config")
    kinds = {SOURCE_KIND[c.source_type.lower()] for c in cfgs}
    if len(kinds) != 1:
        raise ValueError("record_states: sources of one source type only")
    base = {k: v for k, v in cfgs[0].__dict__.items()
            if k not in ("source_type", "source_depth_km", "source_amplitude")}
    for c in cfgs[1:]:
        if {k: v for k, v in c.__dict__.items() if k in base} != base:
            raise ValueError("record_states: configs differ beyond the source")
    write_run_dir(rundir, mesh_dir, cfgs[0])
    lines = []
    for k, (c, sp) in enumerate(zip(cfgs[1:], state_paths[1:]), start=1):
        src = os.path.join(rundir, "inparam_source.%d" % k)
        write_source(src, c)
        lines.append("%s %s\n" % (os.path.abspath(src), os.path.abspath(sp)))
    lst = os.path.join(rundir, "extra_sources")
    with open(lst, "w") as f:
        f.writelines(lines)

    def unlimited_stack():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    env = dict(os.environ, AXH_SAVE_STATE=os.path.abspath(state_paths[0]), AXH_NO_RUN="1",
               AXH_EXTRA_SOURCES=os.path.abspath(lst))
    with open(os.path.join(rundir, "OUTPUT_record"), "w") as log:
        subprocess.run([os.path.abspath(host_exe)], cwd=rundir, stdout=log,
                       stderr=subprocess.STDOUT, check=True, timeout=timeout,
                       preexec_fn=unlimited_stack, env=env)
    return list(state_paths)


def record_state(host_exe: str, rundir: str, state_path: str, timeout: float = 1800,
                 kwf_strain_it: int = 0) -> str:
    """Run the drop-in host in record mode (no GPU): prepare_waves + uploads,
    written to state_path; the time loop is skipped.  kwf_strain_it > 0 also
    records the reference's kernel-wavefield grid (AXH_KWF_STRAIN_IT)."""
    import subprocess
    import resource

    def unlimited_stack():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    env = dict(os.environ, AXH_SAVE_STATE=os.path.abspath(state_path), AXH_NO_RUN="1")
    if kwf_strain_it > 0:
        env["AXH_KWF_STRAIN_IT"] = str(int(kwf_strain_it))
    with open(os.path.join(rundir, "OUTPUT_record"), "w") as log:
        subprocess.run([os.path.abspath(host_exe)], cwd=rundir, stdout=log,
                       stderr=subprocess.STDOUT, check=True, timeout=timeout,
                       preexec_fn=unlimited_stack, env=env)
    return state_path


def read_seismograms(data_dir: str, nrec: int, ncomp: int):
    """Receiver files R###_AX_disp.dat written by the reference or the drop-in
    (first line = the t=0 sample).  Returns float32 array (nsamp, nrec, ncomp)."""
    import numpy as np

    cols = []
    for k in range(nrec):
        a = np.loadtxt(os.path.join(data_dir, "R%03d_AX_disp.dat" % k), dtype=np.float64, ndmin=2)
        cols.append(a[:, :ncomp].astype(np.float32))
    return np.stack(cols, axis=1)


class SourceRecorder:
    """A drop-in host kept running in record mode for one mesh, model, time
    scheme and source type: the reference setup runs once (for the first
    source), then every further source is recorded on demand by the
    reference's own source routines (AXH_EXTRA_SOURCES read from a pipe), in
    seconds instead of a full setup.  record() is serialised per recorder."""

    def __init__(self, host_exe: str, rundir: str, mesh_dir: str, first: RunConfig,
                 first_state: str, timeout: float = 3600):
        import resource
        import subprocess
        import threading

        self.rundir = rundir
        self.base = {k: v for k, v in first.__dict__.items()
                     if k not in ("source_type", "source_depth_km", "source_amplitude")}
        self.kind = SOURCE_KIND[first.source_type.lower()]
        write_run_dir(rundir, mesh_dir, first)
        self.fifo = os.path.join(rundir, "source_pipe")
        if os.path.exists(self.fifo):
            os.remove(self.fifo)
        os.mkfifo(self.fifo)
        self._n = 0
        self._lock = threading.Lock()
        self._first = os.path.abspath(first_state)

        def unlimited_stack():
            resource.setrlimit(resource.RLIMIT_STACK,
                               (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

        env = dict(os.environ, AXH_SAVE_STATE=self._first, AXH_NO_RUN="1",
                   AXH_EXTRA_SOURCES=os.path.abspath(self.fifo))
        self.log = open(os.path.join(rundir, "OUTPUT_record"), "w")
        self.proc = subprocess.Popen([os.path.abspath(host_exe)], cwd=rundir,
                                     stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                     text=True, preexec_fn=unlimited_stack, env=env)
        self.timeout = timeout
        self._wait_for(self._first)
        # the host opens the pipe for reading after the first state
        self._pipe = open(self.fifo, "w")

    def _wait_for(self, path: str) -> None:
        want = "wrote GPU state file " + path
        for line in self.proc.stdout:
            self.log.write(line)
            if line.strip() == want.strip():
                self.log.flush()
                return
        self.log.flush()
        raise RuntimeError("the drop-in host ended before writing %s (see %s)"
                           % (path, self.log.name))

    @property
    def first_state(self) -> str:
        return self._first

    def record(self, cfg: RunConfig, state_path: str) -> str:
        if SOURCE_KIND[cfg.source_type.lower()] != self.kind:
            raise ValueError("a recorder serves one source type")
        if {k: v for k, v in cfg.__dict__.items() if k in self.base} != self.base:
            raise ValueError("a recorder serves configs that differ in the source only")
        with self._lock:
            self._n += 1
            src = os.path.join(self.rundir, "inparam_source.%d" % self._n)
            write_source(src, cfg)
            path = os.path.abspath(state_path)
            self._pipe.write("%s %s\n" % (os.path.abspath(src), path))
            self._pipe.flush()
            self._wait_for(path)
            return path

    def close(self) -> None:
        if self._pipe is not None:
            self._pipe.close()
            self._pipe = None
        if self.proc is not None:
            try:
                rest = self.proc.communicate(timeout=self.timeout)[0]
                self.log.write(rest or "")
            finally:
                self.proc = None
                self.log.close()
