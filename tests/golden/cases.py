"""Golden cases: small runs of the reference SOLVER on the 200 s prem_iso mesh.

Each case fixes the inputs (run directory) of one reference run; the recorded
outputs live in tests/golden/<name>/ref_seis.npz (written by make_golden.py).
"""
from __future__ import annotations

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from axisem_amd.inputs import RunConfig  # noqa: E402

MESH_DIR = os.path.join(HERE, "mesh_prem_iso_200s")
MESH_PERIOD = 200.0
# the transversely isotropic PREM (prem_ani) at 200 s
ANI_MESH_DIR = os.path.join(HERE, "mesh_prem_ani_200s")
# the 50 s prem_iso mesh of the plain CPU-runnable benchmark configuration
MESH50_DIR = os.path.join(HERE, "mesh_prem_iso_50s")
MESH50_PERIOD = 50.0
STATIONS_3DEG = [3.0 * k for k in range(61)]
# a regional mesh (colatitude 0-60 degrees): absorbing boundaries
REGIONAL_MESH_DIR = os.path.join(HERE, "mesh_prem_iso_200s_colat60")
REGIONAL_MAX_COLAT = 60.0
REGIONAL_STATIONS = [2.0 * k for k in range(30)]
STATIONS = [10.0 * k for k in range(19)]
SAMPLE_STRIDE = 5          # every 5th seismogram sample is stored
# (every 20th for the 50 s run of 12569 samples)
STRIDES = {"p50_mono_vertforce_newmark": 20}


def sample_stride(name: str) -> int:
    return STRIDES.get(name, SAMPLE_STRIDE)

CASES = {
    # monopole, surface vertical force, Newmark, elastic
    "mono_vertforce_newmark": RunConfig("vertforce", 0.0, 1e20, "newmark2", False, 900.0, STATIONS),
    # monopole explosion at depth, symplectic PEFRL, anelastic (cg4 memory variables)
    "mono_explosion_symplec4_anel": RunConfig("explosion", 2000.0, 1e20, "symplec4", True, 600.0,
                                              STATIONS),
    # monopole Mtt+Mpp at 300 km, Newmark, anelastic
    "mono_mttpmpp_newmark_anel": RunConfig("mtt_p_mpp", 300.0, 1e20, "newmark2", True, 900.0,
                                           STATIONS),
    # dipole horizontal force at 50 km, Newmark, anelastic
    "dipole_thetaforce_newmark_anel": RunConfig("thetaforce", 50.0, 1e20, "newmark2", True, 900.0,
                                                STATIONS),
    # dipole phi-force at 1000 km, symplectic, anelastic
    "dipole_phiforce_symplec4_anel": RunConfig("phiforce", 1000.0, 1e20, "symplec4", True, 600.0,
                                               STATIONS),
    # dipole Mtr just above the core-mantle boundary, symplectic, elastic
    "dipole_mtr_symplec4": RunConfig("mtr", 2800.0, 1e20, "symplec4", False, 600.0, STATIONS),
    # quadrupole Mtp, Newmark, elastic
    "quad_mtp_newmark": RunConfig("mtp", 100.0, 1e20, "newmark2", False, 900.0, STATIONS),
    # quadrupole Mtt-Mpp at depth, symplectic, anelastic
    "quad_mttmmpp_symplec4_anel": RunConfig("mtt_m_mpp", 600.0, 1e20, "symplec4", True, 600.0,
                                            STATIONS),
    # regional mesh with absorbing boundaries: monopole, Newmark, elastic; and
    # dipole, symplectic, anelastic
    "reg_mono_vertforce_newmark": RunConfig("vertforce", 10.0, 1e20, "newmark2", False, 900.0,
                                            REGIONAL_STATIONS),
    "reg_dipole_thetaforce_symplec4_anel": RunConfig("thetaforce", 30.0, 1e20, "symplec4", True,
                                                     600.0, REGIONAL_STATIONS),
    # memory variables at all 25 points of an element (COARSE_GRAINED false):
    # monopole Newmark, dipole symplectic, quadrupole Newmark
    "mono_explosion_newmark_anelfull": RunConfig("explosion", 100.0, 1e20, "newmark2", True, 600.0,
                                                 STATIONS, coarse_grained=False),
    "dipole_thetaforce_symplec4_anelfull": RunConfig("thetaforce", 50.0, 1e20, "symplec4", True,
                                                     500.0, STATIONS, coarse_grained=False),
    "quad_mtp_newmark_anelfull": RunConfig("mtp", 300.0, 1e20, "newmark2", True, 600.0, STATIONS,
                                           coarse_grained=False),
    # transversely isotropic PREM, anelastic (the large single-GPU
    # configuration's model and physics at 200 s): the moment-tensor runs
    "ani_mono_mrr_symplec4_anel": RunConfig("mrr", 100.0, 1e20, "symplec4", True, 600.0,
                                            STATIONS),
    "ani_dipole_mtr_symplec4_anel": RunConfig("mtr", 100.0, 1e20, "symplec4", True, 600.0,
                                              STATIONS),
    "ani_quad_mtp_symplec4_anel": RunConfig("mtp", 100.0, 1e20, "symplec4", True, 600.0,
                                            STATIONS),
    "ani_dipole_thetaforce_newmark_anel": RunConfig("thetaforce", 50.0, 1e20, "newmark2", True,
                                                    900.0, STATIONS),
    # the plain CPU-runnable benchmark configuration, full length: prem_iso
    # 50 s, elastic, newmark2, monopole vertical surface force, Gaussian STF at
    # the mesh period, 1800 s at 61 surface receivers 3 degrees apart
    "p50_mono_vertforce_newmark": RunConfig("vertforce", 0.0, 1e20, "newmark2", False, 1800.0,
                                            STATIONS_3DEG),
}

# Source batches: several sources on the mesh above with one source type and
# time scheme, stepped together by the batched time loop.  There is no
# reference output for them; each member is checked against its own solo run
# and against the oracle.
BATCH_MEMBERS = {
    "b_dipole_thetaforce_50": RunConfig("thetaforce", 50.0, 1e20, "symplec4", True, 600.0,
                                        STATIONS),
    "b_dipole_phiforce_1000": RunConfig("phiforce", 1000.0, 1e20, "symplec4", True, 600.0,
                                        STATIONS),
    "b_dipole_mtr_700": RunConfig("mtr", 700.0, 3e20, "symplec4", True, 600.0, STATIONS),
    "b_dipole_mpr_2000": RunConfig("mpr", 2000.0, 1e20, "symplec4", True, 600.0,
                                   STATIONS[::2]),
    "b_mono_mtt_p_mpp_300": RunConfig("mtt_p_mpp", 300.0, 1e20, "newmark2", True, 900.0,
                                      STATIONS),
    "b_mono_explosion_100": RunConfig("explosion", 100.0, 1e20, "newmark2", True, 900.0,
                                      STATIONS),
    "b_mono_vertforce_0": RunConfig("vertforce", 0.0, 2e20, "newmark2", True, 700.0, STATIONS),
    "b_quad_mtp_100": RunConfig("mtp", 100.0, 1e20, "symplec4", True, 500.0, STATIONS),
    "b_quad_mtt_m_mpp_1500": RunConfig("mtt_m_mpp", 1500.0, 2e20, "symplec4", True, 500.0,
                                       STATIONS),
    "b_quad_mtp_400": RunConfig("mtp", 400.0, 1e20, "symplec4", True, 500.0, STATIONS[1::2]),
    "b_quad_mtt_m_mpp_20": RunConfig("mtt_m_mpp", 20.0, 1e20, "symplec4", True, 500.0,
                                     STATIONS),
    "b_quad_mtp_2600": RunConfig("mtp", 2600.0, 1e20, "symplec4", True, 500.0, STATIONS),
    # memory variables at all 25 points
    "bf_dipole_thetaforce_50": RunConfig("thetaforce", 50.0, 1e20, "symplec4", True, 400.0,
                                         STATIONS, coarse_grained=False),
    "bf_dipole_mtr_700": RunConfig("mtr", 700.0, 3e20, "symplec4", True, 400.0, STATIONS,
                                   coarse_grained=False),
    "bf_dipole_phiforce_1000": RunConfig("phiforce", 1000.0, 1e20, "symplec4", True, 400.0,
                                         STATIONS, coarse_grained=False),
    "bf_dipole_mpr_2000": RunConfig("mpr", 2000.0, 1e20, "symplec4", True, 400.0,
                                    STATIONS[::2], coarse_grained=False),
    "bf_mono_explosion_100": RunConfig("explosion", 100.0, 1e20, "newmark2", True, 500.0,
                                       STATIONS, coarse_grained=False),
    "bf_mono_vertforce_0": RunConfig("vertforce", 0.0, 2e20, "newmark2", True, 500.0, STATIONS,
                                     coarse_grained=False),
    "bf_mono_mtt_p_mpp_300": RunConfig("mtt_p_mpp", 300.0, 1e20, "newmark2", True, 500.0,
                                       STATIONS, coarse_grained=False),
}
# the four runs of a moment-tensor source (monopole Mzz and Mxx+Myy, dipole
# Mxz, quadrupole Mxy) at several depths: one batch of mixed source types
MT_RUNS = ("mrr", "mtt_p_mpp", "mtr", "mtp")
MT_DEPTHS = (10.0, 100.0, 300.0, 600.0)
for _d in MT_DEPTHS:
    for _r in MT_RUNS:
        BATCH_MEMBERS["mt_%s_%g" % (_r, _d)] = RunConfig(_r, _d, 1e20, "newmark2", False, 600.0,
                                                       STATIONS)
for _r in MT_RUNS:
    BATCH_MEMBERS["mta_%s_100" % _r] = RunConfig(_r, 100.0, 1e20, "symplec4", True, 500.0,
                                                 STATIONS)
    BATCH_MEMBERS["ani_mt_%s_100" % _r] = RunConfig(_r, 100.0, 1e20, "symplec4", True, 400.0,
                                                    STATIONS)
BATCHES = {
    # mixed source types: one moment tensor at 100 km (elastic newmark2), the
    # same at four depths (16 wavefields), and anelastic symplec4 at 100 km
    "mt4_newmark": ["mt_%s_100" % r for r in MT_RUNS],
    "mt16_newmark": ["mt_%s_%g" % (r, d) for d in MT_DEPTHS for r in MT_RUNS],
    "mt4_symplec4_anel": ["mta_%s_100" % r for r in MT_RUNS],
    # the same on transversely isotropic PREM
    "ani_mt4_symplec4_anel": ["ani_mt_%s_100" % r for r in MT_RUNS],
    "dipole_symplec4_anel": ["b_dipole_thetaforce_50", "b_dipole_phiforce_1000",
                             "b_dipole_mtr_700", "b_dipole_mpr_2000"],
    "mono_newmark_anel": ["b_mono_mtt_p_mpp_300", "b_mono_explosion_100", "b_mono_vertforce_0"],
    # five members: one group of 4 and one of 1 at the default width
    "quad_symplec4_anel": ["b_quad_mtp_100", "b_quad_mtt_m_mpp_1500", "b_quad_mtp_400",
                           "b_quad_mtt_m_mpp_20", "b_quad_mtp_2600"],
    "dipole_symplec4_anelfull": ["bf_dipole_thetaforce_50", "bf_dipole_mtr_700",
                                 "bf_dipole_phiforce_1000", "bf_dipole_mpr_2000"],
    "mono_newmark_anelfull": ["bf_mono_explosion_100", "bf_mono_vertforce_0",
                              "bf_mono_mtt_p_mpp_300"],
}
# Cases the reference program cannot run (it lower-cases TIME_SCHEME and its
# symplectic_coefficients then matches neither McLachlan scheme): the GPU is
# checked against the oracle only, whose schemes are pinned by their order.
ORACLE_CASES = {
    "mono_vertforce_mlso4m5": RunConfig("vertforce", 0.0, 1e20, "ML_SO4m5", False, 600.0,
                                        STATIONS),
    "dipole_thetaforce_mlso6m7_anel": RunConfig("thetaforce", 50.0, 1e20, "ML_SO6m7", True,
                                                500.0, STATIONS),
    "quad_mtp_mlso4m5_anel": RunConfig("mtp", 300.0, 1e20, "ML_SO4m5", True, 500.0, STATIONS),
}
# Kernel wavefields (displ_only).  The reference cannot write them without
# NetCDF, but its xdmf snapshots hold the same values: the fluid displacement
# inv_rho_fluid * grad chi and the solid displacement in (s, phi, z), at the
# GLL points XDMF_GLL_I x XDMF_GLL_J; with ipol 0..3 and jpol 1..3 they cover
# every point of the default kernel-wavefield grid (1..3 x 1..3 and the axis
# points).  tests/golden/<case>/ref_xdmf.npz holds one such snapshot.
KWF_CASES = {
    "kwf_dipole_thetaforce_newmark": RunConfig("thetaforce", 50.0, 1e20, "newmark2", False,
                                               720.0, STATIONS, snapshots="xdmf",
                                               snapshot_dt=350.0, xdmf_gll_i="4 0 1 2 3",
                                               xdmf_gll_j="3 1 2 3"),
    "kwf_mono_explosion_symplec4": RunConfig("explosion", 600.0, 1e20, "symplec4", False,
                                             720.0, STATIONS, snapshots="xdmf",
                                             snapshot_dt=350.0, xdmf_gll_i="4 0 1 2 3",
                                             xdmf_gll_j="3 1 2 3"),
}
ALL_CONFIGS = {**CASES, **BATCH_MEMBERS, **ORACLE_CASES, **KWF_CASES}

NCOMP = {"vertforce": 2, "explosion": 2, "mrr": 2, "mtt_p_mpp": 2,
         "thetaforce": 3, "phiforce": 3, "mtr": 3, "mpr": 3, "mtp": 3, "mtt_m_mpp": 3}


# Runs with wavefield snapshots (SAVE_SNAPSHOTS) or energies (SAVE_ENERGY): the
# reference's output files are fixed by their SHA-256 in
# tests/golden/<name>/outputs.json.
OUTPUT_CASES = {
    "snap_mono_vertforce_newmark": RunConfig("vertforce", 0.0, 1e20, "newmark2", False, 300.0,
                                             STATIONS, snapshots="both", snapshot_dt=60.0),
    "snap_dipole_thetaforce_symplec4_anel": RunConfig("thetaforce", 50.0, 1e20, "symplec4", True,
                                                      300.0, STATIONS, snapshots="xdmf",
                                                      snapshot_dt=45.0),
    "energy_dipole_thetaforce_newmark_anel": RunConfig("thetaforce", 50.0, 1e20, "newmark2", True,
                                                       300.0, STATIONS, energy=True),
    "energy_mono_explosion_symplec4": RunConfig("explosion", 600.0, 1e20, "symplec4", False,
                                                300.0, STATIONS, energy=True),
}
# the output files the snapshot comparison covers
OUTPUT_FILES = ("snap_", "xdmf_snap_", "glob_grid_", "xdmf_grid_", "xdmf_points_",
                  "xdmf_xml_", "xdmf_meshonly_", "energy_")


def mesh_of(name: str) -> str:
    """The mesh directory a case runs on."""
    if name.startswith("reg_"):
        return REGIONAL_MESH_DIR
    if name.startswith("p50_"):
        return MESH50_DIR
    if name.startswith("ani"):
        return ANI_MESH_DIR
    return MESH_DIR


def case_dir(name: str) -> str:
    return os.path.join(HERE, name)
