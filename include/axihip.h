/*
 * axihip.h -- C-ABI of the MI355X (gfx950) AxiSEM SOLVER time loop.
 *
 * This is the drop-in boundary for the reference's time loop: the reference
 * program calls `time_loop` (SOLVER/main.f90:92) of module `time_evol_wave`
 * (SOLVER/time_evol_wave.F90:26-41, public prepare_waves/time_loop, no
 * arguments, all state in module globals).  A replacement `time_evol_wave`
 * module written against this header hands those globals to the library
 * through iso_c_binding and runs the loop on one GPU.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Every array is in the reference's Fortran
 *    layout (column major, npol=4):
 *      element fields/coefficients  (0:4,0:4,nel)   index i + 5 j + 25 (e-1)
 *      solid fields with components (0:4,0:4,nel,3) index ... + 25 nel (c-1)
 *      axial vectors                (0:4,nel)
 *      coarse-grained (cg4) arrays  (1:4,nel)       points (1,1),(1,3),(3,1),(3,3)
 *    Element and point indices passed in are 1-based, as in the reference.
 *  - The library copies every array at upload and keeps no caller pointer.
 *  - Functions return AXH_OK (0) or a negative AXH_ERR_* code; the message of
 *    the last failure on the calling thread is returned by axh_last_error().
 *    No C++ exception crosses this boundary.
 *  - Nothing touches the GPU before the first axh_run / axh_download_fields /
 *    axh_device_bytes call, so a context can be filled and written to a state
 *    file on a host without a GPU (axh_save_state).
 *  - One host thread per context; the library owns its HIP stream.
 */
#ifndef AXIHIP_H
#define AXIHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AXH_ABI_VERSION 1
#define AXH_MAX_BATCH 16    /* wavefields per batch */
#define AXH_MAX_GROUP 4     /* wavefields per kernel instance */
#define AXH_MAX_GRAPH_STEPS 64  /* time steps per replayed HIP graph */

enum axh_status {
  AXH_OK = 0,
  AXH_ERR_ARG = -1,          /* bad argument / inconsistent sizes           */
  AXH_ERR_STATE = -2,        /* call out of order, missing upload           */
  AXH_ERR_HIP = -3,          /* HIP runtime failure                         */
  AXH_ERR_IO = -4,           /* state file read/write failure               */
  AXH_ERR_UNSUPPORTED = -5,  /* input the GPU path does not implement       */
  AXH_ERR_BLOWUP = -6        /* |disp(1,1,:,:)| > 10 |magnitude|            */
};

/* src_type(1) of SOLVER/data_source.f90:31 */
enum axh_src_type { AXH_MONOPOLE = 0, AXH_DIPOLE = 1, AXH_QUADPOLE = 2 };

/* time_scheme of SOLVER/data_time.f90 (values of inparam_advanced TIME_SCHEME) */
enum axh_scheme {
  AXH_NEWMARK2 = 0,
  AXH_SYMPLEC4 = 1,   /* PEFRL,            SOLVER/time_evol_wave.F90:763-791 */
  AXH_ML_SO4M5 = 2,   /* McLachlan 4th,    SOLVER/time_evol_wave.F90:794-828 */
  AXH_ML_SO6M7 = 3    /* McLachlan 6th,    SOLVER/time_evol_wave.F90:831-866 */
};

/* Problem dimensions and run constants (filled from data_mesh, data_time,
 * data_source, attenuation by the Fortran host). */
typedef struct axh_dims {
  int32_t abi_version;      /* = AXH_ABI_VERSION                          */
  int32_t npol;             /* must be 4                                  */
  int32_t nel_solid, nel_fluid;
  int32_t nglob_solid, nglob_fluid;   /* ranges of igloc_solid/igloc_fluid */
  int32_t nel_bdry;         /* solid-fluid boundary element pairs         */
  int32_t src_type;         /* enum axh_src_type                          */
  int32_t scheme;           /* enum axh_scheme                            */
  int32_t anel;             /* anel_true                                  */
  int32_t coarse_grained;   /* att_coarse_grained                         */
  int32_t n_sls;            /* n_sls_attenuation                          */
  int32_t niter;            /* number of time steps                       */
  int32_t seis_it;          /* seismogram sampling interval in steps      */
  int32_t num_rec;          /* receivers (recfile_el rows)                */
  int32_t nelsrc;           /* source elements (<= 8)                     */
  int32_t fluid_src;        /* source in the fluid (explosion only)       */
  int32_t reserved[5];
  double deltat;            /* time step [s]                              */
  double magnitude;         /* source magnitude, for the blow-up check    */
} axh_dims;

typedef struct axh_ctx axh_ctx;

/* ---- lifetime ------------------------------------------------------------ */
int axh_create(const axh_dims *dims, axh_ctx **out);
int axh_destroy(axh_ctx *ctx);
const char *axh_last_error(void);

/* ---- uploads (SOLVER/data_spec.f90, data_matr.f90, data_mesh.f90) ------ */

/* G0(0:4), G1/G1T/G2/G2T(0:4,0:4): data_spec (SOLVER/data_spec.f90:35-36). */
int axh_upload_spectral(axh_ctx *ctx, const float *G0, const float *G1,
                        const float *G1T, const float *G2, const float *G2T);

/* Precomputed solid stiffness terms of data_matr (SOLVER/data_matr.f90:45-78),
 * built by def_solid_stiffness_terms (SOLVER/def_precomp_terms.f90:1166-1626).
 * 3-D terms are (0:4,0:4,nel_solid); AXH_M0_W* terms are (0:4,nel_solid). */
enum axh_solid_term {
  AXH_M11S = 0, AXH_M21S, AXH_M41S, AXH_M12S, AXH_M22S, AXH_M42S,
  AXH_M11Z, AXH_M21Z, AXH_M41Z, AXH_M32S, AXH_M13S, AXH_M33S, AXH_M43S,
  AXH_M1PHI, AXH_M2PHI, AXH_M4PHI,
  AXH_M_1, AXH_M_2, AXH_M_3, AXH_M_4, AXH_M_5, AXH_M_6, AXH_M_7, AXH_M_8,
  AXH_M_W1, AXH_M_W2, AXH_M_W3, AXH_M_W4, AXH_M_W5,
  AXH_N_SOLID_3D,
  AXH_M0_W1 = 64, AXH_M0_W2, AXH_M0_W3, AXH_M0_W4, AXH_M0_W5, AXH_M0_W6,
  AXH_M0_W7, AXH_M0_W8, AXH_M0_W9, AXH_M0_W10,
  AXH_SOLID_TERM_END
};
int axh_upload_solid_term(axh_ctx *ctx, int term, const float *data);

/* Fluid terms (SOLVER/data_matr.f90:80-84): M1chi_fl, M2chi_fl, M4chi_fl,
 * M_w_fl are (0:4,0:4,nel_fluid); M0_w_fl is (0:4,nel_fluid). */
enum axh_fluid_term {
  AXH_M1CHI_FL = 0, AXH_M2CHI_FL, AXH_M4CHI_FL, AXH_M_W_FL,
  AXH_M0_W_FL = 64
};
int axh_upload_fluid_term(axh_ctx *ctx, int term, const float *data);

/* Assembled inverse mass matrices inv_mass_rho (0:4,0:4,nel_solid) and
 * inv_mass_fluid (0:4,0:4,nel_fluid) (SOLVER/data_matr.f90:36-37). */
int axh_upload_mass(axh_ctx *ctx, const float *inv_mass_rho,
                    const float *inv_mass_fluid);

/* Global numbering igloc_solid(25 nel_solid), igloc_fluid(25 nel_fluid)
 * (SOLVER/data_mesh.f90:42-43, 1-based) and the axis flags axis_solid /
 * axis_fluid (LOGICAL in SOLVER/data_mesh.f90:88-89, passed as 0/1 int32). */
int axh_upload_mesh_maps(axh_ctx *ctx, const int32_t *igloc_solid,
                         const int32_t *igloc_fluid, const int32_t *axis_solid,
                         const int32_t *axis_fluid);

/* Solid-fluid boundary (SOLVER/data_mesh.f90:93-95, data_matr.f90:90):
 * bdry_solid_el, bdry_fluid_el, bdry_jpol_solid, bdry_jpol_fluid (nel_bdry,
 * 1-based element indices) and bdry_matr(0:4,nel_bdry,2). */
int axh_upload_boundary(axh_ctx *ctx, const int32_t *bdry_solid_el,
                        const int32_t *bdry_fluid_el,
                        const int32_t *bdry_jpol_solid,
                        const int32_t *bdry_jpol_fluid, const float *bdry_matr);

/* fluid_free_surface_mask(0:4,0:4,nel_fluid) (SOLVER/data_mesh.f90:181),
 * built by create_fluid_free_surface_mask (SOLVER/time_evol_wave.F90:1615). */
int axh_upload_fluid_surface_mask(axh_ctx *ctx, const float *mask);

/* Absorbing boundaries of a mesh that does not reach the antipode
 * (LOCAL_MAX_COLAT): solid_absorbing_gamma(0:4,0:4,nel_solid) and
 * fluid_absorbing_gamma(0:4,0:4,nel_fluid) (SOLVER/data_mesh.f90:184-187;
 * may be NULL when nel_fluid is 0), as create_absorbing_gamma builds them
 * (SOLVER/time_evol_wave.F90:1634-1713).  The loop then adds
 * -2 gamma v - gamma^2 u to every acceleration after the inverse mass
 * (SOLVER/time_evol_wave.F90:432-434,468-491,663-711).  Not uploading them
 * means no absorbing boundaries. */
int axh_upload_absorbing(axh_ctx *ctx, const float *gamma_solid,
                         const float *gamma_fluid);

/* Source (SOLVER/data_source.f90:47-49): ielsrc(nelsrc) 1-based, the element
 * source term source_term_el(0:4,0:4,8,3) as compute_src allocates it
 * (SOLVER/source.f90:376; source_term_fl(0:4,0:4,8) when fluid_src), and the
 * source time function table:
 * newmark2: stf(niter) as realkind (SOLVER/data_source.f90:37);
 * symplectic: real(stf_symp(i), realkind) for every step and stage,
 * stf[(iter-1)*nstages + (i-1)] (SOLVER/time_evol_wave.F90:590-593,689). */
int axh_upload_source(axh_ctx *ctx, const int32_t *ielsrc,
                      const float *source_term, const float *stf,
                      int32_t nstf);

/* Coarse-grained anelastic terms (SOLVER/data_matr.f90:103-117,
 * data_pointwise.f90:40-43, attenuation.f90:31-44).  All float arrays are
 * (1:4,nel_solid): Y_cg4, V_s_eta_cg4, V_s_xi_cg4, V_z_eta_cg4, V_z_xi_cg4,
 * DsDeta/DzDeta/DsDxi/DzDxi_over_J_sol_cg4, inv_s_solid at the 4 cg points,
 * delta_mu_cg4, delta_kappa_cg4.  exp_w_j_deltat, ts_fac_t, ts_fac_tm1 are
 * (n_sls) doubles; a_j_mu, a_j_kappa are (n_sls, nel_solid) doubles: the
 * normalised SLS weights of time_step_memvars_cg4
 * (SOLVER/attenuation.f90:105-127) for every element. */
typedef struct axh_attenuation_cg4 {
  const float *Y, *V_s_eta, *V_s_xi, *V_z_eta, *V_z_xi;
  const float *DsDeta_over_J, *DzDeta_over_J, *DsDxi_over_J, *DzDxi_over_J;
  const float *inv_s;
  const float *delta_mu, *delta_kappa;
  const double *exp_w_j_deltat, *ts_fac_t, *ts_fac_tm1;
  const double *a_j_mu, *a_j_kappa;
} axh_attenuation_cg4;
int axh_upload_attenuation_cg4(axh_ctx *ctx, const axh_attenuation_cg4 *att);

/* Anelastic terms at all 25 points, for att_coarse_grained false
 * (COARSE_GRAINED in inparam_advanced, SOLVER/attenuation.f90:812-813;
 * used by glob_anel_stiffness_*_4 and time_step_memvars_4,
 * SOLVER/stiffness_mono.f90:409-507, attenuation.f90:210-334).  Float arrays
 * (0:4,0:4,nel_solid): Y, V_s_eta, V_s_xi, V_z_eta, V_z_xi
 * (SOLVER/data_matr.f90:101-110), DsDeta/DzDeta/DsDxi/DzDxi_over_J_sol and
 * inv_s_solid (SOLVER/data_pointwise.f90), delta_mu, delta_kappa
 * (SOLVER/attenuation.f90:1052-1054); (0:4,nel_solid): Y0, V0_s_eta,
 * V0_s_xi, V0_z_eta, V0_z_xi (SOLVER/def_precomp_terms.f90:1294-1298).  The
 * SLS coefficients are those of axh_upload_attenuation_cg4, with a_j_mu,
 * a_j_kappa of time_step_memvars_4 (SOLVER/attenuation.f90:244-262). */
typedef struct axh_attenuation_full {
  const float *Y, *V_s_eta, *V_s_xi, *V_z_eta, *V_z_xi;
  const float *Y0, *V0_s_eta, *V0_s_xi, *V0_z_eta, *V0_z_xi;
  const float *DsDeta_over_J, *DzDeta_over_J, *DsDxi_over_J, *DzDxi_over_J;
  const float *inv_s;
  const float *delta_mu, *delta_kappa;
  const double *exp_w_j_deltat, *ts_fac_t, *ts_fac_tm1;
  const double *a_j_mu, *a_j_kappa;
} axh_attenuation_full;
int axh_upload_attenuation_full(axh_ctx *ctx, const axh_attenuation_full *att);

/* Receivers: recfile_el(num_rec,3) = (element, ipol, jpol) in the solid,
 * element 1-based, ipol/jpol 0-based (SOLVER/data_mesh.f90:135). */
int axh_upload_receivers(axh_ctx *ctx, const int32_t *recfile_el);

/* Build the library's layout from the uploads without touching the GPU and
 * report inputs the GPU loop cannot represent (AXH_ERR_UNSUPPORTED: e.g.
 * copies of one GLL point with different inverse mass).  axh_run does the
 * same checks; this entry point lets a host fail before any device work. */
int axh_check_inputs(axh_ctx *ctx);

/* The library's numbering of the GLL points of a domain (0 solid, 1 fluid):
 * gid(25 nel) in the reference's element-local order, 0-based, and the number
 * of distinct points.  It is a relabelling of igloc_solid / igloc_fluid
 * (SOLVER/data_mesh.f90:42-43): gid(a) == gid(b) iff igloc(a) == igloc(b). */
int axh_global_numbering(axh_ctx *ctx, int32_t domain, int32_t *gid,
                         int32_t *nglob);

/* ---- running ------------------------------------------------------------- */

/* Advance steps it_begin..it_end (1-based, inclusive; it_begin must be one
 * past the last step run).  On return seis (may be NULL) holds, for every
 * step with mod(iter, seis_it) == 0 inside the range, the receiver samples of
 * compute_recfile_seis_bare (SOLVER/seismograms.f90:742-780) as
 * seis[((k * num_rec) + r) * 3 + c], c in the written order
 * (monopole: u_s, u_z, 0; dipole: u_1+u_2, u_1-u_2, u_3; quadpole: u_s,
 * u_phi, u_z), k counting the sampled steps of this call from 0. */
int axh_run(axh_ctx *ctx, int32_t it_begin, int32_t it_end, float *seis);

/* axh_run that allocates nothing: no hipMalloc, hipFree or other allocating
 * call, for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid it).  No counterpart in the reference.
 * The context must have run (axh_run) before, so that its device part exists;
 * the buffers the run writes besides the fields -- the step counter, the
 * blow-up flag and the staging of the receiver samples -- are a fixed array
 * of AXH_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, shared
 * by every call on that GPU: calls that overlap on one GPU race.  At most
 * (AXH_NOALLOC_SCRATCH_BYTES - 256) / 12 receivers (AXH_ERR_ARG beyond);
 * no per-kernel profiling.  Results equal axh_run's bit for bit. */
#define AXH_NOALLOC_SCRATCH_BYTES (1 << 20)
int axh_run_noalloc(axh_ctx *ctx, int32_t it_begin, int32_t it_end, float *seis);

/* axh_run in two halves, for a host that works while the GPU runs (the
 * drop-in host writes one chunk's samples while the next chunk runs):
 * axh_run_begin queues steps it_begin..it_end on the context's stream and
 * returns; axh_run_end waits for them and writes seis (which must stay valid
 * and untouched in between), with axh_run's results and errors.  Between the
 * two the context takes no other call but axh_destroy. */
int axh_run_begin(axh_ctx *ctx, int32_t it_begin, int32_t it_end, float *seis);
int axh_run_end(axh_ctx *ctx);

/* Number of seismogram samples axh_run(ctx, it_begin, it_end) will write. */
int axh_seis_samples(const axh_ctx *ctx, int32_t it_begin, int32_t it_end);

/* Copy the current fields back in the reference's element-local layout:
 * disp/velo (0:4,0:4,nel_solid,3), chi/dchi/ddchi (0:4,0:4,nel_fluid).
 * ddchi is ddchi0 (newmark2) or ddchi (symplectic).  Any pointer may be NULL. */
int axh_download_fields(axh_ctx *ctx, float *disp, float *velo, float *chi,
                        float *dchi, float *ddchi);

/* Replace the current fields (same layout as axh_download_fields); copies of
 * one global point must agree.  Any pointer may be NULL (left unchanged). */
int axh_upload_fields(axh_ctx *ctx, const float *disp, const float *velo,
                      const float *chi, const float *dchi, const float *ddchi);

/* Time spent in the last axh_run on the device, from HIP events on the
 * library stream [ms]. */
int axh_last_run_ms(const axh_ctx *ctx, double *total_ms);

/* Per-kernel timing.  When on, every kernel launch of axh_run is bracketed by
 * HIP events on the library stream and its duration is accumulated per kernel
 * class; axh_kernel_name(k) names class k (NULL past the last class). */
int axh_set_profiling(axh_ctx *ctx, int on);
const char *axh_kernel_name(int kclass);
int axh_kernel_stats(const axh_ctx *ctx, int kclass, double *total_ms,
                     int64_t *launches);

/* Launch mode of axh_run: steps = 0 launches every kernel directly (the
 * default); steps = S > 0 captures S time steps of the context once into a
 * HIP graph and replays it for every S steps of a run (the rest of a run,
 * and runs with per-kernel profiling on, launch directly).  Each kernel reads
 * its time step from device memory, so the replays need no host work per
 * step; results are bit-identical either way.  Small meshes, whose steps are
 * bound by launch latency, gain.  The reference has no counterpart (its time
 * loop is a host loop, SOLVER/time_evol_wave.F90:353-500). */
int axh_set_graph_steps(axh_ctx *ctx, int32_t steps);

/* Device memory held by the context [bytes] (0 before the first run). */
int axh_device_bytes(axh_ctx *ctx, size_t *bytes);

/* ---- batches of wavefields ------------------------------------------------ */

/* A batch advances several contexts that share one mesh, in lock step, with
 * one launch per kernel for all of them.  The reference runs one source per
 * SOLVER process (SOLVER/main.f90:92, time_loop of
 * SOLVER/time_evol_wave.F90:26-41) and a multi-source study runs the program
 * once per source; a batch is that loop over sources with the mesh terms read
 * from HBM once per step for all members.
 *
 * Members must have equal dims except src_type, nelsrc, fluid_src, num_rec,
 * seis_it, niter and magnitude, and the same global numbering.  Members of
 * one source type must have byte-equal mesh terms (every upload except
 * ielsrc, source_term, stf and recfile_el, the per-source inputs); members of
 * different source types -- the monopole, dipole and quadrupole runs of one
 * moment tensor (SOLVER/source.f90:40-140, one SOLVER run each in the
 * reference's submit.py:276-300) -- each step on their type's terms.  Each
 * member's fields, seismograms and blow-up test are its own and bit-identical
 * to stepping it alone with axh_run.  1 <= n <= AXH_MAX_BATCH.  The batch
 * does not own the members; a member in a batch takes no uploads and
 * axh_destroy refuses it until axh_batch_destroy.  Like axh_create, creating
 * a batch touches no GPU. */
typedef struct axh_batch axh_batch;
int axh_batch_create(axh_ctx *const *members, int32_t n, axh_batch **out);
int axh_batch_destroy(axh_batch *batch);

/* axh_run for every member: all members must be at step it_begin - 1 and
 * have it_end <= their niter.  seis[b] (seis or seis[b] may be NULL) receives
 * member b's samples in the layout of axh_run. */
int axh_batch_run(axh_batch *batch, int32_t it_begin, int32_t it_end,
                  float *const *seis);

/* Advance every member nsteps from its own last completed step (members may
 * be at different steps, e.g. requests admitted at different times); each
 * must have nsteps steps left.  seis as for axh_batch_run, counted over each
 * member's own steps (axh_seis_samples(member, done + 1, done + nsteps)). */
int axh_batch_step(axh_batch *batch, int32_t nsteps, float *const *seis);
/* axh_batch_step, and every member's kernel-wavefield snapshots of the steps
 * it takes written to kwf[i] as axh_run_kwf writes them (the member's
 * axh_set_kwf and grid; the members of one source type must share the grid,
 * strain_it and ring size).  The moment-tensor runs of an Instaseis database
 * stepped as one batch, each with its own dumps. */
int axh_batch_step_kwf(axh_batch *batch, int32_t nsteps, float *const *seis,
                       float *const *kwf);
int axh_batch_last_run_ms(const axh_batch *batch, double *total_ms);

/* Wavefields advanced by one kernel instance (1..AXH_MAX_GROUP, default
 * AXH_MAX_GROUP): the members are split into groups of at most this many,
 * each group with its wavefields interleaved in device memory.  Wider groups
 * read the mesh terms fewer times, narrower ones need fewer registers per
 * lane.  Only before the first axh_batch_run. */
int axh_batch_set_group_width(axh_batch *batch, int32_t width);
int axh_batch_set_profiling(axh_batch *batch, int on);
/* axh_set_graph_steps for the groups of a batch (one graph per group). */
int axh_batch_set_graph_steps(axh_batch *batch, int32_t steps);
int axh_batch_kernel_stats(const axh_batch *batch, int kclass,
                           double *total_ms, int64_t *launches);

/* ---- resident meshes (request loops) --------------------------------------- */

/* A mesh part on the device, shared by the contexts bound to it.  The
 * reference prepares every run from scratch (SOLVER/main.f90:56-80); a request
 * loop instead keeps meshes resident and binds new requests to them.  No
 * counterpart in the reference. */
typedef struct axh_mesh axh_mesh;

/* 128-bit fingerprint of the mesh part ctx would build (every term the
 * kernels read; not the per-source uploads).  Equal fingerprints: one device
 * mesh serves both contexts.  No GPU. */
int axh_mesh_fingerprint(axh_ctx *ctx, uint64_t fp[2]);

/* 128-bit key of the batches ctx may join: contexts with equal keys (same
 * mesh numbering, time scheme, attenuation mode and time step; any source
 * type) may share a batch, provided members of one source type also have
 * equal fingerprints.  Needs only the dims and the mesh maps.  No GPU. */
int axh_batch_key(axh_ctx *ctx, uint64_t key[2]);

/* Device bytes of ctx's mesh part and of one wavefield on it (fields,
 * memory variables, element forces, source and receiver tables and the
 * fixed-size device staging of the receiver samples).  No GPU.  After the
 * first run of a context, axh_run and the batch runs allocate no device
 * memory. */
int axh_footprint(axh_ctx *ctx, size_t *mesh_bytes, size_t *wave_bytes);

/* A handle on ctx's device mesh part (building ctx's device part first).
 * The mesh stays on the device while a handle or a context uses it. */
int axh_mesh_get(axh_ctx *ctx, axh_mesh **out);

/* Build ctx's device part on mesh instead of uploading its own mesh part;
 * before ctx's first device use, with an equal fingerprint. */
int axh_mesh_bind(axh_ctx *ctx, const axh_mesh *mesh);
int axh_mesh_bytes(const axh_mesh *mesh, size_t *bytes);
int axh_mesh_release(axh_mesh *mesh);

/* ---- the MESHER's mesh database ---------------------------------------------- */

/* Reader of Mesh/meshdb.dat0000, the Fortran unformatted sequential file the
 * reference MESHER writes (MESHER/pdb.f90:2213-2386) and the SOLVER reads
 * (SOLVER/get_mesh.f90:47-409 with read_mesh_basics/advanced/axel of
 * SOLVER/data_mesh.f90:195-328).  It replaces the SOLVER's own reading for
 * hosts that drive the library without the Fortran program (request loops,
 * tests).  Single-slice (nproc_mesh = 1), npol = 4, realkind 4 databases. */
typedef struct axh_meshdb axh_meshdb;
typedef struct axh_meshdb_info {
  int32_t nproc, npol, nelem, npoint;
  int32_t nel_solid, nel_fluid, nglob_solid, nglob_fluid;
  int32_t nel_bdry, ndisc, npoin, nglob;
  int32_t naxel, naxel_solid, naxel_fluid;
  int32_t have_bdry, have_fluid, realkind_bytes;
  double pts_wavelngth, period, courant, deltat, router;
  char bkgrdmodel[64];
} axh_meshdb_info;

/* Read the whole database (AXH_ERR_IO: not a mesh database or truncated;
 * AXH_ERR_UNSUPPORTED: several slices, npol /= 4 or realkind /= 4). */
int axh_meshdb_open(const char *path, axh_meshdb **out);
int axh_meshdb_close(axh_meshdb *db);
int axh_meshdb_get_info(const axh_meshdb *db, axh_meshdb_info *info);

/* Copy one array of the database, nbytes must be its exact size:
 * int32: igloc_solid(25 nel_solid), igloc_fluid(25 nel_fluid), ielsolid,
 * ielfluid, axis_solid/axis_fluid (per element 0/1), ax_el_solid,
 * ax_el_fluid, bdry_solid_el, bdry_fluid_el, bdry_jpol_solid,
 * bdry_jpol_fluid (nel_bdry), lnods(8, nelem), coarsing(nelem);
 * float: G0(5), G1, G1T, G2, G2T (25); double: crd_s, crd_z(npoin);
 * char: eltype(6, nelem). */
int axh_meshdb_array(const axh_meshdb *db, const char *name, void *out, size_t nbytes);

/* The mesh-only uploads of ctx from the database (axh_upload_spectral and
 * axh_upload_mesh_maps); ctx's dims must match the database's counts. */
int axh_meshdb_upload(const axh_meshdb *db, axh_ctx *ctx);

/* ---- state files --------------------------------------------------------- */

/* Write every upload of ctx (plus dims) to a self-describing binary file, and
 * rebuild a context from one.  Used to drive the library without the Fortran
 * host (tests, benchmarks, the request loop). */
int axh_save_state(const axh_ctx *ctx, const char *path);
int axh_load_state(const char *path, axh_ctx **out);

/* ---- kernel wavefields (Instaseis databases, dump_type displ_only) ------- */

/* The arrays the fluid displacement of the dumps needs, each
 * (0:npol,0:npol,nel_fluid): inv_rho_fluid and the fluid's pointwise
 * derivative terms over J (SOLVER/def_precomp_terms.f90:148-304,631), read by
 * dump_disp_global through axisym_gradient_fluid
 * (SOLVER/wavefields_io.f90:1082-1090, pointwise_derivatives.f90:509-544).
 * An upload (before the first run). */
int axh_upload_fluid_gradient(axh_ctx *ctx, const float *inv_rho_fluid,
                              const float *DsDeta_over_J_flu, const float *DzDeta_over_J_flu,
                              const float *DsDxi_over_J_flu, const float *DzDxi_over_J_flu);

/* build_kwf_grid (SOLVER/meshes_io.F90:489-610): the kernel-wavefield points
 * are the GLL points ibeg..iend x jbeg..jend (0..4) of the selected elements
 * plus the axis point ipol = 0 of every row of an axial element, numbered
 * 1, 2, ... by the first appearance of their global point, the solid elements
 * first (jpol outer, ipol inner, the axis point after each row), then the
 * fluid elements; the solid points come first.  elem_mask holds
 * nel_solid + nel_fluid flags, solid elements first (the reference's
 * mask_tp_elem, which selects by the radius and colatitude of the element
 * corners); NULL selects every element, as the reference's defaults
 * (KERNEL_RMIN 0, KERNEL_RMAX 7000 km, colatitudes 0..180) do.  Returns the
 * point counts npoint_solid_kwf and npoint_fluid_kwf.  An upload. */
int axh_kwf_grid(axh_ctx *ctx, int32_t ibeg, int32_t iend, int32_t jbeg, int32_t jend,
                 const int32_t *elem_mask, int32_t *npoint_solid, int32_t *npoint_fluid);

/* The grid as the reference's build_kwf_grid leaves it in module data_mesh:
 * mapping_ijel_ikwf and kwf_mask (0 / 1), both (0:npol,0:npol,nel_solid +
 * nel_fluid) with the solid elements first, and npoint_solid_kwf,
 * npoint_fluid_kwf.  Checked: the masked entries number every point exactly
 * once, solid points 1..npoint_solid, fluid points after them; unmasked
 * entries are not read.  An upload. */
int axh_upload_kwf_grid(axh_ctx *ctx, int32_t npoint_solid, int32_t npoint_fluid,
                        const int32_t *mapping_ijel_ikwf, const int32_t *kwf_mask);

/* The context's grid (either of the two above): the point counts, and, for
 * non-NULL arrays, mapping_ijel_ikwf (0 at unmasked entries) and kwf_mask. */
int axh_kwf_get_grid(const axh_ctx *ctx, int32_t *mapping_ijel_ikwf, int32_t *kwf_mask,
                     int32_t *npoint_solid, int32_t *npoint_fluid);

/* Snapshot interval of axh_run_kwf: the steps iter with mod(iter, strain_it)
 * == 0, including the initial fields (iter 0), as dump_stuff
 * (SOLVER/time_evol_wave.F90:1200-1217) calls dump_disp_global; src_mask 1
 * zeroes the solid displacement of the source elements (src_dump_type 'mask',
 * the reference's default; eradicate_src_elem_vec_values,
 * SOLVER/wavefields_io.f90:1123-1135), 0 keeps it ('none').  ring_snapshots
 * sets the device ring the snapshots are gathered in (at least 2; 0: chosen
 * by the library, at most AXH_KWF_RING_BYTES); they are copied out half a
 * ring (at most 16 MiB, at least one snapshot) at a time.  Any time. */
#define AXH_KWF_RING_BYTES ((size_t)1 << 30)
int axh_set_kwf(axh_ctx *ctx, int32_t strain_it, int32_t src_mask, int32_t ring_snapshots);

/* Number of snapshots axh_run_kwf(ctx, it_begin, it_end, ...) writes. */
int axh_kwf_snapshots(const axh_ctx *ctx, int32_t it_begin, int32_t it_end);

/* axh_run, and the kernel-wavefield snapshots of the steps it covers written
 * to kwf: per snapshot, per component, npoint_solid + npoint_fluid floats
 * (point k + 1 of the grid at index k), components s, z for a monopole source
 * and s, phi, z otherwise; the values dump_disp_global gathers with
 * kwf_mapping_sol / kwf_mapping_flu (SOLVER/wavefields_io.f90:743-781,
 * 1040-1097) for the NetCDF variables disp_sol_* / disp_flu_* of a snapshot
 * (SOLVER/nc_routines.F90:945-970): the solid displacement (dipole: u1 + u2,
 * u1 - u2, u3), and in the fluid inv_rho_fluid times the gradient of chi
 * (phi: 0).  The snapshots are gathered on the GPU into a device ring and
 * copied to host memory through pinned staging buffers while the following
 * steps run; kwf is complete when the call returns.  Batches gather them
 * with axh_batch_step_kwf. */
int axh_run_kwf(axh_ctx *ctx, int32_t it_begin, int32_t it_end, float *seis, float *kwf);
/* The same queued (as axh_run_begin): axh_run_end completes it; until then the
 * host is free (e.g. to write the previous chunk's snapshots) and seis and kwf
 * must stay untouched. */
int axh_run_kwf_begin(axh_ctx *ctx, int32_t it_begin, int32_t it_end, float *seis,
                      float *kwf);

/* ---- seismogram post-processing ------------------------------------------ */

/* The convolution of traces with a source time function of the reference's
 * post-processing, convolve_with_stf (SOLVER/UTILS/post_processing.F90:
 * 1014-1084), on the GPU: seis and out are ntraces traces of nt samples, one
 * after the other; for sample i (1-based) of a trace,
 *   out(i) = sum over j = 1 .. min(nj, i-1) of seis(i-j) * stf(j) * dt,
 * each term the REAL(4) product seis*stf times dt in REAL(8), added in
 * REAL(8) to the REAL(4) partial sum in ascending j, and finally
 * out(i) = out(i) * scale in REAL(8): the reference's arithmetic bit for
 * bit.  stf(1..nj) is the tabulated source time function (at most
 * AXH_MAX_STF_TERMS terms).  Allocates its device buffers per call. */
#define AXH_MAX_STF_TERMS 8192
int axh_convolve_traces(int32_t nt, int32_t ntraces, const float *seis, int32_t nj,
                        const float *stf, double dt, double scale, float *out);

/* ---- helpers ------------------------------------------------------------- */

/* coefd(nstages+1), coefv(nstages), coeff(nstages) scaled by deltat, as
 * symplectic_coefficients (SOLVER/time_evol_wave.F90:749-967) computes them.
 * Arrays must hold at least 8 entries; newmark2 returns nstages = 1 and
 * fills nothing. */
int axh_symplectic_coefficients(int32_t scheme, double deltat, double *coefd,
                                double *coefv, double *coeff,
                                int32_t *nstages);

#ifdef __cplusplus
}
#endif
#endif /* AXIHIP_H */
