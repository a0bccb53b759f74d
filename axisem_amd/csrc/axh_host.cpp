// axh_host.cpp -- C-ABI entry points, uploads and state files of the GPU
// time-loop library.  Nothing here touches the GPU; device work starts in
// axh_device.hip on the first axh_run.
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "axh_internal.h"

namespace axh {

const char *const kSolid3dNames[AXH_N_SOLID_3D] = {
    "M11s", "M21s", "M41s", "M12s", "M22s", "M42s", "M11z", "M21z",
    "M41z", "M32s", "M13s", "M33s", "M43s", "M1phi", "M2phi", "M4phi",
    "M_1",  "M_2",  "M_3",  "M_4",  "M_5",  "M_6",  "M_7",  "M_8",
    "M_w1", "M_w2", "M_w3", "M_w4", "M_w5"};
const char *const kFluid3dNames[4] = {"M1chi_fl", "M2chi_fl", "M4chi_fl",
                                      "M_w_fl"};
const char *const kFluidGradNames[5] = {"inv_rho_fluid", "DsDeta_over_J_flu",
                                        "DzDeta_over_J_flu", "DsDxi_over_J_flu",
                                        "DzDxi_over_J_flu"};

[[noreturn]] void fail(int code, const std::string &msg) {
  throw Error{code, msg};
}

// symplectic_coefficients (SOLVER/time_evol_wave.F90:749-967).  The PEFRL and
// McLachlan-6 constants are written in the reference as default-real (single
// precision) literals assigned to double variables, so they are rounded to
// float first exactly as the reference build does.
int symplectic_coefficients(int scheme, double dt, double *cd, double *cv,
                            double *cf) {
  int n = 0;
  if (scheme == AXH_SYMPLEC4) {
    const double z = (double)0.1786178958448091f;
    const double io = (double)-0.2123418310626054f;
    const double k = (double)-0.06626458266981849f;
    n = 4;
    cd[0] = z; cd[1] = k; cd[2] = 1.0 - 2.0 * (z + k); cd[3] = k; cd[4] = z;
    cv[0] = 0.5 - io; cv[1] = io; cv[2] = io; cv[3] = 0.5 - io;
  } else if (scheme == AXH_ML_SO4M5) {
    const double rho = (14.0 - std::sqrt(19.0)) / 108.0;
    const double th = (20.0 - 7.0 * std::sqrt(19.0)) / 108.0;
    const double nu = 2.0 / 5.0, lam = -1.0 / 10.0;
    n = 5;
    cd[0] = rho; cd[1] = th; cd[2] = 0.5 - rho - th; cd[3] = 0.5 - rho - th;
    cd[4] = th; cd[5] = rho;
    cv[0] = nu; cv[1] = lam; cv[2] = 1.0 - 2.0 * (nu + lam); cv[3] = lam;
    cv[4] = nu;
  } else if (scheme == AXH_ML_SO6M7) {
    n = 7;
    cd[0] = (double)-1.01308797891717472981f;
    cd[1] = (double)1.18742957373254270702f;
    cd[2] = (double)-0.01833585209646059034f;
    cd[3] = (double)0.34399425728109261313f;
    for (int i = 5; i <= 8; i++) cd[i - 1] = cd[n + 2 - i - 1];
    cv[0] = (double)0.00016600692650009894f;
    cv[1] = (double)-0.37962421426377360608f;
    cv[2] = (double)0.68913741185181063674f;
    cv[3] = (double)0.38064159097092574080f;
    for (int i = 5; i <= 7; i++) cv[i - 1] = cv[n + 1 - i - 1];
  } else {
    return 0;
  }
  for (int i = 0; i <= n; i++) cd[i] = cd[i] * dt;
  for (int i = 0; i < n; i++) cv[i] = cv[i] * dt;
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int j = 0; j <= i; j++) s = s + cd[j];
    cf[i] = s;
  }
  return n;
}

}  // namespace axh

using namespace axh;

static_assert(offsetof(axh_dims, deltat) == 88, "axh_dims layout");
static_assert(sizeof(axh_dims) == 104, "axh_dims layout");

static thread_local std::string g_last_error;

namespace axh {
int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}
}  // namespace axh

const axh::Array &axh_ctx::need(const std::string &name) const {
  auto *a = get(name);
  if (!a) fail(AXH_ERR_STATE, "missing upload: " + name);
  return *a;
}

void axh_ctx::put(const std::string &name, DType t, std::vector<uint64_t> dims_,
                  const void *src) {
  if (dev) fail(AXH_ERR_STATE, "uploads are closed after the first run: " + name);
  if (batch_refs > 0) fail(AXH_ERR_STATE, "uploads are closed while in a batch: " + name);
  fp_valid = false;
  Array a;
  a.dtype = t;
  a.dims = std::move(dims_);
  size_t el = (t == kF64) ? 8 : 4;
  a.bytes.resize(a.count() * el);
  if (!a.bytes.empty()) {
    if (!src) fail(AXH_ERR_ARG, "null pointer for " + name);
    std::memcpy(a.bytes.data(), src, a.bytes.size());
  }
  arrays[name] = std::move(a);
}

static void check_ctx(const axh_ctx *ctx) {
  if (!ctx) fail(AXH_ERR_ARG, "null context");
}

static void validate_dims(const axh_dims &d) {
  if (d.abi_version != AXH_ABI_VERSION) fail(AXH_ERR_ARG, "ABI version mismatch");
  if (d.npol != kNpol) fail(AXH_ERR_UNSUPPORTED, "only npol = 4 is supported");
  if (d.nel_solid <= 0 || d.nel_fluid < 0) fail(AXH_ERR_ARG, "bad element counts");
  if (d.nglob_solid <= 0 || (d.nel_fluid > 0 && d.nglob_fluid <= 0))
    fail(AXH_ERR_ARG, "bad global point counts");
  if (d.nel_bdry < 0 || (d.nel_bdry > 0 && d.nel_fluid == 0))
    fail(AXH_ERR_ARG, "bad nel_bdry");
  if (d.src_type < AXH_MONOPOLE || d.src_type > AXH_QUADPOLE)
    fail(AXH_ERR_ARG, "bad src_type");
  if (d.scheme < AXH_NEWMARK2 || d.scheme > AXH_ML_SO6M7)
    fail(AXH_ERR_UNSUPPORTED, "unsupported time scheme");
  if (d.anel) {
    if (d.n_sls < 1 || d.n_sls > kMaxSls)
      fail(AXH_ERR_UNSUPPORTED, "n_sls must be in 1..8");
  }
  if (d.niter < 0 || d.seis_it < 1 || d.num_rec < 0) fail(AXH_ERR_ARG, "bad run counters");
  if (d.nelsrc < 0 || d.nelsrc > 8) fail(AXH_ERR_ARG, "nelsrc must be <= 8");
  if (d.fluid_src && d.scheme != AXH_NEWMARK2)
    fail(AXH_ERR_UNSUPPORTED, "fluid sources are only defined for newmark2");
  if (!(d.deltat > 0.0)) fail(AXH_ERR_ARG, "deltat must be positive");
}

extern "C" {

const char *axh_last_error(void) { return g_last_error.c_str(); }

int axh_create(const axh_dims *dims, axh_ctx **out) {
  return guarded([&] {
    if (!dims || !out) fail(AXH_ERR_ARG, "null argument");
    *out = nullptr;
    validate_dims(*dims);
    auto *c = new axh_ctx;
    c->dims = *dims;
    *out = c;
  });
}

int axh_destroy(axh_ctx *ctx) {
  return guarded([&] {
    if (!ctx) return;
    if (ctx->batch_refs > 0) fail(AXH_ERR_STATE, "context is held by a batch");
    delete ctx;
  });
}

int axh_upload_spectral(axh_ctx *ctx, const float *G0, const float *G1,
                        const float *G1T, const float *G2, const float *G2T) {
  return guarded([&] {
    check_ctx(ctx);
    ctx->put("G0", kF32, {5}, G0);
    ctx->put("G1", kF32, {5, 5}, G1);
    ctx->put("G1T", kF32, {5, 5}, G1T);
    ctx->put("G2", kF32, {5, 5}, G2);
    ctx->put("G2T", kF32, {5, 5}, G2T);
  });
}

int axh_upload_solid_term(axh_ctx *ctx, int term, const float *data) {
  return guarded([&] {
    check_ctx(ctx);
    const uint64_t nel = (uint64_t)ctx->dims.nel_solid;
    if (term >= 0 && term < AXH_N_SOLID_3D) {
      ctx->put(kSolid3dNames[term], kF32, {5, 5, nel}, data);
    } else if (term >= AXH_M0_W1 && term <= AXH_M0_W10) {
      std::string nm = "M0_w" + std::to_string(term - AXH_M0_W1 + 1);
      ctx->put(nm, kF32, {5, nel}, data);
    } else {
      fail(AXH_ERR_ARG, "unknown solid term " + std::to_string(term));
    }
  });
}

int axh_upload_fluid_term(axh_ctx *ctx, int term, const float *data) {
  return guarded([&] {
    check_ctx(ctx);
    const uint64_t nel = (uint64_t)ctx->dims.nel_fluid;
    if (term >= AXH_M1CHI_FL && term <= AXH_M_W_FL)
      ctx->put(kFluid3dNames[term], kF32, {5, 5, nel}, data);
    else if (term == AXH_M0_W_FL)
      ctx->put("M0_w_fl", kF32, {5, nel}, data);
    else
      fail(AXH_ERR_ARG, "unknown fluid term " + std::to_string(term));
  });
}

int axh_upload_mass(axh_ctx *ctx, const float *inv_mass_rho,
                    const float *inv_mass_fluid) {
  return guarded([&] {
    check_ctx(ctx);
    ctx->put("inv_mass_rho", kF32, {5, 5, (uint64_t)ctx->dims.nel_solid}, inv_mass_rho);
    ctx->put("inv_mass_fluid", kF32, {5, 5, (uint64_t)ctx->dims.nel_fluid},
             inv_mass_fluid);
  });
}

int axh_upload_mesh_maps(axh_ctx *ctx, const int32_t *igloc_solid,
                         const int32_t *igloc_fluid, const int32_t *axis_solid,
                         const int32_t *axis_fluid) {
  return guarded([&] {
    check_ctx(ctx);
    const auto &d = ctx->dims;
    ctx->put("igloc_solid", kI32, {25 * (uint64_t)d.nel_solid}, igloc_solid);
    ctx->put("igloc_fluid", kI32, {25 * (uint64_t)d.nel_fluid}, igloc_fluid);
    ctx->put("axis_solid", kI32, {(uint64_t)d.nel_solid}, axis_solid);
    ctx->put("axis_fluid", kI32, {(uint64_t)d.nel_fluid}, axis_fluid);
  });
}

int axh_upload_boundary(axh_ctx *ctx, const int32_t *bdry_solid_el,
                        const int32_t *bdry_fluid_el,
                        const int32_t *bdry_jpol_solid,
                        const int32_t *bdry_jpol_fluid, const float *bdry_matr) {
  return guarded([&] {
    check_ctx(ctx);
    const uint64_t nb = (uint64_t)ctx->dims.nel_bdry;
    ctx->put("bdry_solid_el", kI32, {nb}, bdry_solid_el);
    ctx->put("bdry_fluid_el", kI32, {nb}, bdry_fluid_el);
    ctx->put("bdry_jpol_solid", kI32, {nb}, bdry_jpol_solid);
    ctx->put("bdry_jpol_fluid", kI32, {nb}, bdry_jpol_fluid);
    ctx->put("bdry_matr", kF32, {5, nb, 2}, bdry_matr);
  });
}

int axh_upload_fluid_surface_mask(axh_ctx *ctx, const float *mask) {
  return guarded([&] {
    check_ctx(ctx);
    ctx->put("fluid_free_surface_mask", kF32, {5, 5, (uint64_t)ctx->dims.nel_fluid},
             mask);
  });
}

int axh_upload_absorbing(axh_ctx *ctx, const float *gamma_solid, const float *gamma_fluid) {
  return guarded([&] {
    check_ctx(ctx);
    if (!gamma_solid || (ctx->dims.nel_fluid > 0 && !gamma_fluid))
      fail(AXH_ERR_ARG, "absorbing boundaries need the solid and fluid damping factors");
    ctx->put("solid_absorbing_gamma", kF32, {5, 5, (uint64_t)ctx->dims.nel_solid},
             gamma_solid);
    if (ctx->dims.nel_fluid > 0)
      ctx->put("fluid_absorbing_gamma", kF32, {5, 5, (uint64_t)ctx->dims.nel_fluid},
               gamma_fluid);
  });
}

int axh_upload_source(axh_ctx *ctx, const int32_t *ielsrc,
                      const float *source_term, const float *stf, int32_t nstf) {
  return guarded([&] {
    check_ctx(ctx);
    const auto &d = ctx->dims;
    int nst = 1;
    if (d.scheme != AXH_NEWMARK2) {
      double cd[40], cv[40], cf[40];
      nst = symplectic_coefficients(d.scheme, d.deltat, cd, cv, cf);
    }
    if (nstf != d.niter * nst)
      fail(AXH_ERR_ARG, "stf table must hold niter*nstages = " +
                            std::to_string(d.niter * nst) + " values");
    ctx->put("ielsrc", kI32, {(uint64_t)std::max(d.nelsrc, 0)}, ielsrc);
    if (d.fluid_src)
      ctx->put("source_term", kF32, {5, 5, 8}, source_term);
    else
      ctx->put("source_term", kF32, {5, 5, 8, 3}, source_term);
    ctx->put("stf", kF32, {(uint64_t)nstf}, stf);
  });
}

int axh_upload_attenuation_cg4(axh_ctx *ctx, const axh_attenuation_cg4 *a) {
  return guarded([&] {
    check_ctx(ctx);
    if (!a) fail(AXH_ERR_ARG, "null attenuation terms");
    const auto &d = ctx->dims;
    if (!d.anel) fail(AXH_ERR_STATE, "attenuation upload for an elastic run");
    if (!d.coarse_grained)
      fail(AXH_ERR_STATE, "coarse-grained terms for a run with memory variables at all points");
    const uint64_t nel = (uint64_t)d.nel_solid, ns = (uint64_t)d.n_sls;
    ctx->put("Y_cg4", kF32, {4, nel}, a->Y);
    ctx->put("V_s_eta_cg4", kF32, {4, nel}, a->V_s_eta);
    ctx->put("V_s_xi_cg4", kF32, {4, nel}, a->V_s_xi);
    ctx->put("V_z_eta_cg4", kF32, {4, nel}, a->V_z_eta);
    ctx->put("V_z_xi_cg4", kF32, {4, nel}, a->V_z_xi);
    ctx->put("DsDeta_over_J_sol_cg4", kF32, {4, nel}, a->DsDeta_over_J);
    ctx->put("DzDeta_over_J_sol_cg4", kF32, {4, nel}, a->DzDeta_over_J);
    ctx->put("DsDxi_over_J_sol_cg4", kF32, {4, nel}, a->DsDxi_over_J);
    ctx->put("DzDxi_over_J_sol_cg4", kF32, {4, nel}, a->DzDxi_over_J);
    ctx->put("inv_s_solid_cg4", kF32, {4, nel}, a->inv_s);
    ctx->put("delta_mu_cg4", kF32, {4, nel}, a->delta_mu);
    ctx->put("delta_kappa_cg4", kF32, {4, nel}, a->delta_kappa);
    ctx->put("exp_w_j_deltat", kF64, {ns}, a->exp_w_j_deltat);
    ctx->put("ts_fac_t", kF64, {ns}, a->ts_fac_t);
    ctx->put("ts_fac_tm1", kF64, {ns}, a->ts_fac_tm1);
    ctx->put("a_j_mu", kF64, {ns, nel}, a->a_j_mu);
    ctx->put("a_j_kappa", kF64, {ns, nel}, a->a_j_kappa);
  });
}

int axh_upload_attenuation_full(axh_ctx *ctx, const axh_attenuation_full *a) {
  return guarded([&] {
    check_ctx(ctx);
    if (!a) fail(AXH_ERR_ARG, "null attenuation terms");
    const auto &d = ctx->dims;
    if (!d.anel) fail(AXH_ERR_STATE, "attenuation upload for an elastic run");
    if (d.coarse_grained) fail(AXH_ERR_STATE, "full attenuation terms for a coarse-grained run");
    const uint64_t nel = (uint64_t)d.nel_solid, ns = (uint64_t)d.n_sls;
    ctx->put("Y", kF32, {5, 5, nel}, a->Y);
    ctx->put("V_s_eta", kF32, {5, 5, nel}, a->V_s_eta);
    ctx->put("V_s_xi", kF32, {5, 5, nel}, a->V_s_xi);
    ctx->put("V_z_eta", kF32, {5, 5, nel}, a->V_z_eta);
    ctx->put("V_z_xi", kF32, {5, 5, nel}, a->V_z_xi);
    ctx->put("Y0", kF32, {5, nel}, a->Y0);
    ctx->put("V0_s_eta", kF32, {5, nel}, a->V0_s_eta);
    ctx->put("V0_s_xi", kF32, {5, nel}, a->V0_s_xi);
    ctx->put("V0_z_eta", kF32, {5, nel}, a->V0_z_eta);
    ctx->put("V0_z_xi", kF32, {5, nel}, a->V0_z_xi);
    ctx->put("DsDeta_over_J_sol", kF32, {5, 5, nel}, a->DsDeta_over_J);
    ctx->put("DzDeta_over_J_sol", kF32, {5, 5, nel}, a->DzDeta_over_J);
    ctx->put("DsDxi_over_J_sol", kF32, {5, 5, nel}, a->DsDxi_over_J);
    ctx->put("DzDxi_over_J_sol", kF32, {5, 5, nel}, a->DzDxi_over_J);
    ctx->put("inv_s_solid", kF32, {5, 5, nel}, a->inv_s);
    ctx->put("delta_mu", kF32, {5, 5, nel}, a->delta_mu);
    ctx->put("delta_kappa", kF32, {5, 5, nel}, a->delta_kappa);
    ctx->put("exp_w_j_deltat", kF64, {ns}, a->exp_w_j_deltat);
    ctx->put("ts_fac_t", kF64, {ns}, a->ts_fac_t);
    ctx->put("ts_fac_tm1", kF64, {ns}, a->ts_fac_tm1);
    ctx->put("a_j_mu", kF64, {ns, nel}, a->a_j_mu);
    ctx->put("a_j_kappa", kF64, {ns, nel}, a->a_j_kappa);
  });
}

int axh_upload_receivers(axh_ctx *ctx, const int32_t *recfile_el) {
  return guarded([&] {
    check_ctx(ctx);
    ctx->put("recfile_el", kI32, {(uint64_t)ctx->dims.num_rec, 3}, recfile_el);
  });
}

int axh_seis_samples(const axh_ctx *ctx, int32_t it_begin, int32_t it_end) {
  if (!ctx || it_end < it_begin || ctx->dims.num_rec == 0) return 0;
  const int s = ctx->dims.seis_it;
  return it_end / s - (it_begin - 1) / s;
}

// the context has no run queued by axh_run_begin
static void check_idle(const axh_ctx *ctx) {
  if (ctx->running) fail(AXH_ERR_STATE, "a run of the context is queued: call axh_run_end");
}

static void check_range(const axh_ctx *ctx, int32_t it_begin, int32_t it_end, const char *fn) {
  if (it_begin != ctx->iter_done + 1)
    fail(AXH_ERR_ARG, std::string(fn) + " must continue at step " +
                          std::to_string(ctx->iter_done + 1));
  if (it_end < it_begin - 1 || it_end > ctx->dims.niter)
    fail(AXH_ERR_ARG, "step range outside 1..niter");
}

int axh_run(axh_ctx *ctx, int32_t it_begin, int32_t it_end, float *seis) {
  return guarded([&] {
    check_ctx(ctx);
    check_idle(ctx);
    check_range(ctx, it_begin, it_end, "axh_run");
    if (!ctx->dev) device_build(ctx);
    device_run(ctx, it_begin, it_end, seis);
  });
}

int axh_run_begin(axh_ctx *ctx, int32_t it_begin, int32_t it_end, float *seis) {
  return guarded([&] {
    check_ctx(ctx);
    check_idle(ctx);
    check_range(ctx, it_begin, it_end, "axh_run_begin");
    if (!ctx->dev) device_build(ctx);
    device_run_begin(ctx, it_begin, it_end, seis);
    ctx->running = true;
  });
}

int axh_run_end(axh_ctx *ctx) {
  return guarded([&] {
    check_ctx(ctx);
    if (!ctx->running) fail(AXH_ERR_STATE, "no run of the context is queued");
    ctx->running = false;
    device_run_end(ctx);
  });
}

int axh_run_noalloc(axh_ctx *ctx, int32_t it_begin, int32_t it_end, float *seis) {
  return guarded([&] {
    check_ctx(ctx);
    check_idle(ctx);
    if (ctx->dims.num_rec > noalloc_max_receivers())
      fail(AXH_ERR_ARG, "axh_run_noalloc: " + std::to_string(ctx->dims.num_rec) +
                            " receivers; the fixed scratch array holds the samples of at most " +
                            std::to_string(noalloc_max_receivers()));
    if (it_begin != ctx->iter_done + 1)
      fail(AXH_ERR_ARG, "axh_run_noalloc must continue at step " +
                            std::to_string(ctx->iter_done + 1));
    if (it_end < it_begin - 1 || it_end > ctx->dims.niter)
      fail(AXH_ERR_ARG, "step range outside 1..niter");
    if (!ctx->dev)
      fail(AXH_ERR_STATE, "axh_run_noalloc needs a context that has run (axh_run) before");
    device_run_noalloc(ctx, it_begin, it_end, seis);
  });
}

int axh_download_fields(axh_ctx *ctx, float *disp, float *velo, float *chi,
                        float *dchi, float *ddchi) {
  return guarded([&] {
    check_ctx(ctx);
    check_idle(ctx);
    if (!ctx->dev) device_build(ctx);
    device_download(ctx, disp, velo, chi, dchi, ddchi);
  });
}

int axh_upload_fields(axh_ctx *ctx, const float *disp, const float *velo,
                      const float *chi, const float *dchi, const float *ddchi) {
  return guarded([&] {
    check_ctx(ctx);
    check_idle(ctx);
    if (!ctx->dev) device_build(ctx);
    device_upload_fields(ctx, disp, velo, chi, dchi, ddchi);
  });
}

int axh_last_run_ms(const axh_ctx *ctx, double *total_ms) {
  return guarded([&] {
    check_ctx(ctx);
    if (!total_ms) fail(AXH_ERR_ARG, "null pointer");
    *total_ms = ctx->last_run_ms;
  });
}

int axh_check_inputs(axh_ctx *ctx) {
  return guarded([&] {
    check_ctx(ctx);
    check_inputs(ctx);
  });
}

int axh_global_numbering(axh_ctx *ctx, int32_t domain, int32_t *gid, int32_t *nglob) {
  return guarded([&] {
    check_ctx(ctx);
    if (!gid || !nglob || (domain != 0 && domain != 1)) fail(AXH_ERR_ARG, "bad argument");
    if (domain == 1 && ctx->dims.nel_fluid == 0) fail(AXH_ERR_ARG, "no fluid");
    global_numbering(ctx, domain, gid, nglob);
  });
}

int axh_set_profiling(axh_ctx *ctx, int on) {
  return guarded([&] {
    check_ctx(ctx);
    ctx->profile = on != 0;
  });
}

int axh_set_graph_steps(axh_ctx *ctx, int32_t steps) {
  return guarded([&] {
    check_ctx(ctx);
    if (steps < 0 || steps > AXH_MAX_GRAPH_STEPS)
      fail(AXH_ERR_ARG, "graph steps must be 0.." + std::to_string(AXH_MAX_GRAPH_STEPS));
    ctx->graph_steps = steps;
  });
}

const char *axh_kernel_name(int kclass) { return kernel_class_name(kclass); }

int axh_kernel_stats(const axh_ctx *ctx, int kclass, double *total_ms,
                     int64_t *launches) {
  return guarded([&] {
    check_ctx(ctx);
    if (!total_ms || !launches) fail(AXH_ERR_ARG, "null pointer");
    if (!kernel_stats(ctx->dev.get(), kclass, total_ms, launches))
      fail(AXH_ERR_ARG, "unknown kernel class");
  });
}

int axh_device_bytes(axh_ctx *ctx, size_t *bytes) {
  return guarded([&] {
    check_ctx(ctx);
    if (!bytes) fail(AXH_ERR_ARG, "null pointer");
    if (!ctx->dev) device_build(ctx);
    *bytes = device_bytes(ctx->dev.get());
  });
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
}  // extern "C"

// Members of a batch share one mesh and one time scheme: equal dims (the
// source type aside) and equal global numbering.  Members of one source type
// step on one device mesh part, so for them the mesh part the device would
// build (every term the kernels read; not the per-source uploads ielsrc,
// source_term, stf, recfile_el) must agree.  The reference leaves entries it
// never reads unset (e.g. M0_w_fl of non-axial elements,
// SOLVER/def_precomp_terms.f90:2355-2455), so that comparison is on the built
// layout, not on the raw uploads.  Members of different source types (the
// monopole, dipole and quadrupole runs of a moment tensor,
// SOLVER/source.f90:40-140) each use their type's mesh part.
static void check_same_mesh(const axh_ctx *a, axh_ctx *b, int idx) {
  const std::string who = "batch member " + std::to_string(idx) + ": ";
  const axh_dims &x = a->dims, &y = b->dims;
  if (x.abi_version != y.abi_version || x.npol != y.npol || x.nel_solid != y.nel_solid ||
      x.nel_fluid != y.nel_fluid || x.nglob_solid != y.nglob_solid ||
      x.nglob_fluid != y.nglob_fluid || x.nel_bdry != y.nel_bdry ||
      x.scheme != y.scheme || x.anel != y.anel || x.coarse_grained != y.coarse_grained ||
      x.n_sls != y.n_sls || std::memcmp(&x.deltat, &y.deltat, sizeof(double)) != 0)
    fail(AXH_ERR_ARG, who + "mesh or time scheme differs from member 0");
  for (const char *nm : {"igloc_solid", "igloc_fluid"}) {
    const Array *p = a->get(nm), *q = b->get(nm);
    if (!p || !q) fail(AXH_ERR_STATE, who + "missing upload " + nm);
    if (p->bytes.size() != q->bytes.size() ||
        std::memcmp(p->bytes.data(), q->bytes.data(), p->bytes.size()) != 0)
      fail(AXH_ERR_ARG, who + "global numbering differs from member 0");
  }
}

static void check_same_terms(axh_ctx *a, axh_ctx *b, int idx) {
  uint64_t fa[2], fb[2];
  mesh_fingerprint(a, fa);
  mesh_fingerprint(b, fb);
  if (fa[0] != fb[0] || fa[1] != fb[1])
    fail(AXH_ERR_ARG, "batch member " + std::to_string(idx) +
                          ": mesh terms differ from an earlier member of its source type");
}

static void check_batch(const axh_batch *b) {
  if (!b) fail(AXH_ERR_ARG, "null batch");
}

extern "C" {

int axh_batch_create(axh_ctx *const *members, int32_t n, axh_batch **out) {
  return guarded([&] {
    if (!members || !out) fail(AXH_ERR_ARG, "null argument");
    *out = nullptr;
    if (n < 1 || n > AXH_MAX_BATCH)
      fail(AXH_ERR_ARG, "batch size must be 1.." + std::to_string(AXH_MAX_BATCH));
    for (int i = 0; i < n; i++) {
      if (!members[i]) fail(AXH_ERR_ARG, "null batch member");
      if (members[i]->batch_refs > 0) fail(AXH_ERR_STATE, "a context is already in a batch");
      check_idle(members[i]);
      for (int j = 0; j < i; j++)
        if (members[j] == members[i]) fail(AXH_ERR_ARG, "a context appears twice in a batch");
    }
    for (int i = 1; i < n; i++) {
      check_same_mesh(members[0], members[i], i);
      for (int j = 0; j < i; j++)
        if (members[j]->dims.src_type == members[i]->dims.src_type) {
          check_same_terms(members[j], members[i], i);
          break;
        }
    }
    auto *b = new axh_batch;
    b->members.assign(members, members + n);
    for (auto *c : b->members) c->batch_refs++;
    *out = b;
  });
}

int axh_batch_destroy(axh_batch *batch) {
  return guarded([&] {
    if (!batch) return;
    // the members keep their state: it moves back to them
    std::unique_ptr<axh_batch> owned(batch);
    for (auto *c : batch->members) c->batch_refs--;
    batch_detach(batch);
  });
}

int axh_batch_run(axh_batch *batch, int32_t it_begin, int32_t it_end, float *const *seis) {
  return guarded([&] {
    check_batch(batch);
    for (auto *c : batch->members) {
      if (it_begin != c->iter_done + 1)
        fail(AXH_ERR_ARG, "axh_batch_run must continue at step " +
                              std::to_string(c->iter_done + 1));
      if (it_end < it_begin - 1 || it_end > c->dims.niter)
        fail(AXH_ERR_ARG, "step range outside 1..niter of a member");
    }
    batch_attach(batch);
    batch_run(batch, it_end - it_begin + 1, seis);
  });
}

int axh_batch_step(axh_batch *batch, int32_t nsteps, float *const *seis) {
  return guarded([&] {
    check_batch(batch);
    if (nsteps < 0) fail(AXH_ERR_ARG, "negative step count");
    for (auto *c : batch->members)
      if (c->iter_done + nsteps > c->dims.niter)
        fail(AXH_ERR_ARG, "a member has fewer than nsteps steps left");
    batch_attach(batch);
    batch_run(batch, nsteps, seis);
  });
}

int axh_batch_step_kwf(axh_batch *batch, int32_t nsteps, float *const *seis,
                       float *const *kwf) {
  return guarded([&] {
    check_batch(batch);
    if (nsteps < 0) fail(AXH_ERR_ARG, "negative step count");
    if (!kwf) fail(AXH_ERR_ARG, "null kwf array");
    for (size_t i = 0; i < batch->members.size(); i++) {
      axh_ctx *c = batch->members[i];
      if (c->iter_done + nsteps > c->dims.niter)
        fail(AXH_ERR_ARG, "a member has fewer than nsteps steps left");
      if (c->kwf_strain_it < 1)
        fail(AXH_ERR_STATE, "batch member " + std::to_string(i) + " has no axh_set_kwf");
      const int32_t *np = c->need("kwf_npoints").as<int32_t>();
      if (!kwf[i] && np[0] + np[1] > 0 &&
          axh_kwf_snapshots(c, c->iter_done + 1, c->iter_done + nsteps) > 0)
        fail(AXH_ERR_ARG, "null kwf buffer of batch member " + std::to_string(i));
    }
    batch_attach(batch);
    batch_run(batch, nsteps, seis, kwf);
  });
}

int axh_batch_last_run_ms(const axh_batch *batch, double *total_ms) {
  return guarded([&] {
    check_batch(batch);
    if (!total_ms) fail(AXH_ERR_ARG, "null pointer");
    *total_ms = batch->last_run_ms;
  });
}

int axh_batch_set_group_width(axh_batch *batch, int32_t width) {
  return guarded([&] {
    check_batch(batch);
    if (width < 1 || width > AXH_MAX_GROUP)
      fail(AXH_ERR_ARG, "group width must be 1.." + std::to_string(AXH_MAX_GROUP));
    if (batch->state) fail(AXH_ERR_STATE, "the batch has already run");
    batch->group_width = width;
  });
}

int axh_batch_set_graph_steps(axh_batch *batch, int32_t steps) {
  return guarded([&] {
    check_batch(batch);
    if (steps < 0 || steps > AXH_MAX_GRAPH_STEPS)
      fail(AXH_ERR_ARG, "graph steps must be 0.." + std::to_string(AXH_MAX_GRAPH_STEPS));
    batch->graph_steps = steps;
  });
}

int axh_batch_set_profiling(axh_batch *batch, int on) {
  return guarded([&] {
    check_batch(batch);
    batch->profile = on != 0;
  });
}

int axh_batch_kernel_stats(const axh_batch *batch, int kclass, double *total_ms,
                           int64_t *launches) {
  return guarded([&] {
    check_batch(batch);
    if (!total_ms || !launches) fail(AXH_ERR_ARG, "null pointer");
    if (!batch_kernel_stats(batch, kclass, total_ms, launches))
      fail(AXH_ERR_ARG, "unknown kernel class");
  });
}

// ---------------------------------------------------------------------------
// resident meshes and footprints (request loops)
// ---------------------------------------------------------------------------
int axh_batch_key(axh_ctx *ctx, uint64_t *key) {
  return guarded([&] {
    check_ctx(ctx);
    if (!key) fail(AXH_ERR_ARG, "null pointer");
    batch_key(ctx, key);
  });
}

int axh_mesh_fingerprint(axh_ctx *ctx, uint64_t *fp) {
  return guarded([&] {
    check_ctx(ctx);
    if (!fp) fail(AXH_ERR_ARG, "null pointer");
    mesh_fingerprint(ctx, fp);
  });
}

int axh_footprint(axh_ctx *ctx, size_t *mesh_bytes, size_t *wave_bytes) {
  return guarded([&] {
    check_ctx(ctx);
    if (!mesh_bytes || !wave_bytes) fail(AXH_ERR_ARG, "null pointer");
    footprint(ctx, mesh_bytes, wave_bytes);
  });
}

int axh_mesh_get(axh_ctx *ctx, axh_mesh **out) {
  return guarded([&] {
    check_ctx(ctx);
    if (!out) fail(AXH_ERR_ARG, "null pointer");
    *out = mesh_get(ctx);
  });
}

int axh_mesh_bind(axh_ctx *ctx, const axh_mesh *mesh) {
  return guarded([&] {
    check_ctx(ctx);
    if (!mesh) fail(AXH_ERR_ARG, "null mesh");
    mesh_bind(ctx, mesh);
  });
}

int axh_mesh_bytes(const axh_mesh *mesh, size_t *bytes) {
  return guarded([&] {
    if (!mesh || !bytes) fail(AXH_ERR_ARG, "null pointer");
    *bytes = mesh_bytes(mesh);
  });
}

int axh_mesh_release(axh_mesh *mesh) {
  return guarded([&] { mesh_release(mesh); });
}

// ---- kernel wavefields ------------------------------------------------------

int axh_upload_fluid_gradient(axh_ctx *ctx, const float *inv_rho_fluid,
                              const float *DsDeta_over_J_flu, const float *DzDeta_over_J_flu,
                              const float *DsDxi_over_J_flu, const float *DzDxi_over_J_flu) {
  return guarded([&] {
    check_ctx(ctx);
    const uint64_t nel = (uint64_t)ctx->dims.nel_fluid;
    if (nel == 0) return;
    const float *src[5] = {inv_rho_fluid, DsDeta_over_J_flu, DzDeta_over_J_flu,
                           DsDxi_over_J_flu, DzDxi_over_J_flu};
    for (int q = 0; q < 5; q++) {
      if (!src[q]) fail(AXH_ERR_ARG, "null fluid gradient array");
      ctx->put(kFluidGradNames[q], kF32, {5, 5, nel}, src[q]);
    }
  });
}

// The numbering of build_kwf_grid (SOLVER/meshes_io.F90:489-610): global
// point ids as the reference forms them (fluid points 1..nglob_fluid, solid
// points after them), each numbered where it first appears; only the entry
// that numbers a point is kept (mask 1).
static void kwf_grid_build(const axh_ctx *ctx, int ibeg, int iend, int jbeg, int jend,
                           const int32_t *emask, std::vector<int32_t> &mapping,
                           std::vector<int32_t> &mask, int &nps, int &npf) {
  const axh_dims &d = ctx->dims;
  const size_t nel = (size_t)d.nel_solid + d.nel_fluid;
  mapping.assign(nel * kNpp, 0);
  mask.assign(nel * kNpp, 0);
  std::vector<char> seen((size_t)d.nglob_solid + d.nglob_fluid, 0);
  int ct = 0;
  for (int dom = 0; dom < 2; dom++) {
    const bool solid = dom == 0;
    const int n = solid ? d.nel_solid : d.nel_fluid;
    if (n == 0) continue;
    const int32_t *ig = ctx->need(solid ? "igloc_solid" : "igloc_fluid").as<int32_t>();
    const int32_t *ax = ctx->need(solid ? "axis_solid" : "axis_fluid").as<int32_t>();
    const int goff = solid ? d.nglob_fluid : 0;
    const size_t eoff = solid ? 0 : (size_t)d.nel_solid;
    const int ng = solid ? d.nglob_solid : d.nglob_fluid;
    for (int e = 0; e < n; e++) {
      if (emask && !emask[eoff + e]) continue;
      auto visit = [&](int i, int j) {
        const int p = i + 5 * j;
        const int g = ig[(size_t)e * kNpp + p] - 1;
        if (g < 0 || g >= ng) fail(AXH_ERR_ARG, "igloc out of range");
        if (!seen[(size_t)g + goff]) {
          seen[(size_t)g + goff] = 1;
          mapping[(eoff + e) * kNpp + p] = ++ct;
          mask[(eoff + e) * kNpp + p] = 1;
        }
      };
      for (int j = jbeg; j <= jend; j++) {
        for (int i = ibeg; i <= iend; i++) visit(i, j);
        if (ax[e]) visit(0, j);   // the axis point of the row
      }
    }
    if (solid) nps = ct;
  }
  npf = ct - nps;
}

static void kwf_put_grid(axh_ctx *ctx, int nps, int npf, const std::vector<int32_t> &mapping,
                         const std::vector<int32_t> &mask) {
  const uint64_t n = (uint64_t)mapping.size();
  const int32_t np[2] = {nps, npf};
  ctx->put("kwf_mapping", kI32, {n}, mapping.data());
  ctx->put("kwf_mask", kI32, {n}, mask.data());
  ctx->put("kwf_npoints", kI32, {2}, np);
}

int axh_kwf_grid(axh_ctx *ctx, int32_t ibeg, int32_t iend, int32_t jbeg, int32_t jend,
                 const int32_t *elem_mask, int32_t *npoint_solid, int32_t *npoint_fluid) {
  return guarded([&] {
    check_ctx(ctx);
    if (ibeg < 0 || iend > kNpol || ibeg > iend || jbeg < 0 || jend > kNpol || jbeg > jend)
      fail(AXH_ERR_ARG, "GLL index ranges must lie in 0..npol");
    std::vector<int32_t> mapping, mask;
    int nps = 0, npf = 0;
    kwf_grid_build(ctx, ibeg, iend, jbeg, jend, elem_mask, mapping, mask, nps, npf);
    kwf_put_grid(ctx, nps, npf, mapping, mask);
    if (npoint_solid) *npoint_solid = nps;
    if (npoint_fluid) *npoint_fluid = npf;
  });
}

int axh_upload_kwf_grid(axh_ctx *ctx, int32_t npoint_solid, int32_t npoint_fluid,
                        const int32_t *mapping_ijel_ikwf, const int32_t *kwf_mask) {
  return guarded([&] {
    check_ctx(ctx);
    const axh_dims &d = ctx->dims;
    const size_t nel = (size_t)d.nel_solid + d.nel_fluid;
    if (!mapping_ijel_ikwf || !kwf_mask) fail(AXH_ERR_ARG, "null grid array");
    if (npoint_solid < 0 || npoint_fluid < 0 || (d.nel_fluid == 0 && npoint_fluid > 0))
      fail(AXH_ERR_ARG, "bad kernel-wavefield point counts");
    const int npt = npoint_solid + npoint_fluid;
    std::vector<int32_t> mapping(nel * kNpp, 0), mask(nel * kNpp, 0);
    std::vector<char> hit((size_t)npt, 0);
    for (size_t q = 0; q < nel * kNpp; q++) {
      if (!kwf_mask[q]) continue;
      const int m = mapping_ijel_ikwf[q];
      const bool solid = q < (size_t)d.nel_solid * kNpp;
      const bool ok = solid ? (m >= 1 && m <= npoint_solid)
                            : (m > npoint_solid && m <= npt);
      if (!ok || hit[m - 1]) fail(AXH_ERR_ARG, "kwf_mask / mapping_ijel_ikwf do not number "
                                               "each point once");
      hit[m - 1] = 1;
      mapping[q] = m;
      mask[q] = 1;
    }
    for (int k = 0; k < npt; k++)
      if (!hit[k]) fail(AXH_ERR_ARG, "kernel-wavefield point " + std::to_string(k + 1) +
                                         " has no masked entry");
    kwf_put_grid(ctx, npoint_solid, npoint_fluid, mapping, mask);
  });
}

int axh_kwf_get_grid(const axh_ctx *ctx, int32_t *mapping_ijel_ikwf, int32_t *kwf_mask,
                     int32_t *npoint_solid, int32_t *npoint_fluid) {
  return guarded([&] {
    check_ctx(ctx);
    const Array &np = ctx->need("kwf_npoints");
    if (npoint_solid) *npoint_solid = np.as<int32_t>()[0];
    if (npoint_fluid) *npoint_fluid = np.as<int32_t>()[1];
    const Array &mp = ctx->need("kwf_mapping"), &mk = ctx->need("kwf_mask");
    if (mapping_ijel_ikwf) std::memcpy(mapping_ijel_ikwf, mp.bytes.data(), mp.bytes.size());
    if (kwf_mask) std::memcpy(kwf_mask, mk.bytes.data(), mk.bytes.size());
  });
}

int axh_set_kwf(axh_ctx *ctx, int32_t strain_it, int32_t src_mask, int32_t ring_snapshots) {
  return guarded([&] {
    check_ctx(ctx);
    if (strain_it < 1) fail(AXH_ERR_ARG, "strain_it must be >= 1");
    if (src_mask != 0 && src_mask != 1) fail(AXH_ERR_ARG, "src_mask must be 0 or 1");
    if (ring_snapshots < 0 || ring_snapshots == 1)
      fail(AXH_ERR_ARG, "ring_snapshots must be 0 or at least 2");
    ctx->kwf_strain_it = strain_it;
    ctx->kwf_src_mask = src_mask;
    ctx->kwf_ring = ring_snapshots;
  });
}

int axh_kwf_snapshots(const axh_ctx *ctx, int32_t it_begin, int32_t it_end) {
  if (!ctx || ctx->kwf_strain_it < 1 || it_begin < 1 || it_end < it_begin) return 0;
  const int s = ctx->kwf_strain_it;
  return it_end / s - (it_begin - 1) / s + (it_begin == 1 ? 1 : 0);
}

static void check_kwf_run(axh_ctx *ctx, int32_t it_begin, int32_t it_end, const float *kwf,
                          const char *fn) {
  check_ctx(ctx);
  check_idle(ctx);
  check_range(ctx, it_begin, it_end, fn);
  if (ctx->kwf_strain_it < 1) fail(AXH_ERR_STATE, std::string(fn) + " needs axh_set_kwf first");
  const int32_t *np = ctx->need("kwf_npoints").as<int32_t>();
  if (!kwf && axh_kwf_snapshots(ctx, it_begin, it_end) > 0 && np[0] + np[1] > 0)
    fail(AXH_ERR_ARG, "null kwf buffer");
}

int axh_run_kwf(axh_ctx *ctx, int32_t it_begin, int32_t it_end, float *seis, float *kwf) {
  return guarded([&] {
    check_kwf_run(ctx, it_begin, it_end, kwf, "axh_run_kwf");
    if (!ctx->dev) device_build(ctx);
    device_run_kwf(ctx, it_begin, it_end, seis, kwf);
  });
}

int axh_run_kwf_begin(axh_ctx *ctx, int32_t it_begin, int32_t it_end, float *seis,
                      float *kwf) {
  return guarded([&] {
    check_kwf_run(ctx, it_begin, it_end, kwf, "axh_run_kwf_begin");
    if (!ctx->dev) device_build(ctx);
    device_run_kwf_begin(ctx, it_begin, it_end, seis, kwf);
    ctx->running = true;
  });
}

int axh_convolve_traces(int32_t nt, int32_t ntraces, const float *seis, int32_t nj,
                        const float *stf, double dt, double scale, float *out) {
  return guarded([&] {
    if (nt < 0 || ntraces < 0 || nj < 0) fail(AXH_ERR_ARG, "negative size");
    if (nj > AXH_MAX_STF_TERMS)
      fail(AXH_ERR_ARG, "at most " + std::to_string(AXH_MAX_STF_TERMS) +
                            " source time function terms");
    if ((size_t)nt * ntraces > 0 && (!seis || !out || (nj > 0 && !stf)))
      fail(AXH_ERR_ARG, "null pointer");
    convolve_traces(nt, ntraces, seis, nj, stf, dt, scale, out);
  });
}

int axh_symplectic_coefficients(int32_t scheme, double deltat, double *coefd,
                                double *coefv, double *coeff, int32_t *nstages) {
  return guarded([&] {
    if (!coefd || !coefv || !coeff || !nstages) fail(AXH_ERR_ARG, "null pointer");
    if (scheme == AXH_NEWMARK2) {
      *nstages = 1;
      return;
    }
    int n = symplectic_coefficients(scheme, deltat, coefd, coefv, coeff);
    if (n <= 0) fail(AXH_ERR_UNSUPPORTED, "unsupported time scheme");
    *nstages = n;
  });
}

// ---------------------------------------------------------------------------
// state files
//
//   "AXHSTATE" | u32 version (1) | u32 nrecords
//   per record: char name[48] | u32 dtype (1 f32, 2 f64, 3 i32) | u32 ndim |
//               u64 dims[4] | u64 nbytes | data
// dims are written as the records "dims_i32" (22 int32) and "dims_f64" (2).
// ---------------------------------------------------------------------------
static void write_rec(FILE *f, const std::string &name, const Array &a) {
  char nm[48] = {0};
  std::strncpy(nm, name.c_str(), 47);
  uint32_t dt = a.dtype, nd = (uint32_t)a.dims.size();
  uint64_t dims[4] = {0, 0, 0, 0};
  for (size_t i = 0; i < a.dims.size() && i < 4; i++) dims[i] = a.dims[i];
  uint64_t nb = a.bytes.size();
  if (fwrite(nm, 1, 48, f) != 48 || fwrite(&dt, 4, 1, f) != 1 ||
      fwrite(&nd, 4, 1, f) != 1 || fwrite(dims, 8, 4, f) != 4 ||
      fwrite(&nb, 8, 1, f) != 1 ||
      (nb && fwrite(a.bytes.data(), 1, nb, f) != nb))
    fail(AXH_ERR_IO, "write failed for record " + name);
}

int axh_save_state(const axh_ctx *ctx, const char *path) {
  return guarded([&] {
    check_ctx(ctx);
    if (!path) fail(AXH_ERR_ARG, "null path");
    FILE *f = std::fopen(path, "wb");
    if (!f) fail(AXH_ERR_IO, std::string("cannot open ") + path);
    try {
      const uint32_t ver = 1, nrec = (uint32_t)ctx->arrays.size() + 2;
      if (fwrite("AXHSTATE", 1, 8, f) != 8 || fwrite(&ver, 4, 1, f) != 1 ||
          fwrite(&nrec, 4, 1, f) != 1)
        fail(AXH_ERR_IO, "header write failed");
      Array di, dd;
      di.dtype = kI32;
      di.dims = {22};
      di.bytes.resize(88);
      std::memcpy(di.bytes.data(), &ctx->dims, 88);
      dd.dtype = kF64;
      dd.dims = {2};
      dd.bytes.resize(16);
      std::memcpy(dd.bytes.data(), &ctx->dims.deltat, 8);
      std::memcpy(dd.bytes.data() + 8, &ctx->dims.magnitude, 8);
      write_rec(f, "dims_i32", di);
      write_rec(f, "dims_f64", dd);
      for (auto &kv : ctx->arrays) write_rec(f, kv.first, kv.second);
    } catch (...) {
      std::fclose(f);
      throw;
    }
    if (std::fclose(f) != 0) fail(AXH_ERR_IO, "close failed");
  });
}

int axh_load_state(const char *path, axh_ctx **out) {
  return guarded([&] {
    if (!path || !out) fail(AXH_ERR_ARG, "null argument");
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) fail(AXH_ERR_IO, std::string("cannot open ") + path);
    std::unique_ptr<axh_ctx> c(new axh_ctx);
    bool have_i = false, have_d = false;
    try {
      char magic[8];
      uint32_t ver = 0, nrec = 0;
      if (fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "AXHSTATE", 8) != 0 ||
          fread(&ver, 4, 1, f) != 1 || fread(&nrec, 4, 1, f) != 1 || ver != 1)
        fail(AXH_ERR_IO, "not an axh state file");
      for (uint32_t r = 0; r < nrec; r++) {
        char nm[48];
        uint32_t dt, nd;
        uint64_t dims[4], nb;
        if (fread(nm, 1, 48, f) != 48 || fread(&dt, 4, 1, f) != 1 ||
            fread(&nd, 4, 1, f) != 1 || fread(dims, 8, 4, f) != 4 ||
            fread(&nb, 8, 1, f) != 1)
          fail(AXH_ERR_IO, "truncated state file");
        nm[47] = 0;
        if (nd > 4 || (dt != kF32 && dt != kF64 && dt != kI32))
          fail(AXH_ERR_IO, "bad record header");
        Array a;
        a.dtype = (DType)dt;
        a.dims.assign(dims, dims + nd);
        if (a.count() * (dt == kF64 ? 8 : 4) != nb) fail(AXH_ERR_IO, "bad record size");
        a.bytes.resize(nb);
        if (nb && fread(a.bytes.data(), 1, nb, f) != nb)
          fail(AXH_ERR_IO, "truncated state file");
        std::string name(nm);
        if (name == "dims_i32") {
          if (nb != 88) fail(AXH_ERR_IO, "bad dims record");
          std::memcpy(&c->dims, a.bytes.data(), 88);
          have_i = true;
        } else if (name == "dims_f64") {
          if (nb != 16) fail(AXH_ERR_IO, "bad dims record");
          std::memcpy(&c->dims.deltat, a.bytes.data(), 8);
          std::memcpy(&c->dims.magnitude, a.bytes.data() + 8, 8);
          have_d = true;
        } else {
          c->arrays[name] = std::move(a);
        }
      }
    } catch (...) {
      std::fclose(f);
      throw;
    }
    std::fclose(f);
    if (!have_i || !have_d) fail(AXH_ERR_IO, "state file without dims");
    validate_dims(c->dims);
    *out = c.release();
  });
}

}  // extern "C"
