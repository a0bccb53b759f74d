/*
 * axh_host_glue.c -- small C helpers for the Fortran side of the drop-in.
 *
 * 1. The standard-linear-solid coefficients of the reference's attenuation
 *    module (SOLVER/attenuation.f90:31-37: exp_w_j_deltat, ts_fac_t,
 *    ts_fac_tm1, y_j_attenuation, do_corr_lowq) are PRIVATE module variables,
 *    fitted by simulated annealing with an unseeded random generator
 *    (attenuation.f90:1313), so they cannot be recomputed.  amdflang gives
 *    module variables external linkage whatever their Fortran visibility
 *    (symbols _QM<module>E<name>), and describes allocatable arrays with the
 *    ISO_Fortran_binding descriptor; reading the descriptors gives the exact
 *    values the reference loop would use.  This ties the drop-in host to a
 *    build of the reference with amdflang, which is how it is built.
 * 2. Flat-argument wrappers around axh_upload_attenuation_cg4/_full (Fortran
 *    cannot easily build a struct of pointers to allocatable arrays).
 * 3. getenv for the run-mode switches of the host, and a file copy for the
 *    record mode with several sources (the reference reads its source from
 *    ./inparam_source, SOLVER/source.f90:51).
 */
#include <ISO_Fortran_binding.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "axihip.h"

extern CFI_CDESC_T(1) _QMattenuationEexp_w_j_deltat;
extern CFI_CDESC_T(1) _QMattenuationEts_fac_t;
extern CFI_CDESC_T(1) _QMattenuationEts_fac_tm1;
extern CFI_CDESC_T(1) _QMattenuationEy_j_attenuation;
extern int _QMattenuationEdo_corr_lowq; /* LOGICAL(4) */

static int copy_desc(const void *desc, double *out, int nmax) {
  const CFI_CDESC_T(1) *d = (const CFI_CDESC_T(1) *)desc;
  if (!d->base_addr || d->rank != 1 || d->elem_len != sizeof(double)) return -1;
  long n = (long)d->dim[0].extent;
  if (n < 0 || n > nmax) return -1;
  const char *b = (const char *)d->base_addr;
  for (long i = 0; i < n; i++)
    memcpy(&out[i], b + i * d->dim[0].sm, sizeof(double));
  return (int)n;
}

/* Returns n_sls (>0) or -1 when the module arrays are not allocated. */
int axh_glue_reference_sls(int nmax, double *exp_w, double *ts_t,
                           double *ts_tm1, double *y_j, int *do_corr) {
  int n1 = copy_desc(&_QMattenuationEexp_w_j_deltat, exp_w, nmax);
  int n2 = copy_desc(&_QMattenuationEts_fac_t, ts_t, nmax);
  int n3 = copy_desc(&_QMattenuationEts_fac_tm1, ts_tm1, nmax);
  int n4 = copy_desc(&_QMattenuationEy_j_attenuation, y_j, nmax);
  if (n1 < 1 || n1 != n2 || n1 != n3 || n1 != n4) return -1;
  *do_corr = _QMattenuationEdo_corr_lowq != 0;
  return n1;
}

int axh_glue_upload_att(axh_ctx *ctx, const float *Y, const float *vse,
                        const float *vsx, const float *vze, const float *vzx,
                        const float *dsde, const float *dzde, const float *dsdx,
                        const float *dzdx, const float *invs, const float *dmu,
                        const float *dkap, const double *expw, const double *tst,
                        const double *tstm1, const double *ajmu,
                        const double *ajkap) {
  axh_attenuation_cg4 a;
  a.Y = Y; a.V_s_eta = vse; a.V_s_xi = vsx; a.V_z_eta = vze; a.V_z_xi = vzx;
  a.DsDeta_over_J = dsde; a.DzDeta_over_J = dzde; a.DsDxi_over_J = dsdx;
  a.DzDxi_over_J = dzdx; a.inv_s = invs; a.delta_mu = dmu; a.delta_kappa = dkap;
  a.exp_w_j_deltat = expw; a.ts_fac_t = tst; a.ts_fac_tm1 = tstm1;
  a.a_j_mu = ajmu; a.a_j_kappa = ajkap;
  return axh_upload_attenuation_cg4(ctx, &a);
}

/* The same for axh_upload_attenuation_full. */
int axh_glue_upload_att_full(axh_ctx *ctx, const float *Y, const float *vse,
                             const float *vsx, const float *vze, const float *vzx,
                             const float *y0, const float *v0se, const float *v0sx,
                             const float *v0ze, const float *v0zx, const float *dsde,
                             const float *dzde, const float *dsdx, const float *dzdx,
                             const float *invs, const float *dmu, const float *dkap,
                             const double *expw, const double *tst, const double *tstm1,
                             const double *ajmu, const double *ajkap) {
  axh_attenuation_full a;
  a.Y = Y; a.V_s_eta = vse; a.V_s_xi = vsx; a.V_z_eta = vze; a.V_z_xi = vzx;
  a.Y0 = y0; a.V0_s_eta = v0se; a.V0_s_xi = v0sx; a.V0_z_eta = v0ze; a.V0_z_xi = v0zx;
  a.DsDeta_over_J = dsde; a.DzDeta_over_J = dzde; a.DsDxi_over_J = dsdx;
  a.DzDxi_over_J = dzdx; a.inv_s = invs; a.delta_mu = dmu; a.delta_kappa = dkap;
  a.exp_w_j_deltat = expw; a.ts_fac_t = tst; a.ts_fac_tm1 = tstm1;
  a.a_j_mu = ajmu; a.a_j_kappa = ajkap;
  return axh_upload_attenuation_full(ctx, &a);
}

/* Copies the value of environment variable `name` (NUL-terminated) into val
 * (blank padded, length n) and returns its length, or -1 if unset. */
int axh_glue_getenv(const char *name, char *val, int n) {
  const char *v = getenv(name);
  memset(val, ' ', (size_t)n);
  if (!v) return -1;
  int l = (int)strlen(v);
  if (l > n) l = n;
  memcpy(val, v, (size_t)l);
  return l;
}

/* Copies file `from` to `to` (both NUL-terminated); 0 on success. */
int axh_glue_copy_file(const char *from, const char *to) {
  FILE *a = fopen(from, "rb");
  if (!a) return -1;
  FILE *b = fopen(to, "wb");
  if (!b) {
    fclose(a);
    return -1;
  }
  char buf[4096];
  size_t n;
  int rc = 0;
  while ((n = fread(buf, 1, sizeof buf, a)) > 0)
    if (fwrite(buf, 1, n, b) != n) rc = -1;
  if (ferror(a)) rc = -1;
  fclose(a);
  if (fclose(b) != 0) rc = -1;
  return rc;
}
