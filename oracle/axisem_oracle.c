/*
 * axisem_oracle.c -- CPU restatement of the AxiSEM SOLVER time loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the GPU library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  It is never part of the product path.
 *
 * It restates, element by element and in the reference's own evaluation
 * order, the routines the time loop calls (file:line below are relative to
 * the reference tree).  Storage is element-local exactly as in the reference:
 *   fields (0:4,0:4,nel[,3]) column major, index i + 5 j + 25 e [+ 25 nel c].
 * Floating point follows the reference build (realkind = 4): single precision
 * where the Fortran operands are real(4), double precision where a dp
 * constant (deltat, half_dt, two, third, ts_fac_*, coefd/coefv) enters the
 * expression, rounded to single on assignment.  Compile with
 * -ffp-contract=off so no fused multiply-add changes a rounding, and FTZ is
 * enabled like SOLVER/ftz.c:44-48.
 *
 * Inputs come from an axh state file (the uploads of the C-ABI written by
 * axh_save_state; the format is parsed here independently of the library).
 *
 * Pinning: the seismograms of this restatement equal the reference program's
 * (built from the reference sources) bit for bit on the golden runs of
 * tests/golden (monopole, dipole, quadrupole; Newmark and PEFRL; elastic,
 * coarse-grained and all-point attenuation; a regional mesh with absorbing
 * boundaries).  The McLachlan schemes, which the reference program cannot
 * select, are pinned by their order of accuracy and against the Newmark run.
 * add_source_fl (a source in the fluid), which the reference program cannot
 * place on its meshes (SOLVER/source.f90:531), is pinned by a closed form:
 * with every operator zeroed, a Newmark step from rest is a point-mass update
 * of the source elements' points (tests/test_oracle_math.py).  The
 * kernel-wavefield restatement (orc_kwf_grid, orc_kwf_sample), whose NetCDF
 * output the reference cannot write here, is pinned against the reference's
 * own build_kwf_grid (run by the drop-in host) and against the xdmf
 * snapshots of the reference program, which hold the same solid and fluid
 * displacements at the same GLL points (tests/test_kwf.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#if defined(__x86_64__)
#include <xmmintrin.h>
#endif

#define NP 5
#define NPP 25

/* ------------------------------------------------------------------------ */
/* state file records                                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
  char name[48];
  uint32_t dtype; /* 1 f32, 2 f64, 3 i32 */
  uint32_t ndim;
  uint64_t dims[4];
  uint64_t nbytes;
  void *data;
} rec_t;

typedef struct orc {
  /* dims */
  int nel_s, nel_f, nglob_s, nglob_f, nel_bdry, src_type, scheme, anel, cg,
      n_sls, niter, seis_it, num_rec, nelsrc, fluid_src;
  double deltat, magnitude;
  int nrec;
  rec_t *recs;
  /* inputs */
  const float *G0, *G1, *G1T, *G2, *G2T;
  const float *M[40];    /* 3-D solid terms by enum axh_solid_term    */
  const float *M0[10];   /* axial terms M0_w1..M0_w10               */
  const float *Mfl[4], *M0fl;
  const float *inv_mass_rho, *inv_mass_fluid;
  /* absorbing boundaries (NULL when the mesh reaches the antipode) */
  const float *gamma_s, *gamma_f;
  const int32_t *igloc_s, *igloc_f, *axis_s, *axis_f;
  const int32_t *bdry_solid_el, *bdry_fluid_el, *bdry_jpol_solid,
      *bdry_jpol_fluid;
  const float *bdry_matr, *fsmask;
  const int32_t *ielsrc;
  const float *source_term, *stf;
  int nstf;
  const float *Ycg, *Vseta, *Vsxi, *Vzeta, *Vzxi, *DsDeta, *DzDeta, *DsDxi,
      *DzDxi, *inv_s, *dmu, *dkappa;
  /* COARSE_GRAINED false: the terms above at all 25 points, (0:4,0:4,nel),
   * and the axial vectors (0:4,nel) */
  const float *Y0, *V0seta, *V0sxi, *V0zeta, *V0zxi;
  const double *expw, *tsfac_t, *tsfac_tm1, *aj_mu, *aj_kappa;
  const int32_t *recfile_el;
  int32_t *ax_el_s, *ax_el_f;
  int naxel_s, naxel_f;
  /* fields */
  float *disp, *velo, *acc0, *acc1, *chi, *dchi, *ddchi0, *ddchi1;
  float *memvar, *src_tr_tm1, *src_dev_tm1;
  float *gvec_s, *gvec_f;
  double t;
  int iter;
  /* symplectic */
  int nstages;
  double coefd[40], coefv[40], coeff[40];
} orc_t;

enum { MONO = 0, DI = 1, QUAD = 2 };
enum {
  M11S = 0, M21S, M41S, M12S, M22S, M42S, M11Z, M21Z, M41Z, M32S, M13S, M33S,
  M43S, M1PHI, M2PHI, M4PHI, M_1, M_2, M_3, M_4, M_5, M_6, M_7, M_8, M_W1,
  M_W2, M_W3, M_W4, M_W5
};

static const char *solid3d_names[29] = {
    "M11s", "M21s", "M41s", "M12s", "M22s", "M42s", "M11z", "M21z",
    "M41z", "M32s", "M13s", "M33s", "M43s", "M1phi", "M2phi", "M4phi",
    "M_1",  "M_2",  "M_3",  "M_4",  "M_5",  "M_6",  "M_7",  "M_8",
    "M_w1", "M_w2", "M_w3", "M_w4", "M_w5"};
static const char *fluid3d_names[4] = {"M1chi_fl", "M2chi_fl", "M4chi_fl",
                                       "M_w_fl"};

static const void *find(orc_t *o, const char *name) {
  for (int i = 0; i < o->nrec; i++)
    if (strcmp(o->recs[i].name, name) == 0) return o->recs[i].data;
  return NULL;
}
static int find_n(orc_t *o, const char *name) {
  for (int i = 0; i < o->nrec; i++)
    if (strcmp(o->recs[i].name, name) == 0) {
      uint64_t n = 1;
      for (uint32_t d = 0; d < o->recs[i].ndim; d++) n *= o->recs[i].dims[d];
      return (int)n;
    }
  return 0;
}

/* symplectic_coefficients (SOLVER/time_evol_wave.F90:749-967): PEFRL,
 * ML_SO4m5, ML_SO6m7; values stated in Omelyan, Mryglod & Folk (2002) and
 * McLachlan (1995) as tabulated by the reference. */
static int symp_coeffs(int scheme, double dt, double *cd, double *cv,
                       double *cf) {
  int n = 0;
  if (scheme == 1) {
    /* the reference writes these as default-real (single precision)
       literals assigned to dp variables (time_evol_wave.F90:776-778) */
    const double z = (double)0.1786178958448091f, io = (double)-0.2123418310626054f,
                 k = (double)-0.06626458266981849f;
    n = 4;
    cd[0] = z; cd[1] = k; cd[2] = 1.0 - 2.0 * (z + k); cd[3] = k; cd[4] = z;
    cv[0] = 0.5 - io; cv[1] = io; cv[2] = io; cv[3] = 0.5 - io;
  } else if (scheme == 2) {
    double rho = (14.0 - sqrt(19.0)) / 108.0, th = (20.0 - 7.0 * sqrt(19.0)) / 108.0;
    double nu = 2.0 / 5.0, lam = -1.0 / 10.0;
    n = 5;
    cd[0] = rho; cd[1] = th; cd[2] = 0.5 - rho - th; cd[3] = 0.5 - rho - th;
    cd[4] = th; cd[5] = rho;
    cv[0] = nu; cv[1] = lam; cv[2] = 1.0 - 2.0 * (nu + lam); cv[3] = lam; cv[4] = nu;
  } else if (scheme == 3) {
    n = 7;
    /* single precision literals, as in time_evol_wave.F90:848-851,856-859 */
    cd[0] = (double)-1.01308797891717472981f; cd[1] = (double)1.18742957373254270702f;
    cd[2] = (double)-0.01833585209646059034f; cd[3] = (double)0.34399425728109261313f;
    for (int i = 5; i <= 8; i++) cd[i - 1] = cd[n + 2 - i - 1];
    cv[0] = (double)0.00016600692650009894f; cv[1] = (double)-0.37962421426377360608f;
    cv[2] = (double)0.68913741185181063674f; cv[3] = (double)0.38064159097092574080f;
    for (int i = 5; i <= 7; i++) cv[i - 1] = cv[n + 1 - i - 1];
  } else {
    return 0;
  }
  for (int i = 0; i < n + 1; i++) cd[i] *= dt;
  for (int i = 0; i < n; i++) cv[i] *= dt;
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int j = 0; j <= i; j++) s += cd[j];
    cf[i] = s;
  }
  return n;
}

/* ------------------------------------------------------------------------ */
/* SOLVER/unrolled_loops.f90                                                 */
/* ------------------------------------------------------------------------ */
/* mxm_4 (unrolled_loops.f90:164-188): c(i,j) = sum(a(i,:)*b(:,j)), summed
 * from zero in index order (the order amdflang's SUM uses for this loop). */
static void mxm(const float *a, const float *b, float *c) {
  for (int j = 0; j < NP; j++)
    for (int i = 0; i < NP; i++) {
      float s = 0.0f;
      for (int k = 0; k < NP; k++) s = s + a[i + 5 * k] * b[k + 5 * j];
      c[i + 5 * j] = s;
    }
}
/* vxm_4 (unrolled_loops.f90:194-206): c(j) = sum(a * b(:,j)) */
static void vxm(const float *a, const float *b, float *c) {
  for (int j = 0; j < NP; j++) {
    float s = 0.0f;
    for (int k = 0; k < NP; k++) s = s + a[k] * b[k + 5 * j];
    c[j] = s;
  }
}
/* outerprod_4 (unrolled_loops.f90:221-229) */
static void outer(const float *a, const float *b, float *c) {
  for (int j = 0; j < NP; j++)
    for (int i = 0; i < NP; i++) c[i + 5 * j] = a[i] * b[j];
}
/* mxm_cg4_sparse_a (unrolled_loops.f90:76-97): a sparse (1:4) at
 * (1,1),(1,3),(3,1),(3,3) */
static void sparse_a(const float *a, const float *b, float *c) {
  memset(c, 0, sizeof(float) * NPP);
  for (int j = 0; j < NP; j++) {
    c[1 + 5 * j] = a[0] * b[1 + 5 * j] + a[1] * b[3 + 5 * j];
    c[3 + 5 * j] = a[2] * b[1 + 5 * j] + a[3] * b[3 + 5 * j];
  }
}
/* mxm_cg4_sparse_b (unrolled_loops.f90:101-121) */
static void sparse_b(const float *a, const float *b, float *c) {
  memset(c, 0, sizeof(float) * NPP);
  for (int i = 0; i < NP; i++) {
    c[i + 5 * 1] = a[i + 5 * 1] * b[0] + a[i + 5 * 3] * b[2];
    c[i + 5 * 3] = a[i + 5 * 1] * b[1] + a[i + 5 * 3] * b[3];
  }
}
/* mxm_cg4_sparse_c (unrolled_loops.f90:125-158) */
static void sparse_c(const float *a, const float *b, float *c) {
  const int rows[4] = {1, 1, 3, 3}, cols[4] = {1, 3, 1, 3};
  for (int q = 0; q < 4; q++) {
    int r = rows[q], cl = cols[q];
    float s = 0.0f;
    for (int k = 0; k < NP; k++) s = s + a[r + 5 * k] * b[k + 5 * cl];
    c[q] = s;
  }
}

/* cg4 point q -> local index (1,1),(1,3),(3,1),(3,3) */
static const int cgidx[4] = {1 + 5 * 1, 1 + 5 * 3, 3 + 5 * 1, 3 + 5 * 3};

/* ------------------------------------------------------------------------ */
/* solid stiffness                                                          */
/* ------------------------------------------------------------------------ */
#define ME(name) (o->M[name] + (size_t)e * NPP)

/* glob_stiffness_mono_4 (SOLVER/stiffness_mono.f90:60-157) */
static void stiff_mono(orc_t *o, const float *u, float *out) {
  const int nel = o->nel_s;
  float X1[NPP], X2[NPP], X3[NPP], X4[NPP], S1s[NPP], S2s[NPP], S1z[NPP],
      S2z[NPP], ls[NPP], lz[NPP], V1[5], V2[5], V3[5], V4[5], uz0[5], T[NPP];
  for (int e = 0; e < nel; e++) {
    const float *us = u + (size_t)e * NPP, *uz = u + (size_t)(e + 2 * nel) * NPP;
    const float *m1 = ME(M_1), *m2 = ME(M_2), *m3 = ME(M_3), *m4 = ME(M_4),
                *mw1 = ME(M_W1), *m11s = ME(M11S), *m21s = ME(M21S),
                *m41s = ME(M41S), *m12s = ME(M12S), *m22s = ME(M22S),
                *m32s = ME(M32S), *m42s = ME(M42S), *m11z = ME(M11Z),
                *m21z = ME(M21Z), *m41z = ME(M41Z);
    int ax = o->axis_s[e];
    mxm(ax ? o->G1T : o->G2T, us, X1);
    mxm(ax ? o->G1T : o->G2T, uz, X2);
    mxm(us, o->G2, X3);
    mxm(uz, o->G2, X4);
    for (int p = 0; p < NPP; p++) {
      ls[p] = m4[p] * X4[p] + m2[p] * X3[p] + m1[p] * X1[p] + m3[p] * X2[p] +
              us[p] * mw1[p];
      S1s[p] = m11s[p] * X3[p] + m21s[p] * X1[p] + m12s[p] * X4[p] +
               m22s[p] * X2[p] + m1[p] * us[p];
      S2s[p] = m11s[p] * X1[p] + m41s[p] * X3[p] + m32s[p] * X2[p] +
               m42s[p] * X4[p] + m2[p] * us[p];
      S1z[p] = m11z[p] * X4[p] + m21z[p] * X2[p] + m32s[p] * X3[p] +
               m22s[p] * X1[p] + m3[p] * us[p];
      S2z[p] = m11z[p] * X2[p] + m41z[p] * X4[p] + m12s[p] * X1[p] +
               m42s[p] * X3[p] + m4[p] * us[p];
    }
    mxm(S2s, o->G2T, X2);
    mxm(S2z, o->G2T, X4);
    mxm(ax ? o->G1 : o->G2, S1s, X1);
    mxm(ax ? o->G1 : o->G2, S1z, X3);
    for (int p = 0; p < NPP; p++) {
      ls[p] = ls[p] + X1[p] + X2[p];
      lz[p] = X3[p] + X4[p];
    }
    if (ax) {
      const float *w1 = o->M0[0] + 5 * e, *w2 = o->M0[1] + 5 * e,
                  *w3 = o->M0[2] + 5 * e;
      for (int j = 0; j < 5; j++) uz0[j] = uz[5 * j];
      vxm(o->G0, us, V1);
      vxm(uz0, o->G2, V2);
      for (int j = 0; j < 5; j++) V4[j] = w1[j] * V1[j] + w3[j] * V2[j];
      vxm(o->G0, uz, V3);
      for (int j = 0; j < 5; j++) V4[j] = V4[j] + w2[j] * V3[j];
      float tmp[5];
      for (int j = 0; j < 5; j++) tmp[j] = w2[j] * V1[j];
      outer(o->G0, tmp, X2);
      for (int j = 0; j < 5; j++) V2[j] = w3[j] * V1[j];
      vxm(V2, o->G2T, V1);
      for (int j = 0; j < 5; j++) X2[5 * j] = X2[5 * j] + V1[j];
      outer(o->G0, V4, T);
      for (int p = 0; p < NPP; p++) {
        ls[p] = ls[p] + T[p];
        lz[p] = X2[p] + lz[p];
      }
    }
    memcpy(out + (size_t)e * NPP, ls, sizeof ls);
    memcpy(out + (size_t)(e + 2 * nel) * NPP, lz, sizeof lz);
  }
}

/* glob_stiffness_di_4 (SOLVER/stiffness_di.f90:60-256) */
static void stiff_di(orc_t *o, const float *u, float *out) {
  const int nel = o->nel_s;
  float X1[NPP], X2[NPP], X3[NPP], X4[NPP], X5[NPP], X6[NPP], X7[NPP],
      X8[NPP], ls2[NPP], ls3[NPP], S1p[NPP], S1m[NPP], S2p[NPP], S2m[NPP],
      S1z[NPP], S2z[NPP], l1[NPP], l2[NPP], l3[NPP];
  for (int e = 0; e < nel; e++) {
    const float *u1 = u + (size_t)e * NPP, *u2 = u + (size_t)(e + nel) * NPP,
                *u3 = u + (size_t)(e + 2 * nel) * NPP;
    const float *m1 = ME(M_1), *m2 = ME(M_2), *m3 = ME(M_3), *m4 = ME(M_4),
                *m5 = ME(M_5), *m6 = ME(M_6), *m7 = ME(M_7), *m8 = ME(M_8),
                *mw1 = ME(M_W1), *mw2 = ME(M_W2), *mw3 = ME(M_W3),
                *m11s = ME(M11S), *m21s = ME(M21S), *m41s = ME(M41S),
                *m12s = ME(M12S), *m22s = ME(M22S), *m42s = ME(M42S),
                *m13s = ME(M13S), *m23s = ME(M32S), *m33s = ME(M33S),
                *m43s = ME(M43S), *m11z = ME(M11Z), *m21z = ME(M21Z),
                *m41z = ME(M41Z);
    int ax = o->axis_s[e];
    const float *GT = ax ? o->G1T : o->G2T, *G = ax ? o->G1 : o->G2;
    mxm(u1, o->G2, X4);
    mxm(u2, o->G2, X5);
    mxm(u3, o->G2, X6);
    mxm(GT, u1, X1);
    mxm(GT, u2, X2);
    mxm(GT, u3, X3);
    for (int p = 0; p < NPP; p++) {
      X7[p] = X1[p] + X2[p];
      X8[p] = X4[p] + X5[p];
      ls2[p] = m8[p] * X6[p] + m7[p] * X3[p] + m1[p] * X1[p] + m5[p] * X2[p] +
               m2[p] * X4[p] + m6[p] * X5[p] + mw1[p] * u2[p] + mw2[p] * u3[p];
      ls3[p] = m4[p] * X4[p] - m4[p] * X5[p] + m3[p] * X1[p] - m3[p] * X2[p] +
               mw2[p] * u2[p] + mw3[p] * u3[p];
      float c1 = m13s[p] * X6[p], c2 = m23s[p] * X3[p], c3 = m3[p] * u3[p];
      S1p[p] = c1 + c2 + c3 + m11s[p] * X4[p] + m21s[p] * X1[p] +
               m12s[p] * X5[p] + m22s[p] * X2[p] + m1[p] * u2[p];
      S1m[p] = c1 + c2 - c3 + m11s[p] * X5[p] + m21s[p] * X2[p] +
               m12s[p] * X4[p] + m22s[p] * X1[p] + m5[p] * u2[p];
      c1 = m33s[p] * X3[p];
      c2 = m43s[p] * X6[p];
      c3 = m4[p] * u3[p];
      S2p[p] = c1 + c2 + c3 + m11s[p] * X1[p] + m41s[p] * X4[p] +
               m12s[p] * X2[p] + m42s[p] * X5[p] + m2[p] * u2[p];
      S2m[p] = c1 + c2 - c3 + m11s[p] * X2[p] + m41s[p] * X5[p] +
               m12s[p] * X1[p] + m42s[p] * X4[p] + m6[p] * u2[p];
      S1z[p] = m33s[p] * X8[p] + m23s[p] * X7[p] + m11z[p] * X6[p] +
               m21z[p] * X3[p] + m7[p] * u2[p];
      S2z[p] = m13s[p] * X7[p] + m43s[p] * X8[p] + m11z[p] * X3[p] +
               m41z[p] * X6[p] + m8[p] * u2[p];
    }
    mxm(G, S1p, X1);
    mxm(G, S1m, X3);
    mxm(G, S1z, X5);
    mxm(S2p, o->G2T, X2);
    mxm(S2m, o->G2T, X4);
    mxm(S2z, o->G2T, X6);
    for (int p = 0; p < NPP; p++) {
      l1[p] = X1[p] + X2[p];
      l2[p] = X3[p] + X4[p] + ls2[p];
      l3[p] = X5[p] + X6[p] + ls3[p];
    }
    if (ax) {
      const float *w1 = o->M0[0] + 5 * e, *w2 = o->M0[1] + 5 * e,
                  *w3 = o->M0[2] + 5 * e, *w4 = o->M0[3] + 5 * e,
                  *w6 = o->M0[5] + 5 * e, *w7 = o->M0[6] + 5 * e,
                  *w8 = o->M0[7] + 5 * e, *w9 = o->M0[8] + 5 * e,
                  *w10 = o->M0[9] + 5 * e;
      float u10[5], V1[5], V2[5], V3[5], V4[5], a[5], b[5], c[5];
      for (int j = 0; j < 5; j++) u10[j] = u1[5 * j];
      vxm(o->G0, u1, V1);
      vxm(o->G0, u2, V2);
      vxm(o->G0, u3, V3);
      vxm(u10, o->G2, V4);
      for (int j = 0; j < 5; j++) {
        a[j] = w1[j] * V2[j] + w3[j] * V3[j];
        b[j] = w1[j] * V1[j] + (w2[j] + w6[j]) * V4[j] + w9[j] * V2[j] +
               w10[j] * V3[j];
        c[j] = w3[j] * V1[j] + (w4[j] + w8[j]) * V4[j] + w7[j] * V3[j] +
               w10[j] * V2[j];
      }
      outer(o->G0, a, S1p);
      outer(o->G0, b, S1m);
      outer(o->G0, c, S1z);
      for (int j = 0; j < 5; j++)
        V4[j] = (w2[j] + w6[j]) * V2[j] + (w4[j] + w8[j]) * V3[j];
      vxm(V4, o->G2T, V1);
      for (int j = 0; j < 5; j++) S1p[5 * j] = S1p[5 * j] + V1[j];
      for (int p = 0; p < NPP; p++) {
        l1[p] = l1[p] + S1p[p];
        l2[p] = l2[p] + S1m[p];
        l3[p] = l3[p] + S1z[p];
      }
    }
    memcpy(out + (size_t)e * NPP, l1, sizeof l1);
    memcpy(out + (size_t)(e + nel) * NPP, l2, sizeof l2);
    memcpy(out + (size_t)(e + 2 * nel) * NPP, l3, sizeof l3);
  }
}

/* glob_stiffness_quad_4 (SOLVER/stiffness_quad.f90:238-412) */
static void stiff_quad(orc_t *o, const float *u, float *out) {
  const int nel = o->nel_s;
  float X1[NPP], X2[NPP], X3[NPP], X4[NPP], X5[NPP], X6[NPP], ls[NPP],
      lp[NPP], lz[NPP], S1s[NPP], S2s[NPP], S1z[NPP], S2z[NPP], S1p[NPP],
      S2p[NPP];
  for (int e = 0; e < nel; e++) {
    const float *us = u + (size_t)e * NPP, *up = u + (size_t)(e + nel) * NPP,
                *uz = u + (size_t)(e + 2 * nel) * NPP;
    const float *m1 = ME(M_1), *m2 = ME(M_2), *m3 = ME(M_3), *m4 = ME(M_4),
                *m5 = ME(M_5), *m6 = ME(M_6), *m7 = ME(M_7), *m8 = ME(M_8),
                *mw1 = ME(M_W1), *mw2 = ME(M_W2), *mw3 = ME(M_W3),
                *mw4 = ME(M_W4), *mw5 = ME(M_W5), *m11s = ME(M11S),
                *m21s = ME(M21S), *m41s = ME(M41S), *m12s = ME(M12S),
                *m22s = ME(M22S), *m32s = ME(M32S), *m42s = ME(M42S),
                *m11z = ME(M11Z), *m21z = ME(M21Z), *m41z = ME(M41Z),
                *m1phi = ME(M1PHI), *m2phi = ME(M2PHI), *m4phi = ME(M4PHI);
    int ax = o->axis_s[e];
    const float *GT = ax ? o->G1T : o->G2T, *G = ax ? o->G1 : o->G2;
    mxm(GT, us, X1);
    mxm(GT, up, X2);
    mxm(GT, uz, X3);
    mxm(us, o->G2, X4);
    mxm(up, o->G2, X5);
    mxm(uz, o->G2, X6);
    for (int p = 0; p < NPP; p++) {
      float c1 = m2[p] * X4[p], c2 = m1[p] * X1[p], c3 = m6[p] * X5[p],
            c4 = m5[p] * X2[p], c5 = m4[p] * X6[p], c6 = m3[p] * X3[p];
      ls[p] = c1 + c2 + 2.0f * (c3 + c4) + c5 + c6 + mw1[p] * us[p] +
              mw2[p] * up[p] + 2.0f * mw3[p] * uz[p];
      lp[p] = -2.0f * (c1 + c2 + c5 + c6) - (c3 + c4) + mw2[p] * us[p] +
              mw4[p] * up[p] - mw3[p] * uz[p];
      lz[p] = 2.0f * (m8[p] * X5[p] + m7[p] * X2[p]) +
              mw3[p] * (2.0f * us[p] - up[p]) + mw5[p] * uz[p];
      float usp = us[p] - 2.0f * up[p];
      S1s[p] = m11s[p] * X4[p] + m21s[p] * X1[p] + m12s[p] * X6[p] +
               m22s[p] * X3[p] + m1[p] * usp;
      S2s[p] = m11s[p] * X1[p] + m41s[p] * X4[p] + m32s[p] * X3[p] +
               m42s[p] * X6[p] + m2[p] * usp;
      S1z[p] = m11z[p] * X6[p] + m21z[p] * X3[p] + m32s[p] * X4[p] +
               m22s[p] * X1[p] + m3[p] * usp;
      S2z[p] = m11z[p] * X3[p] + m41z[p] * X6[p] + m12s[p] * X1[p] +
               m42s[p] * X4[p] + m4[p] * usp;
      float tus = 2.0f * us[p] - up[p];
      S1p[p] = m1phi[p] * X5[p] + m2phi[p] * X2[p] + m5[p] * tus +
               2.0f * m7[p] * uz[p];
      S2p[p] = m1phi[p] * X2[p] + m4phi[p] * X5[p] + m6[p] * tus +
               2.0f * m8[p] * uz[p];
    }
    mxm(S2s, o->G2T, X2);
    mxm(S2p, o->G2T, X4);
    mxm(S2z, o->G2T, X6);
    mxm(G, S1s, X1);
    mxm(G, S1p, X3);
    mxm(G, S1z, X5);
    for (int p = 0; p < NPP; p++) {
      ls[p] = ls[p] + X1[p] + X2[p];
      lp[p] = lp[p] + X3[p] + X4[p];
      lz[p] = lz[p] + X5[p] + X6[p];
    }
    if (ax) {
      const float *w1 = o->M0[0] + 5 * e, *w2 = o->M0[1] + 5 * e,
                  *w3 = o->M0[2] + 5 * e, *w4 = o->M0[3] + 5 * e,
                  *w5 = o->M0[4] + 5 * e, *w6 = o->M0[5] + 5 * e;
      float V1[5], V2[5], V3[5], a[5], b[5], c[5], T[NPP];
      vxm(o->G0, us, V1);
      vxm(o->G0, up, V2);
      vxm(o->G0, uz, V3);
      for (int j = 0; j < 5; j++) {
        a[j] = w1[j] * V1[j] + w2[j] * V2[j] + w3[j] * V3[j];
        b[j] = w2[j] * V1[j] + w4[j] * V2[j] + w5[j] * V3[j];
        c[j] = w3[j] * V1[j] + w5[j] * V2[j] + w6[j] * V3[j];
      }
      outer(o->G0, a, T);
      for (int p = 0; p < NPP; p++) ls[p] = ls[p] + T[p];
      outer(o->G0, b, T);
      for (int p = 0; p < NPP; p++) lp[p] = lp[p] + T[p];
      outer(o->G0, c, T);
      for (int p = 0; p < NPP; p++) lz[p] = lz[p] + T[p];
    }
    memcpy(out + (size_t)e * NPP, ls, sizeof ls);
    memcpy(out + (size_t)(e + nel) * NPP, lp, sizeof lp);
    memcpy(out + (size_t)(e + 2 * nel) * NPP, lz, sizeof lz);
  }
}

/* glob_fluid_stiffness_4 (SOLVER/stiffness_fluid.f90:139-216) */
static void stiff_fluid(orc_t *o, const float *chi, float *out) {
  float X1[NPP], X2[NPP], S1[NPP], S2[NPP], loc[NPP];
  for (int e = 0; e < o->nel_f; e++) {
    const float *c = chi + (size_t)e * NPP;
    const float *m1 = o->Mfl[0] + (size_t)e * NPP, *m2 = o->Mfl[1] + (size_t)e * NPP,
                *m4 = o->Mfl[2] + (size_t)e * NPP;
    int ax = o->axis_f[e];
    mxm(ax ? o->G1T : o->G2T, c, X1);
    mxm(c, o->G2, X2);
    for (int p = 0; p < NPP; p++) {
      S1[p] = m1[p] * X2[p] + m2[p] * X1[p];
      S2[p] = m1[p] * X1[p] + m4[p] * X2[p];
    }
    mxm(ax ? o->G1 : o->G2, S1, X1);
    mxm(S2, o->G2T, X2);
    for (int p = 0; p < NPP; p++) loc[p] = X1[p] + X2[p];
    if (o->src_type != MONO) {
      const float *mw = o->Mfl[3] + (size_t)e * NPP;
      for (int p = 0; p < NPP; p++) loc[p] = loc[p] + mw[p] * c[p];
      if (ax) {
        float V1[5], a[5], T[NPP];
        const float *w = o->M0fl + 5 * e;
        vxm(o->G0, c, V1);
        for (int j = 0; j < 5; j++) a[j] = w[j] * V1[j];
        outer(o->G0, a, T);
        for (int p = 0; p < NPP; p++) loc[p] = loc[p] + T[p];
      }
    }
    memcpy(out + (size_t)e * NPP, loc, sizeof loc);
  }
}

/* ------------------------------------------------------------------------ */
/* anelastic stiffness, coarse grained                                      */
/* ------------------------------------------------------------------------ */
/* memvar layout (1:4,6,n_sls,nel_solid) (SOLVER/time_evol_wave.F90:312) */
#define MV(q, c, j, e) \
  o->memvar[(q) + 4 * ((c) + 6 * ((j) + (size_t)o->n_sls * (e)))]

static void rsum(orc_t *o, int e, float r[6][4]) {
  for (int c = 0; c < 6; c++)
    for (int q = 0; q < 4; q++) {
      float s = 0.0f;
      for (int j = 0; j < o->n_sls; j++) s = s + MV(q, c, j, e);
      r[c][q] = s;
    }
}

/* glob_anel_stiffness_{mono,di,quad}_cg4 (SOLVER/stiffness_mono.f90:511-589,
 * stiffness_di.f90:704-825, stiffness_quad.f90:665-774); subtracts from acc */
static void anel_stiff_cg4(orc_t *o, float *acc) {
  const int nel = o->nel_s;
  for (int e = 0; e < nel; e++) {
    float r[6][4], yl[4], vse[4], vsx[4], vze[4], vzx[4];
    float S1a[4], S2a[4], S1b[4], S2b[4], S1c[4], S2c[4];
    float X1[NPP], X2[NPP], X3[NPP], X4[NPP], X5[NPP], X6[NPP];
    float l1[NPP], l2[NPP], l3[NPP];
    rsum(o, e, r);
    float *r1 = r[0], *r2 = r[1], *r3 = r[2], *r4 = r[3], *r5 = r[4], *r6 = r[5];
    for (int q = 0; q < 4; q++) {
      yl[q] = o->Ycg[4 * e + q];
      vse[q] = o->Vseta[4 * e + q];
      vsx[q] = o->Vsxi[4 * e + q];
      vze[q] = o->Vzeta[4 * e + q];
      vzx[q] = o->Vzxi[4 * e + q];
    }
    const float *G = o->axis_s[e] ? o->G1 : o->G2;
    float *a1 = acc + (size_t)e * NPP, *a2 = acc + (size_t)(e + nel) * NPP,
          *a3 = acc + (size_t)(e + 2 * nel) * NPP;
    if (o->src_type == MONO) {
      for (int q = 0; q < 4; q++) {
        S1a[q] = vze[q] * r1[q] + vse[q] * r5[q];
        S2a[q] = vzx[q] * r1[q] + vsx[q] * r5[q];
        S1c[q] = vze[q] * r5[q] + vse[q] * r3[q];
        S2c[q] = vzx[q] * r5[q] + vsx[q] * r3[q];
      }
      sparse_b(G, S1a, X1);
      sparse_b(G, S1c, X3);
      sparse_a(S2a, o->G2T, X2);
      sparse_a(S2c, o->G2T, X4);
      for (int p = 0; p < NPP; p++) {
        l1[p] = X1[p] + X2[p];
        l3[p] = X3[p] + X4[p];
      }
      for (int q = 0; q < 4; q++) l1[cgidx[q]] = l1[cgidx[q]] + yl[q] * r2[q];
      for (int p = 0; p < NPP; p++) {
        a1[p] = a1[p] - l1[p];
        a3[p] = a3[p] - l3[p];
      }
    } else if (o->src_type == DI) {
      for (int q = 0; q < 4; q++) {
        S1a[q] = vze[q] * (r1[q] - r6[q]) + vse[q] * (r5[q] - r4[q]);
        S2a[q] = vzx[q] * (r1[q] - r6[q]) + vsx[q] * (r5[q] - r4[q]);
        S1b[q] = vze[q] * (r1[q] + r6[q]) + vse[q] * (r5[q] + r4[q]);
        S2b[q] = vzx[q] * (r1[q] + r6[q]) + vsx[q] * (r5[q] + r4[q]);
        S1c[q] = vze[q] * r5[q] + vse[q] * r3[q];
        S2c[q] = vzx[q] * r5[q] + vsx[q] * r3[q];
      }
      sparse_b(G, S1a, X1);
      sparse_b(G, S1b, X3);
      sparse_b(G, S1c, X5);
      sparse_a(S2a, o->G2T, X2);
      sparse_a(S2b, o->G2T, X4);
      sparse_a(S2c, o->G2T, X6);
      for (int p = 0; p < NPP; p++) {
        l1[p] = X1[p] + X2[p];
        l2[p] = X3[p] + X4[p];
        l3[p] = X5[p] + X6[p];
      }
      for (int q = 0; q < 4; q++) {
        int k = cgidx[q];
        l2[k] = l2[k] + 2.0f * yl[q] * (r2[q] - r6[q]);
        l3[k] = l3[k] - yl[q] * r4[q];
      }
      for (int p = 0; p < NPP; p++) {
        a1[p] = a1[p] - l1[p];
        a2[p] = a2[p] - l2[p];
        a3[p] = a3[p] - l3[p];
      }
    } else {
      for (int q = 0; q < 4; q++) {
        S1a[q] = vze[q] * r1[q] + vse[q] * r5[q];
        S2a[q] = vzx[q] * r1[q] + vsx[q] * r5[q];
        S1b[q] = vze[q] * r6[q] + vse[q] * r4[q];
        S2b[q] = vzx[q] * r6[q] + vsx[q] * r4[q];
        S1c[q] = vze[q] * r5[q] + vse[q] * r3[q];
        S2c[q] = vzx[q] * r5[q] + vsx[q] * r3[q];
      }
      sparse_b(G, S1a, X1);
      sparse_b(G, S1b, X3);
      sparse_b(G, S1c, X5);
      sparse_a(S2a, o->G2T, X2);
      sparse_a(S2b, o->G2T, X4);
      sparse_a(S2c, o->G2T, X6);
      for (int p = 0; p < NPP; p++) {
        l1[p] = X1[p] + X2[p];
        l2[p] = -X3[p] - X4[p];
        l3[p] = X5[p] + X6[p];
      }
      for (int q = 0; q < 4; q++) {
        int k = cgidx[q];
        l1[k] = l1[k] + yl[q] * (r2[q] - 2.0f * r6[q]);
        l2[k] = l2[k] + yl[q] * (r6[q] - 2.0f * r2[q]);
        l3[k] = l3[k] - 2.0f * yl[q] * r4[q];
      }
      for (int p = 0; p < NPP; p++) {
        a1[p] = a1[p] - l1[p];
        a2[p] = a2[p] - l2[p];
        a3[p] = a3[p] - l3[p];
      }
    }
  }
}

/* axisym_gradient_solid_el_cg4 (SOLVER/pointwise_derivatives.f90:216-246) */
static void grad_cg4(orc_t *o, const float *f, int e, float g[2][4]) {
  float m1[4], m2[4];
  sparse_c(o->axis_s[e] ? o->G1T : o->G2T, f, m1);
  sparse_c(f, o->G2, m2);
  for (int q = 0; q < 4; q++) {
    g[0][q] = o->DzDeta[4 * e + q] * m1[q] + o->DzDxi[4 * e + q] * m2[q];
    g[1][q] = o->DsDeta[4 * e + q] * m1[q] + o->DsDxi[4 * e + q] * m2[q];
  }
}
/* f_over_s_solid_el_cg4 (SOLVER/pointwise_derivatives.f90:55-75) */
static void fos_cg4(orc_t *o, const float *f, int e, float r[4]) {
  for (int q = 0; q < 4; q++) r[q] = o->inv_s[4 * e + q] * f[cgidx[q]];
}

/* compute_strain_att_el_cg4 (SOLVER/attenuation.f90:472-537):
 * Voigt strain at the 4 cg points, g[c][q], c = 0..5 for E1..E6 */
void orc_strain_cg4(orc_t *o, const float *u, int e, float g[6][4]) {
  const int nel = o->nel_s;
  const float *u1 = u + (size_t)e * NPP, *u2 = u + (size_t)(e + nel) * NPP,
              *u3 = u + (size_t)(e + 2 * nel) * NPP;
  float b1[2][4], b2[2][4], tmp[NPP], f[4], f2[4];
  memset(g, 0, sizeof(float) * 24);
  if (o->src_type == DI) {
    for (int p = 0; p < NPP; p++) tmp[p] = u1[p] + u2[p];
    grad_cg4(o, tmp, e, b1);
  } else {
    grad_cg4(o, u1, e, b1);
  }
  grad_cg4(o, u3, e, b2);
  for (int q = 0; q < 4; q++) {
    g[0][q] = b1[0][q];
    g[2][q] = b2[1][q];
    g[4][q] = b1[1][q] + b2[0][q];
  }
  if (o->src_type == MONO) {
    fos_cg4(o, u1, e, f);
    for (int q = 0; q < 4; q++) g[1][q] = f[q];
  } else if (o->src_type == DI) {
    fos_cg4(o, u2, e, f);
    for (int q = 0; q < 4; q++) g[1][q] = 2.0f * f[q];
    for (int p = 0; p < NPP; p++) tmp[p] = u1[p] - u2[p];
    grad_cg4(o, tmp, e, b1);
    fos_cg4(o, u3, e, f);
    for (int q = 0; q < 4; q++) {
      g[3][q] = -f[q] - b1[1][q];
      g[5][q] = -g[1][q] - b1[0][q];
    }
  } else {
    for (int p = 0; p < NPP; p++) tmp[p] = u1[p] - 2.0f * u2[p];
    fos_cg4(o, tmp, e, f);
    for (int q = 0; q < 4; q++) g[1][q] = f[q];
    grad_cg4(o, u2, e, b1);
    fos_cg4(o, u3, e, f);
    for (int p = 0; p < NPP; p++) tmp[p] = u2[p] - 2.0f * u1[p];
    fos_cg4(o, tmp, e, f2);
    for (int q = 0; q < 4; q++) {
      g[3][q] = -(2.0f * f[q]) - b1[1][q];
      g[5][q] = f2[q] - b1[0][q];
    }
  }
}

/* time_step_memvars_cg4 (SOLVER/attenuation.f90:81-202), with the per-element
 * a_j of attenuation.f90:105-127 taken from the upload. */
static void memvars_cg4(orc_t *o, const float *disp) {
  const int nsls = o->n_sls;
  const double third = 1.0 / 3.0;
  for (int e = 0; e < o->nel_s; e++) {
    float g[6][4], tr[4], str[4], sdev[6][4];
    orc_strain_cg4(o, disp, e, g);
    for (int q = 0; q < 4; q++) tr[q] = g[0][q] + g[1][q] + g[2][q];
    memset(sdev, 0, sizeof sdev);
    for (int q = 0; q < 4; q++) {
      float dm = o->dmu[4 * e + q], dk = o->dkappa[4 * e + q];
      str[q] = dk * tr[q];
      for (int c = 0; c < 3; c++)
        sdev[c][q] = (float)((double)(dm * 2.0f) *
                             ((double)g[c][q] - (double)tr[q] * third));
      sdev[4][q] = dm * g[4][q];
      if (o->src_type != MONO) {
        sdev[3][q] = dm * g[3][q];
        sdev[5][q] = dm * g[5][q];
      }
    }
    float *trm1 = o->src_tr_tm1 + 4 * (size_t)e;
    float *devm1 = o->src_dev_tm1 + 24 * (size_t)e; /* (1:4,6,nel) */
    const double *amu = o->aj_mu + (size_t)nsls * e,
                 *akap = o->aj_kappa + (size_t)nsls * e;
    for (int j = 0; j < nsls; j++) {
      double Tt = o->tsfac_t[j], Tm = o->tsfac_tm1[j], E = o->expw[j];
      for (int q = 0; q < 4; q++) {
        float trb = (float)(Tt * akap[j] * (double)str[q] +
                            Tm * akap[j] * (double)trm1[q]);
        float db[6];
        for (int c = 0; c < 6; c++) {
          if (o->src_type == MONO && (c == 3 || c == 5)) continue;
          db[c] = (float)(Tt * amu[j] * (double)sdev[c][q] +
                          Tm * amu[j] * (double)devm1[q + 4 * c]);
        }
        for (int c = 0; c < 3; c++)
          MV(q, c, j, e) = (float)((E * (double)MV(q, c, j, e) + (double)db[c]) +
                                   (double)trb);
        MV(q, 4, j, e) = (float)(E * (double)MV(q, 4, j, e) + (double)db[4]);
        if (o->src_type != MONO) {
          MV(q, 3, j, e) = (float)(E * (double)MV(q, 3, j, e) + (double)db[3]);
          MV(q, 5, j, e) = (float)(E * (double)MV(q, 5, j, e) + (double)db[5]);
        }
      }
    }
    for (int q = 0; q < 4; q++) trm1[q] = str[q];
    for (int c = 0; c < 6; c++)
      for (int q = 0; q < 4; q++) devm1[q + 4 * c] = sdev[c][q];
  }
}

/* ------------------------------------------------------------------------ */
/* anelastic stiffness and memory variables at all 25 points                */
/* (COARSE_GRAINED false)                                                   */
/* ------------------------------------------------------------------------ */
/* memory_var layout (0:4,0:4,6,n_sls,nel_solid) (SOLVER/time_evol_wave.F90:315) */
#define MVF(p, c, j, e) \
  o->memvar[(p) + NPP * ((c) + 6 * ((j) + (size_t)o->n_sls * (e)))]
#define EF(a) ((a) + (size_t)e * NPP)

/* glob_anel_stiffness_{mono,di,quad}_4 (SOLVER/stiffness_mono.f90:409-507,
 * stiffness_di.f90:604-723, stiffness_quad.f90:555-668); subtracts from acc.
 * The dipole z-component adds outerprod(G0, V2) of the p-component
 * (stiffness_di.f90:706-707), as the reference does. */
static void anel_stiff_full(orc_t *o, float *acc) {
  const int nel = o->nel_s;
  for (int e = 0; e < nel; e++) {
    float r[6][NPP], Sa1[NPP], Sa2[NPP], Sb1[NPP], Sb2[NPP], Sc1[NPP], Sc2[NPP];
    float X1[NPP], X2[NPP], X3[NPP], X4[NPP], X5[NPP], X6[NPP];
    float l1[NPP], l2[NPP], l3[NPP], T[NPP];
    float V1[5], V2[5], V3[5], V4[5], r10[5], r20[5], r30[5], r40[5], r50[5], r60[5];
    for (int c = 0; c < 6; c++)
      for (int p = 0; p < NPP; p++) {
        float sm = 0.0f;
        for (int j = 0; j < o->n_sls; j++) sm = sm + MVF(p, c, j, e);
        r[c][p] = sm;
      }
    const float *r1 = r[0], *r2 = r[1], *r3 = r[2], *r4 = r[3], *r5 = r[4], *r6 = r[5];
    const float *yl = EF(o->Ycg), *vse = EF(o->Vseta), *vsx = EF(o->Vsxi),
                *vze = EF(o->Vzeta), *vzx = EF(o->Vzxi);
    const int ax = o->axis_s[e];
    const float *G = ax ? o->G1 : o->G2;
    const float *y0 = o->Y0 + 5 * (size_t)e, *v0se = o->V0seta + 5 * (size_t)e,
                *v0sx = o->V0sxi + 5 * (size_t)e, *v0ze = o->V0zeta + 5 * (size_t)e,
                *v0zx = o->V0zxi + 5 * (size_t)e;
    for (int jj = 0; jj < 5; jj++) {
      r10[jj] = r1[5 * jj]; r20[jj] = r2[5 * jj]; r30[jj] = r3[5 * jj];
      r40[jj] = r4[5 * jj]; r50[jj] = r5[5 * jj]; r60[jj] = r6[5 * jj];
    }
    float *a1 = acc + (size_t)e * NPP, *a2 = acc + (size_t)(e + nel) * NPP,
          *a3 = acc + (size_t)(e + 2 * nel) * NPP;
    if (o->src_type == MONO) {
      for (int p = 0; p < NPP; p++) {
        Sa1[p] = vze[p] * r1[p] + vse[p] * r5[p];
        Sa2[p] = vzx[p] * r1[p] + vsx[p] * r5[p];
        Sc1[p] = vze[p] * r5[p] + vse[p] * r3[p];
        Sc2[p] = vzx[p] * r5[p] + vsx[p] * r3[p];
      }
      mxm(G, Sa1, X1);
      mxm(G, Sc1, X3);
      mxm(Sa2, o->G2T, X2);
      mxm(Sc2, o->G2T, X4);
      for (int p = 0; p < NPP; p++) {
        l1[p] = X1[p] + X2[p] + yl[p] * r2[p];
        l3[p] = X3[p] + X4[p];
      }
      if (ax) {
        for (int jj = 0; jj < 5; jj++) {
          V1[jj] = v0ze[jj] * r10[jj] + v0se[jj] * r50[jj] + y0[jj] * r20[jj];
          V2[jj] = v0ze[jj] * r50[jj] + v0se[jj] * r30[jj];
          V3[jj] = v0zx[jj] * r50[jj] + v0sx[jj] * r30[jj];
        }
        vxm(V3, o->G2T, V4);
        outer(o->G0, V1, T);
        for (int p = 0; p < NPP; p++) l1[p] = l1[p] + T[p];
        outer(o->G0, V2, T);
        for (int p = 0; p < NPP; p++) l3[p] = l3[p] + T[p];
        for (int jj = 0; jj < 5; jj++) l3[5 * jj] = l3[5 * jj] + V4[jj];
      }
      for (int p = 0; p < NPP; p++) {
        a1[p] = a1[p] - l1[p];
        a3[p] = a3[p] - l3[p];
      }
    } else if (o->src_type == DI) {
      for (int p = 0; p < NPP; p++) {
        Sa1[p] = vze[p] * (r1[p] - r6[p]) + vse[p] * (r5[p] - r4[p]);
        Sa2[p] = vzx[p] * (r1[p] - r6[p]) + vsx[p] * (r5[p] - r4[p]);
        Sb1[p] = vze[p] * (r1[p] + r6[p]) + vse[p] * (r5[p] + r4[p]);
        Sb2[p] = vzx[p] * (r1[p] + r6[p]) + vsx[p] * (r5[p] + r4[p]);
        Sc1[p] = vze[p] * r5[p] + vse[p] * r3[p];
        Sc2[p] = vzx[p] * r5[p] + vsx[p] * r3[p];
      }
      mxm(G, Sa1, X1);
      mxm(G, Sb1, X3);
      mxm(G, Sc1, X5);
      mxm(Sa2, o->G2T, X2);
      mxm(Sb2, o->G2T, X4);
      mxm(Sc2, o->G2T, X6);
      for (int p = 0; p < NPP; p++) {
        l1[p] = X1[p] + X2[p];
        l2[p] = X3[p] + X4[p] + 2.0f * yl[p] * (r2[p] - r6[p]);
        l3[p] = X5[p] + X6[p] - yl[p] * r4[p];
      }
      if (ax) {
        for (int jj = 0; jj < 5; jj++) {
          V1[jj] = v0ze[jj] * (r10[jj] - r60[jj]) + v0se[jj] * (r50[jj] - r40[jj]);
          V2[jj] = v0zx[jj] * (r10[jj] - r60[jj]) + v0sx[jj] * (r50[jj] - r40[jj]);
        }
        vxm(V2, o->G2T, V3);
        outer(o->G0, V1, T);
        for (int p = 0; p < NPP; p++) l1[p] = l1[p] + T[p];
        for (int jj = 0; jj < 5; jj++) l1[5 * jj] = l1[5 * jj] + V3[jj];
        for (int jj = 0; jj < 5; jj++)
          V1[jj] = v0ze[jj] * (r10[jj] + r60[jj]) + v0se[jj] * (r50[jj] + r40[jj]) +
                   y0[jj] * 2.0f * (r20[jj] - r60[jj]);
        outer(o->G0, V1, T);
        for (int p = 0; p < NPP; p++) l2[p] = l2[p] + T[p];
        /* V1 of the z-component is computed and not used */
        outer(o->G0, V2, T);
        for (int p = 0; p < NPP; p++) l3[p] = l3[p] + T[p];
      }
      for (int p = 0; p < NPP; p++) {
        a1[p] = a1[p] - l1[p];
        a2[p] = a2[p] - l2[p];
        a3[p] = a3[p] - l3[p];
      }
    } else {
      for (int p = 0; p < NPP; p++) {
        Sa1[p] = vze[p] * r1[p] + vse[p] * r5[p];
        Sa2[p] = vzx[p] * r1[p] + vsx[p] * r5[p];
        Sb1[p] = vze[p] * r6[p] + vse[p] * r4[p];
        Sb2[p] = vzx[p] * r6[p] + vsx[p] * r4[p];
        Sc1[p] = vze[p] * r5[p] + vse[p] * r3[p];
        Sc2[p] = vzx[p] * r5[p] + vsx[p] * r3[p];
      }
      mxm(G, Sa1, X1);
      mxm(G, Sb1, X3);
      mxm(G, Sc1, X5);
      mxm(Sa2, o->G2T, X2);
      mxm(Sb2, o->G2T, X4);
      mxm(Sc2, o->G2T, X6);
      for (int p = 0; p < NPP; p++) {
        l1[p] = X1[p] + X2[p] + yl[p] * (r2[p] - 2.0f * r6[p]);
        l2[p] = -X3[p] - X4[p] + yl[p] * (r6[p] - 2.0f * r2[p]);
        l3[p] = X5[p] + X6[p] - 2.0f * yl[p] * r4[p];
      }
      if (ax) {
        for (int jj = 0; jj < 5; jj++)
          V1[jj] = v0ze[jj] * r10[jj] + y0[jj] * (r20[jj] - 2.0f * r60[jj]);
        outer(o->G0, V1, T);
        for (int p = 0; p < NPP; p++) l1[p] = l1[p] + T[p];
        for (int jj = 0; jj < 5; jj++)
          V1[jj] = -(v0ze[jj] * r60[jj]) + y0[jj] * (r60[jj] - 2.0f * r20[jj]);
        outer(o->G0, V1, T);
        for (int p = 0; p < NPP; p++) l2[p] = l2[p] + T[p];
        for (int jj = 0; jj < 5; jj++) V1[jj] = v0se[jj] * r30[jj];
        outer(o->G0, V1, T);
        for (int p = 0; p < NPP; p++) l3[p] = l3[p] + T[p];
      }
      for (int p = 0; p < NPP; p++) {
        a1[p] = a1[p] - l1[p];
        a2[p] = a2[p] - l2[p];
        a3[p] = a3[p] - l3[p];
      }
    }
  }
}

/* axisym_gradient_solid_el_4 (SOLVER/pointwise_derivatives.f90:252-286):
 * g[0] = d/ds, g[1] = d/dz at the 25 points */
static void grad_full(orc_t *o, const float *f, int e, float g[2][NPP]) {
  float m1[NPP], m2[NPP];
  mxm(o->axis_s[e] ? o->G1T : o->G2T, f, m1);
  mxm(f, o->G2, m2);
  const float *dze = EF(o->DzDeta), *dzx = EF(o->DzDxi), *dse = EF(o->DsDeta),
              *dsx = EF(o->DsDxi);
  for (int p = 0; p < NPP; p++) {
    g[0][p] = dze[p] * m1[p] + dzx[p] * m2[p];
    g[1][p] = dse[p] * m1[p] + dsx[p] * m2[p];
  }
}
/* f_over_s_solid_el_4 (SOLVER/pointwise_derivatives.f90:104-124): f/s, and
 * on the axis (ipol = 0 of axial elements) df/ds from dsdf_elem_solid
 * (pointwise_derivatives.f90:419-445) */
static void fos_full(orc_t *o, const float *f, int e, float r[NPP]) {
  const float *is = EF(o->inv_s);
  for (int p = 0; p < NPP; p++) r[p] = is[p] * f[p];
  if (o->axis_s[e]) {
    float g[2][NPP];
    grad_full(o, f, e, g);
    for (int jj = 0; jj < 5; jj++) r[5 * jj] = g[0][5 * jj];
  }
}

/* compute_strain_att_el_4 (SOLVER/attenuation.f90:542-606): Voigt strain at
 * the 25 points, g[c][p], c = 0..5 for E1..E6 */
void orc_strain_full(orc_t *o, const float *u, int e, float g[6][NPP]) {
  const int nel = o->nel_s;
  const float *u1 = u + (size_t)e * NPP, *u2 = u + (size_t)(e + nel) * NPP,
              *u3 = u + (size_t)(e + 2 * nel) * NPP;
  float b1[2][NPP], b2[2][NPP], tmp[NPP], f[NPP], f2[NPP];
  memset(g, 0, sizeof(float) * 6 * NPP);
  if (o->src_type == DI) {
    for (int p = 0; p < NPP; p++) tmp[p] = u1[p] + u2[p];
    grad_full(o, tmp, e, b1);
  } else {
    grad_full(o, u1, e, b1);
  }
  grad_full(o, u3, e, b2);
  for (int p = 0; p < NPP; p++) {
    g[0][p] = b1[0][p];
    g[2][p] = b2[1][p];
    g[4][p] = b1[1][p] + b2[0][p];
  }
  if (o->src_type == MONO) {
    fos_full(o, u1, e, f);
    for (int p = 0; p < NPP; p++) g[1][p] = f[p];
  } else if (o->src_type == DI) {
    fos_full(o, u2, e, f);
    for (int p = 0; p < NPP; p++) g[1][p] = 2.0f * f[p];
    for (int p = 0; p < NPP; p++) tmp[p] = u1[p] - u2[p];
    grad_full(o, tmp, e, b1);
    fos_full(o, u3, e, f);
    for (int p = 0; p < NPP; p++) {
      g[3][p] = -f[p] - b1[1][p];
      g[5][p] = -g[1][p] - b1[0][p];
    }
  } else {
    for (int p = 0; p < NPP; p++) tmp[p] = u1[p] - 2.0f * u2[p];
    fos_full(o, tmp, e, f);
    for (int p = 0; p < NPP; p++) g[1][p] = f[p];
    grad_full(o, u2, e, b1);
    fos_full(o, u3, e, f);
    for (int p = 0; p < NPP; p++) tmp[p] = u2[p] - 2.0f * u1[p];
    fos_full(o, tmp, e, f2);
    for (int p = 0; p < NPP; p++) {
      g[3][p] = -(2.0f * f[p]) - b1[1][p];
      g[5][p] = f2[p] - b1[0][p];
    }
  }
}

/* time_step_memvars_4 (SOLVER/attenuation.f90:210-334), with the per-element
 * a_j of attenuation.f90:244-262 taken from the upload. */
static void memvars_full(orc_t *o, const float *disp) {
  const int nsls = o->n_sls;
  const double third = 1.0 / 3.0;
  for (int e = 0; e < o->nel_s; e++) {
    float g[6][NPP], tr[NPP], str[NPP], sdev[6][NPP];
    orc_strain_full(o, disp, e, g);
    for (int p = 0; p < NPP; p++) tr[p] = g[0][p] + g[1][p] + g[2][p];
    memset(sdev, 0, sizeof sdev);
    const float *dmu = EF(o->dmu), *dkap = EF(o->dkappa);
    for (int p = 0; p < NPP; p++) {
      float dm = dmu[p];
      str[p] = dkap[p] * tr[p];
      for (int c = 0; c < 3; c++)
        sdev[c][p] = (float)((double)(dm * 2.0f) *
                             ((double)g[c][p] - (double)tr[p] * third));
      sdev[4][p] = dm * g[4][p];
      if (o->src_type != MONO) {
        sdev[3][p] = dm * g[3][p];
        sdev[5][p] = dm * g[5][p];
      }
    }
    float *trm1 = o->src_tr_tm1 + NPP * (size_t)e;
    float *devm1 = o->src_dev_tm1 + 6 * NPP * (size_t)e; /* (0:4,0:4,6,nel) */
    const double *amu = o->aj_mu + (size_t)nsls * e,
                 *akap = o->aj_kappa + (size_t)nsls * e;
    for (int j = 0; j < nsls; j++) {
      double Tt = o->tsfac_t[j], Tm = o->tsfac_tm1[j], E = o->expw[j];
      for (int p = 0; p < NPP; p++) {
        float trb = (float)(Tt * akap[j] * (double)str[p] +
                            Tm * akap[j] * (double)trm1[p]);
        float db[6];
        for (int c = 0; c < 6; c++) {
          if (o->src_type == MONO && (c == 3 || c == 5)) continue;
          db[c] = (float)(Tt * amu[j] * (double)sdev[c][p] +
                          Tm * amu[j] * (double)devm1[p + NPP * c]);
        }
        for (int c = 0; c < 3; c++)
          MVF(p, c, j, e) = (float)((E * (double)MVF(p, c, j, e) + (double)db[c]) +
                                    (double)trb);
        MVF(p, 4, j, e) = (float)(E * (double)MVF(p, 4, j, e) + (double)db[4]);
        if (o->src_type != MONO) {
          MVF(p, 3, j, e) = (float)(E * (double)MVF(p, 3, j, e) + (double)db[3]);
          MVF(p, 5, j, e) = (float)(E * (double)MVF(p, 5, j, e) + (double)db[5]);
        }
      }
    }
    for (int p = 0; p < NPP; p++) trm1[p] = str[p];
    for (int c = 0; c < 6; c++)
      for (int p = 0; p < NPP; p++) devm1[p + NPP * c] = sdev[c][p];
  }
}

/* time_step_memvars / glob_anel_stiffness_* dispatch on att_coarse_grained
 * (SOLVER/attenuation.f90:55-74, stiffness_mono.f90:285-303) */
static void anel_stiff(orc_t *o, float *acc) {
  if (o->cg) anel_stiff_cg4(o, acc);
  else anel_stiff_full(o, acc);
}
static void memvars(orc_t *o, const float *disp) {
  if (o->cg) memvars_cg4(o, disp);
  else memvars_full(o, disp);
}
/* memory-variable storage per element: (1:4,6,n_sls) or (0:4,0:4,6,n_sls) */
static size_t mv_per_el(const orc_t *o) { return (size_t)(o->cg ? 4 : NPP) * 6 * o->n_sls; }
static size_t pts_per_el(const orc_t *o) { return o->cg ? 4 : NPP; }

/* ------------------------------------------------------------------------ */
/* assembly: pdistsum_solid / pdistsum_fluid (SOLVER/commun.F90:69-283)      */
/* ------------------------------------------------------------------------ */
/* The gather walks the 16 edge points of every element in the order of
 * commun.F90:101-128 (jpol=0 row, the ipol=0/npol columns, jpol=npol row);
 * igloc holds an entry for all 25 points, index i + 5 j. */
static void pdistsum(const int32_t *igloc, float *gvec, int nglob, float *vec,
                     int nel, int nc) {
  static const int walk[16][2] = {{0, 0}, {1, 0}, {2, 0}, {3, 0}, {4, 0},
                                  {0, 1}, {4, 1}, {0, 2}, {4, 2}, {0, 3},
                                  {4, 3}, {0, 4}, {1, 4}, {2, 4}, {3, 4},
                                  {4, 4}};
  for (int c = 0; c < nc; c++) {
    float *g = gvec + (size_t)c * nglob;
    float *v = vec + (size_t)c * nel * NPP;
    memset(g, 0, sizeof(float) * (size_t)nglob);
    for (int e = 0; e < nel; e++)
      for (int w = 0; w < 16; w++) {
        int k = walk[w][0] + 5 * walk[w][1];
        int id = igloc[(size_t)e * NPP + k] - 1;
        g[id] = g[id] + v[(size_t)e * NPP + k];
      }
    for (int e = 0; e < nel; e++)
      for (int w = 0; w < 16; w++) {
        int k = walk[w][0] + 5 * walk[w][1];
        v[(size_t)e * NPP + k] = g[igloc[(size_t)e * NPP + k] - 1];
      }
  }
}

/* ------------------------------------------------------------------------ */
/* masks, coupling, sources (SOLVER/apply_masks.f90, time_evol_wave.F90)     */
/* ------------------------------------------------------------------------ */
/* apply_axis_mask_{scal,onecomp,twocomp,threecomp} (apply_masks.f90:40-101):
 * zero row ipol=0 of axial elements for components c0..c1 */
static void axis_mask(float *u, int nel, const int32_t *ax, int nax, int c0,
                      int c1) {
  for (int c = c0; c <= c1; c++)
    for (int a = 0; a < nax; a++) {
      int e = ax[a];
      for (int j = 0; j < NP; j++) u[((size_t)e + (size_t)c * nel) * NPP + 5 * j] = 0.0f;
    }
}
static void solid_mask(orc_t *o, float *u) {
  if (o->src_type == MONO) axis_mask(u, o->nel_s, o->ax_el_s, o->naxel_s, 0, 0);
  else if (o->src_type == DI) axis_mask(u, o->nel_s, o->ax_el_s, o->naxel_s, 1, 2);
  else axis_mask(u, o->nel_s, o->ax_el_s, o->naxel_s, 0, 2);
}

/* bdry_copy2fluid (SOLVER/time_evol_wave.F90:1532-1570) */
static void copy2fluid(orc_t *o, float *uflu, const float *usol) {
  const int ns = o->nel_s, nf = o->nel_f;
  for (int b = 0; b < o->nel_bdry; b++) {
    int js = o->bdry_jpol_solid[b], jf = o->bdry_jpol_fluid[b];
    int es = o->bdry_solid_el[b] - 1, ef = o->bdry_fluid_el[b] - 1;
    const float *bm1 = o->bdry_matr + 5 * (size_t)b,
                *bm2 = o->bdry_matr + 5 * ((size_t)b + o->nel_bdry);
    for (int i = 0; i < NP; i++) {
      float *f = &uflu[(size_t)ef * NPP + i + 5 * jf];
      float s1 = usol[(size_t)es * NPP + i + 5 * js];
      float s2 = usol[((size_t)es + ns) * NPP + i + 5 * js];
      float s3 = usol[((size_t)es + 2 * ns) * NPP + i + 5 * js];
      if (o->src_type == DI) *f = *f - bm1[i] * (s1 + s2) - bm2[i] * s3;
      else *f = *f - bm1[i] * s1 - bm2[i] * s3;
    }
  }
  (void)nf;
}
/* bdry_copy2solid (SOLVER/time_evol_wave.F90:1576-1611) */
static void copy2solid(orc_t *o, float *usol, const float *uflu) {
  const int ns = o->nel_s;
  for (int b = 0; b < o->nel_bdry; b++) {
    int js = o->bdry_jpol_solid[b], jf = o->bdry_jpol_fluid[b];
    int es = o->bdry_solid_el[b] - 1, ef = o->bdry_fluid_el[b] - 1;
    const float *bm1 = o->bdry_matr + 5 * (size_t)b,
                *bm2 = o->bdry_matr + 5 * ((size_t)b + o->nel_bdry);
    for (int i = 0; i < NP; i++) {
      float fl = uflu[(size_t)ef * NPP + i + 5 * jf];
      size_t p1 = (size_t)es * NPP + i + 5 * js;
      usol[p1] = usol[p1] + bm1[i] * fl;
      if (o->src_type == DI) usol[p1 + (size_t)ns * NPP] = usol[p1 + (size_t)ns * NPP] + bm1[i] * fl;
      usol[p1 + 2 * (size_t)ns * NPP] = usol[p1 + 2 * (size_t)ns * NPP] + bm2[i] * fl;
    }
  }
}
/* add_source_el / add_source_fl (SOLVER/time_evol_wave.F90:1062-1097) */
static void add_source_el(orc_t *o, float *acc, float stf1) {
  if (stf1 == 0.0f) return;
  for (int k = 0; k < o->nelsrc; k++) {
    int e = o->ielsrc[k] - 1;
    for (int c = 0; c < 3; c++)
      for (int p = 0; p < NPP; p++) {
        size_t a = ((size_t)e + (size_t)c * o->nel_s) * NPP + p;
        /* source_term_el(0:4,0:4,8,3), source.f90:376 */
        acc[a] = acc[a] - o->source_term[p + NPP * ((size_t)k + 8 * (size_t)c)] * stf1;
      }
  }
}
static void add_source_fl(orc_t *o, float *chi, float stf1) {
  if (stf1 == 0.0f) return;
  for (int k = 0; k < o->nelsrc; k++) {
    int e = o->ielsrc[k] - 1;
    for (int p = 0; p < NPP; p++)
      chi[(size_t)e * NPP + p] = chi[(size_t)e * NPP + p] -
                                 o->source_term[(size_t)k * NPP + p] * stf1;
  }
}

/* ------------------------------------------------------------------------ */
/* loop                                                                      */
/* ------------------------------------------------------------------------ */
static void solid_stiffness(orc_t *o, const float *u, float *acc) {
  if (o->src_type == MONO) stiff_mono(o, u, acc);
  else if (o->src_type == DI) stiff_di(o, u, acc);
  else stiff_quad(o, u, acc);
}

/* receiver sample of compute_recfile_seis_bare (SOLVER/seismograms.f90:742) */
static void sample(orc_t *o, float *out) {
  const int ns = o->nel_s;
  for (int r = 0; r < o->num_rec; r++) {
    int e = o->recfile_el[r] - 1, i = o->recfile_el[r + o->num_rec],
        j = o->recfile_el[r + 2 * o->num_rec];
    size_t p = (size_t)e * NPP + i + 5 * j;
    float d1 = o->disp[p], d2 = o->disp[p + (size_t)ns * NPP],
          d3 = o->disp[p + 2 * (size_t)ns * NPP];
    float *s = out + 3 * (size_t)r;
    if (o->src_type == MONO) { s[0] = d1; s[1] = d3; s[2] = 0.0f; }
    else if (o->src_type == DI) { s[0] = d1 + d2; s[1] = d1 - d2; s[2] = d3; }
    else { s[0] = d1; s[1] = d2; s[2] = d3; }
  }
}

/* runtime_info blow-up test (SOLVER/time_evol_wave.F90:1042): disp(1,1,:,:) */
static int blown_up(orc_t *o) {
  float m = 0.0f;
  size_t n = 3 * (size_t)o->nel_s;
  for (size_t e = 0; e < n; e++) {
    float v = fabsf(o->disp[e * NPP + 1 + 5]);
    if (v > m) m = v;
  }
  return (double)m > 10.0 * fabs(o->magnitude);
}

/* One newmark2 step, sf_time_loop_newmark (SOLVER/time_evol_wave.F90:353-495) */
/* Absorbing boundaries after the inverse mass (SOLVER/time_evol_wave.F90:432-434,
 * 468-491, 663-665, 690-711): a = a - 2*gamma*v - gamma**2*u in FP32, left to right
 * ((2*gamma)*v first, gamma**2 = gamma*gamma), for n element-local points. */
static void absorb(float *a, const float *gamma, const float *v, const float *u, size_t n) {
  for (size_t p = 0; p < n; p++) {
    const float g = gamma[p];
    a[p] = (a[p] - (2.0f * g) * v[p]) - (g * g) * u[p];
  }
}

static void newmark_step(orc_t *o) {
  const size_t ns = (size_t)o->nel_s * NPP, nf = (size_t)o->nel_f * NPP;
  /* half_dt, half_dt_sq as parameters.F90:1124-1125 */
  const double dt = o->deltat, hdt = 0.5 * dt, hdt2 = 0.5 * (dt * dt);
  const int iter = o->iter;
  for (size_t p = 0; p < nf; p++)
    o->chi[p] = (float)((double)o->chi[p] + dt * (double)o->dchi[p] +
                        hdt2 * (double)o->ddchi0[p]);
  for (int c = 0; c < 3; c++) {
    if (c == 1 && o->src_type == MONO) continue;
    float *d = o->disp + c * ns, *v = o->velo + c * ns, *a = o->acc0 + c * ns;
    for (size_t p = 0; p < ns; p++)
      d[p] = (float)((double)d[p] + dt * (double)v[p] + hdt2 * (double)a[p]);
  }
  if (o->src_type != MONO) axis_mask(o->chi, o->nel_f, o->ax_el_f, o->naxel_f, 0, 0);
  if (o->nel_f) stiff_fluid(o, o->chi, o->ddchi1);
  if (o->fluid_src) add_source_fl(o, o->ddchi1, o->stf[iter - 1]);
  copy2fluid(o, o->ddchi1, o->disp);
  if (o->src_type != MONO) axis_mask(o->ddchi1, o->nel_f, o->ax_el_f, o->naxel_f, 0, 0);
  for (size_t p = 0; p < nf; p++) o->ddchi1[p] = o->ddchi1[p] * o->fsmask[p];
  if (o->nel_f) pdistsum(o->igloc_f, o->gvec_f, o->nglob_f, o->ddchi1, o->nel_f, 1);
  solid_mask(o, o->disp);
  solid_stiffness(o, o->disp, o->acc1);
  if (o->anel) anel_stiff(o, o->acc1);
  for (size_t p = 0; p < nf; p++) o->ddchi1[p] = -o->inv_mass_fluid[p] * o->ddchi1[p];
  if (o->gamma_f) absorb(o->ddchi1, o->gamma_f, o->dchi, o->chi, nf);
  copy2solid(o, o->acc1, o->ddchi1);
  solid_mask(o, o->acc1);
  pdistsum(o->igloc_s, o->gvec_s, o->nglob_s, o->acc1, o->nel_s, 3);
  if (o->anel) memvars(o, o->disp);
  for (size_t p = 0; p < nf; p++) {
    o->dchi[p] = (float)((double)o->dchi[p] +
                         hdt * (double)(o->ddchi0[p] + o->ddchi1[p]));
    o->ddchi0[p] = o->ddchi1[p];
  }
  if (!o->fluid_src) add_source_el(o, o->acc1, o->stf[iter - 1]);
  static const int order[3] = {0, 2, 1};
  for (int k = 0; k < 3; k++) {
    int c = order[k];
    if (c == 1 && o->src_type == MONO) continue;
    float *a1 = o->acc1 + c * ns, *a0 = o->acc0 + c * ns, *v = o->velo + c * ns;
    for (size_t p = 0; p < ns; p++) {
      if (c == 2 && o->src_type == DI)
        a1[p] = (float)(-2.0 * (double)o->inv_mass_rho[p] * (double)a1[p]);
      else
        a1[p] = -o->inv_mass_rho[p] * a1[p];
      if (o->gamma_s) absorb(a1 + p, o->gamma_s + p, v + p, o->disp + c * ns + p, 1);
      v[p] = (float)((double)v[p] + hdt * (double)(a0[p] + a1[p]));
      a0[p] = a1[p];
    }
  }
}

/* One symplectic step, symplectic_time_loop (SOLVER/time_evol_wave.F90:584-731) */
static void symplectic_step(orc_t *o) {
  const size_t ns = (size_t)o->nel_s * NPP, nf = (size_t)o->nel_f * NPP;
  const int iter = o->iter;
  float *acc = o->acc1, *ddchi = o->ddchi0;
  for (int i = 0; i < o->nstages; i++) {
    double cd = o->coefd[i], cv = o->coefv[i];
    for (size_t p = 0; p < nf; p++)
      o->chi[p] = (float)((double)o->chi[p] + (double)o->dchi[p] * cd);
    for (int c = 0; c < 3; c++) {
      if (c == 1 && o->src_type == MONO) continue;
      float *d = o->disp + c * ns, *v = o->velo + c * ns;
      for (size_t p = 0; p < ns; p++) d[p] = (float)((double)d[p] + (double)v[p] * cd);
    }
    if (o->src_type != MONO) axis_mask(o->chi, o->nel_f, o->ax_el_f, o->naxel_f, 0, 0);
    if (o->nel_f) stiff_fluid(o, o->chi, ddchi);
    copy2fluid(o, ddchi, o->disp);
    if (o->src_type != MONO) axis_mask(ddchi, o->nel_f, o->ax_el_f, o->naxel_f, 0, 0);
    if (o->nel_f) pdistsum(o->igloc_f, o->gvec_f, o->nglob_f, ddchi, o->nel_f, 1);
    solid_mask(o, o->disp);
    solid_stiffness(o, o->disp, acc);
    if (o->anel) anel_stiff(o, acc);
    for (size_t p = 0; p < nf; p++) ddchi[p] = -ddchi[p] * o->inv_mass_fluid[p];
    if (o->gamma_f) absorb(ddchi, o->gamma_f, o->dchi, o->chi, nf);
    copy2solid(o, acc, ddchi);
    solid_mask(o, acc);
    pdistsum(o->igloc_s, o->gvec_s, o->nglob_s, acc, o->nel_s, 3);
    for (size_t p = 0; p < nf; p++)
      o->dchi[p] = (float)((double)o->dchi[p] + cv * (double)ddchi[p]);
    add_source_el(o, acc, o->stf[(size_t)(iter - 1) * o->nstages + i]);
    static const int order[3] = {0, 1, 2};
    for (int k = 0; k < 3; k++) {
      int c = order[k];
      if (c == 1 && o->src_type == MONO) continue;
      float *a = acc + c * ns, *v = o->velo + c * ns;
      for (size_t p = 0; p < ns; p++) {
        a[p] = -o->inv_mass_rho[p] * a[p];
        if (o->gamma_s) absorb(a + p, o->gamma_s + p, v + p, o->disp + c * ns + p, 1);
        if (c == 2 && o->src_type == DI)
          v[p] = (float)((double)v[p] + 2.0 * (double)a[p] * cv);
        else
          v[p] = (float)((double)v[p] + (double)a[p] * cv);
      }
    }
  }
  double cd = o->coefd[o->nstages];
  for (size_t p = 0; p < nf; p++)
    o->chi[p] = (float)((double)o->chi[p] + (double)o->dchi[p] * cd);
  for (int c = 0; c < 3; c++) {
    if (c == 1 && o->src_type == MONO) continue;
    float *d = o->disp + c * ns, *v = o->velo + c * ns;
    for (size_t p = 0; p < ns; p++) d[p] = (float)((double)d[p] + (double)v[p] * cd);
  }
  if (o->anel) memvars(o, o->disp);
}

/* ------------------------------------------------------------------------ */
/* public entry points (ctypes)                                              */
/* ------------------------------------------------------------------------ */
static void enable_ftz(void) {
#if defined(__x86_64__)
  _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_ON);
#endif
}

void orc_free(orc_t *o) {
  if (!o) return;
  for (int i = 0; i < o->nrec; i++) free(o->recs[i].data);
  free(o->recs);
  free(o->ax_el_s); free(o->ax_el_f);
  free(o->disp); free(o->velo); free(o->acc0); free(o->acc1);
  free(o->chi); free(o->dchi); free(o->ddchi0); free(o->ddchi1);
  free(o->memvar); free(o->src_tr_tm1); free(o->src_dev_tm1);
  free(o->gvec_s); free(o->gvec_f);
  free(o);
}

static float *zalloc(size_t n) {
  float *p = (float *)calloc(n ? n : 1, sizeof(float));
  return p;
}

orc_t *orc_load(const char *path) {
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  char magic[8];
  uint32_t ver, nrec;
  if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "AXHSTATE", 8) != 0 ||
      fread(&ver, 4, 1, f) != 1 || fread(&nrec, 4, 1, f) != 1) {
    fclose(f);
    return NULL;
  }
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  o->nrec = (int)nrec;
  o->recs = (rec_t *)calloc(nrec, sizeof(rec_t));
  for (uint32_t r = 0; r < nrec; r++) {
    rec_t *R = &o->recs[r];
    if (fread(R->name, 1, 48, f) != 48 || fread(&R->dtype, 4, 1, f) != 1 ||
        fread(&R->ndim, 4, 1, f) != 1 || fread(R->dims, 8, 4, f) != 4 ||
        fread(&R->nbytes, 8, 1, f) != 1) {
      fclose(f); orc_free(o); return NULL;
    }
    R->name[47] = 0;
    R->data = malloc(R->nbytes ? R->nbytes : 1);
    if (R->nbytes && fread(R->data, 1, R->nbytes, f) != R->nbytes) {
      fclose(f); orc_free(o); return NULL;
    }
  }
  fclose(f);
  const int32_t *di = (const int32_t *)find(o, "dims_i32");
  const double *dd = (const double *)find(o, "dims_f64");
  if (!di || !dd) { orc_free(o); return NULL; }
  o->nel_s = di[2]; o->nel_f = di[3]; o->nglob_s = di[4]; o->nglob_f = di[5];
  o->nel_bdry = di[6]; o->src_type = di[7]; o->scheme = di[8]; o->anel = di[9];
  o->cg = di[10]; o->n_sls = di[11]; o->niter = di[12]; o->seis_it = di[13];
  o->num_rec = di[14]; o->nelsrc = di[15]; o->fluid_src = di[16];
  o->deltat = dd[0]; o->magnitude = dd[1];
  o->G0 = find(o, "G0"); o->G1 = find(o, "G1"); o->G1T = find(o, "G1T");
  o->G2 = find(o, "G2"); o->G2T = find(o, "G2T");
  for (int k = 0; k < 29; k++) o->M[k] = find(o, solid3d_names[k]);
  for (int k = 0; k < 10; k++) {
    char nm[16];
    snprintf(nm, sizeof nm, "M0_w%d", k + 1);
    o->M0[k] = find(o, nm);
  }
  for (int k = 0; k < 4; k++) o->Mfl[k] = find(o, fluid3d_names[k]);
  o->M0fl = find(o, "M0_w_fl");
  o->inv_mass_rho = find(o, "inv_mass_rho");
  o->inv_mass_fluid = find(o, "inv_mass_fluid");
  o->gamma_s = find(o, "solid_absorbing_gamma");
  o->gamma_f = find(o, "fluid_absorbing_gamma");
  o->igloc_s = find(o, "igloc_solid"); o->igloc_f = find(o, "igloc_fluid");
  o->axis_s = find(o, "axis_solid"); o->axis_f = find(o, "axis_fluid");
  o->bdry_solid_el = find(o, "bdry_solid_el");
  o->bdry_fluid_el = find(o, "bdry_fluid_el");
  o->bdry_jpol_solid = find(o, "bdry_jpol_solid");
  o->bdry_jpol_fluid = find(o, "bdry_jpol_fluid");
  o->bdry_matr = find(o, "bdry_matr");
  o->fsmask = find(o, "fluid_free_surface_mask");
  o->ielsrc = find(o, "ielsrc"); o->source_term = find(o, "source_term");
  o->stf = find(o, "stf"); o->nstf = find_n(o, "stf");
  o->recfile_el = find(o, "recfile_el");
  if (o->anel && !o->cg) {
    o->Ycg = find(o, "Y"); o->Vseta = find(o, "V_s_eta");
    o->Vsxi = find(o, "V_s_xi"); o->Vzeta = find(o, "V_z_eta");
    o->Vzxi = find(o, "V_z_xi");
    o->Y0 = find(o, "Y0"); o->V0seta = find(o, "V0_s_eta");
    o->V0sxi = find(o, "V0_s_xi"); o->V0zeta = find(o, "V0_z_eta");
    o->V0zxi = find(o, "V0_z_xi");
    o->DsDeta = find(o, "DsDeta_over_J_sol");
    o->DzDeta = find(o, "DzDeta_over_J_sol");
    o->DsDxi = find(o, "DsDxi_over_J_sol");
    o->DzDxi = find(o, "DzDxi_over_J_sol");
    o->inv_s = find(o, "inv_s_solid");
    o->dmu = find(o, "delta_mu"); o->dkappa = find(o, "delta_kappa");
    o->expw = find(o, "exp_w_j_deltat"); o->tsfac_t = find(o, "ts_fac_t");
    o->tsfac_tm1 = find(o, "ts_fac_tm1"); o->aj_mu = find(o, "a_j_mu");
    o->aj_kappa = find(o, "a_j_kappa");
    if (!o->Ycg || !o->Y0 || !o->V0zxi || !o->DsDeta || !o->inv_s || !o->dkappa ||
        !o->expw || !o->aj_mu) {
      orc_free(o);
      return NULL;
    }
  } else if (o->anel) {
    o->Ycg = find(o, "Y_cg4"); o->Vseta = find(o, "V_s_eta_cg4");
    o->Vsxi = find(o, "V_s_xi_cg4"); o->Vzeta = find(o, "V_z_eta_cg4");
    o->Vzxi = find(o, "V_z_xi_cg4");
    o->DsDeta = find(o, "DsDeta_over_J_sol_cg4");
    o->DzDeta = find(o, "DzDeta_over_J_sol_cg4");
    o->DsDxi = find(o, "DsDxi_over_J_sol_cg4");
    o->DzDxi = find(o, "DzDxi_over_J_sol_cg4");
    o->inv_s = find(o, "inv_s_solid_cg4");
    o->dmu = find(o, "delta_mu_cg4"); o->dkappa = find(o, "delta_kappa_cg4");
    o->expw = find(o, "exp_w_j_deltat"); o->tsfac_t = find(o, "ts_fac_t");
    o->tsfac_tm1 = find(o, "ts_fac_tm1"); o->aj_mu = find(o, "a_j_mu");
    o->aj_kappa = find(o, "a_j_kappa");
    if (!o->Ycg || !o->DsDeta || !o->expw || !o->aj_mu) { orc_free(o); return NULL; }
  }
  if (!o->G0 || !o->inv_mass_rho || !o->igloc_s || !o->axis_s || !o->stf) {
    orc_free(o);
    return NULL;
  }
  /* ax_el lists: axial elements in ascending order (def_grid.f90:64-77) */
  o->ax_el_s = (int32_t *)malloc(sizeof(int32_t) * (o->nel_s + 1));
  o->ax_el_f = (int32_t *)malloc(sizeof(int32_t) * (o->nel_f + 1));
  for (int e = 0; e < o->nel_s; e++) if (o->axis_s[e]) o->ax_el_s[o->naxel_s++] = e;
  for (int e = 0; e < o->nel_f; e++) if (o->axis_f[e]) o->ax_el_f[o->naxel_f++] = e;
  size_t ns = (size_t)o->nel_s * NPP, nf = (size_t)o->nel_f * NPP;
  o->disp = zalloc(3 * ns); o->velo = zalloc(3 * ns);
  o->acc0 = zalloc(3 * ns); o->acc1 = zalloc(3 * ns);
  o->chi = zalloc(nf); o->dchi = zalloc(nf); o->ddchi0 = zalloc(nf); o->ddchi1 = zalloc(nf);
  o->gvec_s = zalloc(3 * (size_t)o->nglob_s); o->gvec_f = zalloc((size_t)o->nglob_f);
  if (o->anel) {
    o->memvar = zalloc(mv_per_el(o) * o->nel_s);
    o->src_tr_tm1 = zalloc(pts_per_el(o) * o->nel_s);
    o->src_dev_tm1 = zalloc(6 * pts_per_el(o) * o->nel_s);
  }
  o->nstages = 1;
  if (o->scheme != 0) {
    o->nstages = symp_coeffs(o->scheme, o->deltat, o->coefd, o->coefv, o->coeff);
    if (o->nstages <= 0) { orc_free(o); return NULL; }
  }
  o->iter = 0;
  o->t = 0.0;
  return o;
}

int orc_nel_solid(orc_t *o) { return o->nel_s; }
int orc_nel_fluid(orc_t *o) { return o->nel_f; }
int orc_iter(orc_t *o) { return o->iter; }
int orc_src_type(orc_t *o) { return o->src_type; }

/* Run steps it0..it1 (1-based, it0 == iter+1); write receiver samples for
 * every step with mod(iter, seis_it) == 0.  Returns the number of samples,
 * or -(iter) when the blow-up check of runtime_info fires. */
int orc_run(orc_t *o, int it0, int it1, float *seis) {
  enable_ftz();
  if (it0 != o->iter + 1 || it1 > o->niter) return -1000000;
  int ns = 0;
  for (int it = it0; it <= it1; it++) {
    o->t = o->t + o->deltat;
    if (blown_up(o)) return -it;
    o->iter = it;
    if (o->scheme == 0) newmark_step(o);
    else symplectic_step(o);
    if (o->num_rec > 0 && it % o->seis_it == 0) {
      if (seis) sample(o, seis + (size_t)ns * 3 * o->num_rec);
      ns++;
    }
  }
  return ns;
}

void orc_get_fields(orc_t *o, float *disp, float *velo, float *chi, float *dchi,
                    float *ddchi) {
  size_t ns = (size_t)o->nel_s * NPP, nf = (size_t)o->nel_f * NPP;
  if (disp) memcpy(disp, o->disp, 3 * ns * sizeof(float));
  if (velo) memcpy(velo, o->velo, 3 * ns * sizeof(float));
  if (chi) memcpy(chi, o->chi, nf * sizeof(float));
  if (dchi) memcpy(dchi, o->dchi, nf * sizeof(float));
  if (ddchi) memcpy(ddchi, o->ddchi0, nf * sizeof(float));
}

void orc_set_fields(orc_t *o, const float *disp, const float *velo,
                    const float *chi, const float *dchi, const float *ddchi) {
  size_t ns = (size_t)o->nel_s * NPP, nf = (size_t)o->nel_f * NPP;
  if (disp) memcpy(o->disp, disp, 3 * ns * sizeof(float));
  if (velo) memcpy(o->velo, velo, 3 * ns * sizeof(float));
  if (chi) memcpy(o->chi, chi, nf * sizeof(float));
  if (dchi) memcpy(o->dchi, dchi, nf * sizeof(float));
  if (ddchi) memcpy(o->ddchi0, ddchi, nf * sizeof(float));
}

/* kernel-level entry points for tests */
void orc_solid_stiffness(orc_t *o, const float *u, float *out) {
  enable_ftz();
  solid_stiffness(o, u, out);
}
void orc_fluid_stiffness(orc_t *o, const float *chi, float *out) {
  enable_ftz();
  stiff_fluid(o, chi, out);
}
void orc_anel_stiffness(orc_t *o, const float *memvar, float *acc) {
  enable_ftz();
  float *keep = o->memvar;
  o->memvar = (float *)memvar;
  anel_stiff(o, acc);
  o->memvar = keep;
}
void orc_memvar_update(orc_t *o, const float *disp) {
  enable_ftz();
  memvars(o, disp);
}
void orc_get_memvars(orc_t *o, float *memvar, float *tr, float *dev) {
  if (!o->anel) return;
  if (memvar) memcpy(memvar, o->memvar, mv_per_el(o) * o->nel_s * sizeof(float));
  if (tr) memcpy(tr, o->src_tr_tm1, pts_per_el(o) * o->nel_s * sizeof(float));
  if (dev) memcpy(dev, o->src_dev_tm1, 6 * pts_per_el(o) * o->nel_s * sizeof(float));
}
void orc_assemble_solid(orc_t *o, float *vec, int nc) {
  enable_ftz();
  pdistsum(o->igloc_s, o->gvec_s, o->nglob_s, vec, o->nel_s, nc);
}
void orc_assemble_fluid(orc_t *o, float *vec) {
  enable_ftz();
  pdistsum(o->igloc_f, o->gvec_f, o->nglob_f, vec, o->nel_f, 1);
}
int orc_symplectic_coefficients(int scheme, double dt, double *cd, double *cv,
                                double *cf) {
  return symp_coeffs(scheme, dt, cd, cv, cf);
}
void orc_mxm(const float *a, const float *b, float *c) { mxm(a, b, c); }

/* ------------------------------------------------------------------------ */
/* kernel wavefields, displ_only (SOLVER/meshes_io.F90, wavefields_io.f90)   */
/* ------------------------------------------------------------------------ */
/* build_kwf_grid (SOLVER/meshes_io.F90:489-610): the points ibeg..iend x
 * jbeg..jend of the elements selected by elem_mask (solid elements 1..nel_s,
 * then fluid elements as nel_s + 1..; NULL selects all), plus the axis point
 * ipol = 0 of each row of an axial element, numbered by the first appearance
 * of their global point: solid elements first, jpol outer, ipol inner, the
 * axis point after each row; then the fluid elements.  mapping[p + 25 e]
 * (e over solid then fluid elements) receives the 1-based number (0 where
 * the entry is not a kernel-wavefield point) and mask[p + 25 e] 1 at the
 * entry that numbered its point.  Returns the solid point count through
 * nps and the fluid count through npf. */
void orc_kwf_grid(orc_t *o, int ibeg, int iend, int jbeg, int jend,
                  const int32_t *elem_mask, int32_t *mapping, int32_t *mask,
                  int32_t *nps, int32_t *npf) {
  const int nel = o->nel_s + o->nel_f;
  const size_t ng = (size_t)o->nglob_s + o->nglob_f;
  int32_t *num = calloc(ng ? ng : 1, sizeof(int32_t)); /* 0: unseen */
  memset(mapping, 0, (size_t)nel * NPP * sizeof(int32_t));
  memset(mask, 0, (size_t)nel * NPP * sizeof(int32_t));
  int ct = 0;
  for (int dom = 0; dom < 2; dom++) {
    const int n = dom == 0 ? o->nel_s : o->nel_f;
    const int32_t *ig = dom == 0 ? o->igloc_s : o->igloc_f;
    const int32_t *ax = dom == 0 ? o->axis_s : o->axis_f;
    const int off = dom == 0 ? o->nglob_f : 0; /* idest: fluid ids first */
    const int eoff = dom == 0 ? 0 : o->nel_s;
    for (int e = 0; e < n; e++) {
      if (elem_mask && !elem_mask[eoff + e]) continue;
      for (int j = jbeg; j <= jend; j++) {
        for (int pass = 0; pass < 2; pass++) {
          const int i0 = pass == 0 ? ibeg : 0, i1 = pass == 0 ? iend : 0;
          if (pass == 1 && !ax[e]) break;
          for (int i = i0; i <= i1; i++) {
            const int p = i + NP * j;
            const int g = ig[(size_t)e * NPP + p] - 1 + off;
            if (!num[g]) {
              num[g] = ++ct;
              mask[(size_t)(eoff + e) * NPP + p] = 1;
            }
            mapping[(size_t)(eoff + e) * NPP + p] = num[g];
          }
        }
      }
    }
    if (dom == 0) *nps = ct;
  }
  *npf = ct - *nps;
  free(num);
}

/* axisym_gradient_fluid (SOLVER/pointwise_derivatives.f90:509-544) of one
 * fluid element: grad(:,:,1) = dzdeta mxm1 + dzdxi mxm2, grad(:,:,2) =
 * dsdeta mxm1 + dsdxi mxm2 with mxm1 = G1T f (axial) or G2T f, mxm2 = f G2;
 * every term an over-J array of the element. */
static void grad_fluid(orc_t *o, const float *f, int e, const float *dsdeta,
                       const float *dzdeta, const float *dsdxi,
                       const float *dzdxi, float *gs, float *gz) {
  float m1[NPP], m2[NPP];
  mxm(o->axis_f[e] ? o->G1T : o->G2T, f, m1);
  mxm(f, o->G2, m2);
  const size_t b = (size_t)e * NPP;
  for (int p = 0; p < NPP; p++) {
    gs[p] = dzdeta[b + p] * m1[p] + dzdxi[b + p] * m2[p];
    gz[p] = dsdeta[b + p] * m1[p] + dsdxi[b + p] * m2[p];
  }
}

/* dump_disp_global with dump_type displ_only (SOLVER/wavefields_io.f90:
 * 1019-1097) for the current fields: the solid displacement (dipole: s =
 * u1 + u2, phi = u1 - u2, z = u3; else u as it is), with every entry of the
 * source elements ielsrc(1:nelsrc) zeroed when src_mask (src_dump_type
 * 'mask', eradicate_src_elem_vec_values :1123-1135), gathered by
 * kwf_mapping_sol (:743-760); the fluid displacement inv_rho_fluid * grad
 * chi (s, 0, z) gathered by kwf_mapping_flu (:764-781).  out[c][k] holds
 * component c (monopole: s, z; else s, phi, z) of kernel-wavefield point
 * k + 1 (solid points first).  Returns -1 when the fluid arrays are missing
 * from the state. */
int orc_kwf_sample(orc_t *o, const int32_t *mapping, const int32_t *mask,
                   int nps, int npf, int src_mask, float *out) {
  enable_ftz();
  const int npt = nps + npf;
  const int mono = o->src_type == MONO;
  const int nc = mono ? 2 : 3;
  const size_t ns = (size_t)o->nel_s * NPP;
  float *f = malloc(3 * (ns ? ns : 1) * sizeof(float));
  for (size_t q = 0; q < ns; q++) {
    const float u1 = o->disp[q], u2 = o->disp[ns + q], u3 = o->disp[2 * ns + q];
    if (o->src_type == DI) {
      f[q] = u1 + u2; f[ns + q] = u1 - u2; f[2 * ns + q] = u3;
    } else {
      f[q] = u1; f[ns + q] = u2; f[2 * ns + q] = u3;
    }
  }
  if (src_mask)
    for (int k = 0; k < o->nelsrc; k++) {
      const size_t b = (size_t)(o->ielsrc[k] - 1) * NPP;
      for (int c = 0; c < 3; c++)
        for (int p = 0; p < NPP; p++) f[c * ns + b + p] = 0.0f;
    }
  /* solid: components 1, (2), 3 of f */
  for (int e = 0; e < o->nel_s; e++)
    for (int p = 0; p < NPP; p++) {
      const size_t q = (size_t)e * NPP + p;
      if (!mask[q]) continue;
      const int k = mapping[q] - 1;
      for (int c = 0, oc = 0; c < 3; c++) {
        if (mono && c == 1) continue;
        out[(size_t)oc++ * npt + k] = f[c * ns + q];
      }
    }
  free(f);
  if (o->nel_f > 0 && npf > 0) {
    const float *rho = find(o, "inv_rho_fluid"), *dsdeta = find(o, "DsDeta_over_J_flu"),
                *dzdeta = find(o, "DzDeta_over_J_flu"), *dsdxi = find(o, "DsDxi_over_J_flu"),
                *dzdxi = find(o, "DzDxi_over_J_flu");
    if (!rho || !dsdeta || !dzdeta || !dsdxi || !dzdxi) return -1;
    for (int e = 0; e < o->nel_f; e++) {
      const int32_t *mk = mask + (size_t)(o->nel_s + e) * NPP;
      const int32_t *mp = mapping + (size_t)(o->nel_s + e) * NPP;
      int any = 0;
      for (int p = 0; p < NPP; p++) any |= mk[p];
      if (!any) continue;
      float gs[NPP], gz[NPP];
      grad_fluid(o, o->chi + (size_t)e * NPP, e, dsdeta, dzdeta, dsdxi, dzdxi, gs, gz);
      for (int p = 0; p < NPP; p++) {
        if (!mk[p]) continue;
        const int k = mp[p] - 1; /* the fluid points follow the solid ones */
        const float r = rho[(size_t)e * NPP + p];
        out[k] = r * gs[p];
        if (!mono) out[(size_t)npt + k] = 0.0f;
        out[(size_t)(nc - 1) * npt + k] = r * gz[p];
      }
    }
  }
  return 0;
}
