// axh_device.hip -- device layout and gfx950 kernels of the AxiSEM time loop.
//
// Layout in HBM (one GPU, whole mesh resident):
//  * Fields are stored once per GLOBAL point (disp/velo/acc0 [3][nglob_s],
//    chi/dchi/ddchi [nglob_f]) instead of once per element-local point as in
//    the reference; every copy of a point in the reference holds the same
//    value (checked at build), so this is the same state with 16/25 of the
//    field traffic.
//  * Element data is grouped in tiles of 10 elements = 250 of 256 lanes of a
//    workgroup ("slots"); slot = tile*256 + (e%10)*25 + i + 5 j.  The
//    precomputed stiffness terms are packed [tile][term][256], so each term of
//    a tile is one contiguous 1 KB row and every load of a wavefront is
//    coalesced.
//  * Assembly (pdistsum, SOLVER/commun.F90:69-283) is a gather: every global
//    point sums its element contributions through a CSR list ordered by
//    element number, which is the reference's summation order, so the FP32
//    sums are the reference's bit for bit.  No atomics.
//  * Arithmetic order follows the reference expression by expression; the
//    file is compiled with -ffp-contract=off (no FMA), and dp constants enter
//    in double precision where the Fortran promotes to dp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "axh_internal.h"

#pragma clang fp contract(off)

namespace axh {

#define HIPCHECK(x)                                                          \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess)                                                    \
      fail(AXH_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));     \
  } while (0)

// number of packed element terms per source type (order: see pack_terms)
constexpr int kNcMono = 15, kNcDi = 24, kNcQuad = 26;
constexpr int kNcFluidMono = 3, kNcFluidMulti = 4;
constexpr int kCg4Terms = 12;  // packed cg4 coefficients per element

// flag bits per global point
constexpr uint8_t kFlagAxis = 1;      // on the symmetry axis (masked)
constexpr uint8_t kFlagSource = 2;    // carries a source term (solid)
constexpr uint8_t kFlagBlow = 4;      // GLL point (1,1) of an element
constexpr uint8_t kFlagSurface = 8;   // fluid free surface (masked)
constexpr uint8_t kFlagAxisBdry = 16; // axis point on the solid-fluid boundary
constexpr int kMaxAxisBdry = 16;

// kernel classes for timing
enum KClass { KC_PRED = 0, KC_FLUID, KC_COUPLE_F, KC_SOLID, KC_ASM_F,
              KC_COUPLE_S, KC_ASM_S, KC_DRIFT, KC_MEMVAR, KC_REC, KC_N };
static const char *const kClassName[KC_N] = {
    "predict", "fluid_stiffness", "copy2fluid", "solid_stiffness",
    "fluid_assemble_update", "copy2solid", "solid_assemble_update",
    "final_drift", "memvar_update", "receivers"};

template <class T> struct DBuf {
  T *p = nullptr;
  size_t n = 0;
};

struct Device {
  hipStream_t stream = nullptr;
  int src = 0, scheme = 0, anel = 0, nsls = 0, nstages = 1;
  int nel_s = 0, nel_f = 0, ntile_s = 0, ntile_f = 0;
  int nglob_s = 0, nglob_f = 0, nbdry = 0, nrec = 0, nsrcpts = 0, naxb = 0;
  double coefd[16] = {0}, coefv[16] = {0};
  double expw[kMaxSls] = {0}, tst[kMaxSls] = {0}, tstm1[kMaxSls] = {0};
  std::vector<float> stf;  // host copy of the stf table
  size_t bytes = 0;
  // constants
  float *G = nullptr;  // G0 | G1 | G1T | G2 | G2T
  // solid
  float *coef_s = nullptr, *m0_s = nullptr, *minv_s = nullptr;
  int *einfo_s = nullptr, *gid_s = nullptr, *gstart_s = nullptr, *gslot_s = nullptr;
  uint8_t *flags_s = nullptr;
  float *disp = nullptr, *velo = nullptr, *acc0 = nullptr, *force_s = nullptr;
  int *src_g = nullptr;
  float *src_val = nullptr;
  // fluid
  float *coef_f = nullptr, *m0_f = nullptr, *minv_f = nullptr;
  int *einfo_f = nullptr, *gid_f = nullptr, *gstart_f = nullptr, *gslot_f = nullptr;
  uint8_t *flags_f = nullptr;
  float *chi = nullptr, *dchi = nullptr, *ddchi = nullptr, *force_f = nullptr;
  int *srcel_f = nullptr;     // per fluid element: source slot k or -1
  float *src_fl = nullptr;    // source_term_fl (25 x 8)
  // coupling
  int *cpl_ss = nullptr, *cpl_gs = nullptr, *cpl_sf = nullptr, *cpl_gf = nullptr;
  float *cpl_bm = nullptr;    // [nbdry*5][2]
  int *cpl_axb = nullptr;     // side-buffer slot of a masked axis point or -1
  int *axb_g = nullptr;       // global ids of the masked axis boundary points
  float *axb_val = nullptr;   // their unmasked predicted displacement [naxb][3]
  // attenuation
  float *cg4 = nullptr, *memvar = nullptr, *rsum = nullptr, *trm1 = nullptr,
        *devm1 = nullptr;
  double *aj_mu = nullptr, *aj_k = nullptr;
  // receivers
  int *rec_g = nullptr;
  float *seis_d = nullptr;
  size_t seis_cap = 0;
  int *blow = nullptr;
  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool profile = false;
  std::vector<hipEvent_t> kev;      // pairs
  std::vector<int> kev_class;
  double kms[KC_N] = {0};
  long kcount[KC_N] = {0};
  std::vector<void *> allocs;

  template <class T> T *alloc(size_t n) {
    if (n == 0) n = 1;
    void *p = nullptr;
    HIPCHECK(hipMalloc(&p, n * sizeof(T)));
    allocs.push_back(p);
    bytes += n * sizeof(T);
    return (T *)p;
  }
  template <class T> T *upload(const std::vector<T> &v) {
    T *p = alloc<T>(v.size());
    if (!v.empty())
      HIPCHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
  }
  template <class T> T *zeros(size_t n) {
    T *p = alloc<T>(n);
    HIPCHECK(hipMemset(p, 0, (n ? n : 1) * sizeof(T)));
    return p;
  }
};

// ---------------------------------------------------------------------------
// kernel argument blocks
// ---------------------------------------------------------------------------
struct SolidArgs {
  int nel, nsls, iter_unused;
  const float *G, *coef, *m0;
  const int *einfo, *gid;
  const float *disp;      // [3][nglob]
  int nglob;
  float *force;           // [3][nslot]
  int nslot;
  // anelastic
  const float *cg4;       // [nel][12][4]
  float *memvar;          // [nel][nsls][6][4]
  float *rsum;            // [nel][6][4]
  float *trm1, *devm1;    // [nel][4], [nel][6][4]
  const double *aj_mu, *aj_k;  // [nel][nsls]
  double expw[kMaxSls], tst[kMaxSls], tstm1[kMaxSls];
};

struct FluidArgs {
  int nel, multipole, fluid_src;
  const float *G, *coef, *m0;
  const int *einfo, *gid, *srcel;
  const float *src_fl;
  float stf;
  const float *chi;
  float *force;
};

// ---------------------------------------------------------------------------
// 5x5 products on LDS tiles (mxm_4 / vxm_4, SOLVER/unrolled_loops.f90:164-206)
// ---------------------------------------------------------------------------
// Spectral operators in LDS (105 floats, copied once per workgroup); lanes
// index them with their own (i, j), which keeps the register footprint small.
struct GRegs {
  const float *gT;   // (axial ? G1T : G2T), column major
  const float *g;    // (axial ? G1 : G2)
  const float *g2;   // G2
  const float *g2T;  // G2T
  const float *g0;   // G0
  int i, j;
  float g0i;         // G0(i)
};

__device__ __forceinline__ void load_g(const float *sG, bool axial, int i, int j, GRegs &r) {
  r.g0 = sG;
  r.gT = axial ? sG + 30 : sG + 80;
  r.g = axial ? sG + 5 : sG + 55;
  r.g2 = sG + 55;
  r.g2T = sG + 80;
  r.i = i;
  r.j = j;
  r.g0i = sG[i];
}

__device__ __forceinline__ void stage_g(const float *G, float *sG) {
  for (int t = threadIdx.x; t < 105; t += blockDim.x) sG[t] = G[t];
}

// (A M)(i,j) = sum_k A(i,k) M(k,j)  for a 5x5 operator A in LDS
__device__ __forceinline__ float amul(const float *A, const float *M, int i, int j) {
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) s = s + A[i + 5 * k] * M[k + 5 * j];
  return s;
}
// (M B)(i,j) = sum_k M(i,k) B(k,j)  for a 5x5 operator B in LDS
__device__ __forceinline__ float mulb(const float *M, const float *B, int i, int j) {
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) s = s + M[i + 5 * k] * B[k + 5 * j];
  return s;
}
// sum_k v(k) B(k,j)
__device__ __forceinline__ float vmul(const float *v, const float *B, int j) {
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) s = s + v[k] * B[k + 5 * j];
  return s;
}

// ---------------------------------------------------------------------------
// coarse-grained strain and memory variables (SOLVER/attenuation.f90:81-202,
// 472-537; SOLVER/pointwise_derivatives.f90:55-75,216-246)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int cg_r(int q) { return (q < 2) ? 1 : 3; }
__device__ __forceinline__ int cg_c(int q) { return (q & 1) ? 3 : 1; }

// gradient of f at cg point q: returns (dsdf, dzdf)
template <class F>
__device__ __forceinline__ void grad_cg4(const float *G, bool axial,
                                         const float *cg, int q, F f,
                                         float &gs, float &gz) {
  const float *GT = axial ? (G + 30) : (G + 80);
  const float *G2 = G + 55;
  const int r = cg_r(q), c = cg_c(q);
  float m1 = 0.0f, m2 = 0.0f;
  // mxm_cg4_sparse_c(GT, f): row r of GT times column c of f
#pragma unroll
  for (int k = 0; k < 5; k++) m1 = m1 + GT[r + 5 * k] * f(k + 5 * c);
  // mxm_cg4_sparse_c(f, G2): row r of f times column c of G2
#pragma unroll
  for (int k = 0; k < 5; k++) m2 = m2 + f(r + 5 * k) * G2[k + 5 * c];
  // cg layout [12][4]: 5 DsDeta, 6 DzDeta, 7 DsDxi, 8 DzDxi
  gs = cg[6 * 4 + q] * m1 + cg[8 * 4 + q] * m2;
  gz = cg[5 * 4 + q] * m1 + cg[7 * 4 + q] * m2;
}

// Voigt strain E1..E6 at cg point q from the element's displacement in LDS.
template <int SRC>
__device__ void strain_cg4(const float *G, bool axial, const float *cg, int q,
                           const float *U1, const float *U2, const float *U3,
                           float g[6]) {
  const float *invs = cg + 9 * 4;
  const int p = cg_r(q) + 5 * cg_c(q);
  float b1s, b1z, b2s, b2z;
#pragma unroll
  for (int c = 0; c < 6; c++) g[c] = 0.0f;
  if (SRC == AXH_DIPOLE)
    grad_cg4(G, axial, cg, q, [&](int k) { return U1[k] + U2[k]; }, b1s, b1z);
  else
    grad_cg4(G, axial, cg, q, [&](int k) { return U1[k]; }, b1s, b1z);
  grad_cg4(G, axial, cg, q, [&](int k) { return U3[k]; }, b2s, b2z);
  g[0] = b1s;
  g[2] = b2z;
  g[4] = b1z + b2s;
  if (SRC == AXH_MONOPOLE) {
    g[1] = invs[q] * U1[p];
  } else if (SRC == AXH_DIPOLE) {
    g[1] = 2.0f * (invs[q] * U2[p]);
    grad_cg4(G, axial, cg, q, [&](int k) { return U1[k] - U2[k]; }, b1s, b1z);
    float f3 = invs[q] * U3[p];
    g[3] = -f3 - b1z;
    g[5] = -g[1] - b1s;
  } else {
    g[1] = invs[q] * (U1[p] - 2.0f * U2[p]);
    grad_cg4(G, axial, cg, q, [&](int k) { return U2[k]; }, b1s, b1z);
    float f3 = invs[q] * U3[p];
    float f6 = invs[q] * (U2[p] - 2.0f * U1[p]);
    g[3] = -(2.0f * f3) - b1z;
    g[5] = f6 - b1s;
  }
}

// source terms of the memory-variable update at cg point q
template <int SRC>
__device__ __forceinline__ void memvar_sources(const float *cg, int q,
                                               const float g[6], float &str,
                                               float sdev[6]) {
  const double third = 1.0 / 3.0;
  const float dm = cg[10 * 4 + q], dk = cg[11 * 4 + q];
  const float tr = g[0] + g[1] + g[2];
  str = dk * tr;
#pragma unroll
  for (int c = 0; c < 6; c++) sdev[c] = 0.0f;
#pragma unroll
  for (int c = 0; c < 3; c++)
    sdev[c] = (float)((double)(dm * 2.0f) * ((double)g[c] - (double)tr * third));
  sdev[4] = dm * g[4];
  if (SRC != AXH_MONOPOLE) {
    sdev[3] = dm * g[3];
    sdev[5] = dm * g[5];
  }
}

// Update of the memory variables of one element.  Work of the element is
// spread over its lanes: lane k handles entries idx = k, k+25, ... of
// (q, c, j) = (idx%4, (idx/4)%6, idx/24).  mv/old in LDS, src_* in LDS.
template <int SRC>
__device__ __forceinline__ void memvar_update_elem(
    const SolidArgs &A, int e, int k, const float *mv_lds, const float *src_lds,
    const float *old_lds, float *rs_lds) {
  const int nsls = A.nsls;
  const int n = 24 * nsls;
  for (int idx = k; idx < n; idx += 25) {
    const int q = idx & 3, c = (idx >> 2) % 6, jj = idx / 24;
    if (SRC == AXH_MONOPOLE && (c == 3 || c == 5)) continue;
    const double Tt = A.tst[jj], Tm = A.tstm1[jj], E = A.expw[jj];
    const double amu = A.aj_mu[(size_t)e * nsls + jj];
    const double m = (double)mv_lds[idx];
    const float db = (float)(Tt * amu * (double)src_lds[4 + 4 * c + q] +
                             Tm * amu * (double)old_lds[4 + 4 * c + q]);
    float nv;
    if (c < 3) {
      const double ak = A.aj_k[(size_t)e * nsls + jj];
      const float trb = (float)(Tt * ak * (double)src_lds[q] +
                                Tm * ak * (double)old_lds[q]);
      nv = (float)((E * m + (double)db) + (double)trb);
    } else {
      nv = (float)(E * m + (double)db);
    }
    A.memvar[(size_t)e * n + idx] = nv;
    if (rs_lds) rs_lds[idx] = nv;
  }
}

// ---------------------------------------------------------------------------
// solid stiffness: glob_stiffness_{mono,di,quad}_4 (SOLVER/stiffness_mono.f90:
// 60-157, stiffness_di.f90:60-256, stiffness_quad.f90:238-412) followed by
// glob_anel_stiffness_*_cg4 and, for newmark2, time_step_memvars_cg4.
// MODE 0: newmark2 (memory variables read, used, updated)
// MODE 1: symplectic stage (sums of memory variables read)
// ---------------------------------------------------------------------------
template <int SRC, int MODE, bool ANEL>
__global__ __launch_bounds__(256) void k_solid(SolidArgs A) {
  constexpr int NCOMP = (SRC == AXH_MONOPOLE) ? 2 : 3;
  constexpr int NC = (SRC == AXH_MONOPOLE) ? kNcMono
                     : (SRC == AXH_DIPOLE) ? kNcDi : kNcQuad;
  constexpr int NS = (SRC == AXH_MONOPOLE) ? 4 : 6;
  __shared__ float sU[NCOMP][256];
  __shared__ float sS[NS][256];
  __shared__ float sV[kElemPerTile][8];
  __shared__ float sMv[ANEL ? (MODE == 0 ? kElemPerTile * 24 * kMaxSls : 1) : 1];
  __shared__ float sR[ANEL ? kElemPerTile * 24 : 1];
  __shared__ float sSa[ANEL ? kElemPerTile * 24 : 1];
  __shared__ float sSrc[ANEL ? kElemPerTile * 28 : 1];
  __shared__ float sOld[ANEL ? kElemPerTile * 28 : 1];
  __shared__ float sG[105];

  const int t = threadIdx.x, tile = blockIdx.x;
  const int le = t / 25, k = t - 25 * le;
  const int i = k % 5, j = k / 5;
  const int e = tile * kElemPerTile + le;
  const bool valid = (t < 250) && (e < A.nel);
  const size_t slot = (size_t)tile * 256 + t;
  const int ax = valid ? A.einfo[e] : -1;
  const bool axial = ax >= 0;

  stage_g(A.G, sG);
  // gather the (predicted, masked) displacement of the element
  if (valid) {
    const int g = A.gid[slot];
    if (SRC == AXH_MONOPOLE) {
      sU[0][t] = A.disp[g];
      sU[1][t] = A.disp[2 * (size_t)A.nglob + g];
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) sU[c][t] = A.disp[c * (size_t)A.nglob + g];
    }
  } else {
#pragma unroll
    for (int c = 0; c < NCOMP; c++) sU[c][t] = 0.0f;
  }
  if (ANEL && MODE == 0) {
    // memory variables of the tile (contiguous per element)
    const int n = 24 * A.nsls;
    for (int idx = k; idx < n && valid; idx += 25)
      sMv[le * 24 * kMaxSls + idx] = A.memvar[(size_t)e * n + idx];
  }
  __syncthreads();

  GRegs G;
  load_g(sG, axial, i, j, G);
  const float *coef = A.coef + (size_t)tile * NC * 256 + t;
  float acc[3];

  if (SRC == AXH_MONOPOLE) {
    const float *U0 = &sU[0][le * 25], *U1 = &sU[1][le * 25];
    const float us = U0[k], uz = U1[k];
    const float X1 = amul(G.gT, U0, i, j), X2 = amul(G.gT, U1, i, j);
    const float X3 = mulb(U0, G.g2, i, j), X4 = mulb(U1, G.g2, i, j);
    const float m1 = coef[0 * 256], m2 = coef[1 * 256], m3 = coef[2 * 256],
                m4 = coef[3 * 256], mw1 = coef[4 * 256], m11s = coef[5 * 256],
                m21s = coef[6 * 256], m41s = coef[7 * 256], m12s = coef[8 * 256],
                m22s = coef[9 * 256], m32s = coef[10 * 256], m42s = coef[11 * 256],
                m11z = coef[12 * 256], m21z = coef[13 * 256], m41z = coef[14 * 256];
    float ls = m4 * X4 + m2 * X3 + m1 * X1 + m3 * X2 + us * mw1;
    sS[0][t] = m11s * X3 + m21s * X1 + m12s * X4 + m22s * X2 + m1 * us;
    sS[1][t] = m11s * X1 + m41s * X3 + m32s * X2 + m42s * X4 + m2 * us;
    sS[2][t] = m11z * X4 + m21z * X2 + m32s * X3 + m22s * X1 + m3 * us;
    sS[3][t] = m11z * X2 + m41z * X4 + m12s * X1 + m42s * X3 + m4 * us;
    __syncthreads();
    const float *S1s = &sS[0][le * 25], *S2s = &sS[1][le * 25],
                *S1z = &sS[2][le * 25], *S2z = &sS[3][le * 25];
    const float Y2 = mulb(S2s, G.g2T, i, j), Y4 = mulb(S2z, G.g2T, i, j);
    const float Y1 = amul(G.g, S1s, i, j), Y3 = amul(G.g, S1z, i, j);
    ls = ls + Y1 + Y2;
    float lz = Y3 + Y4;
    // axial terms with M0_w1..3
    float w1 = 0, w2 = 0, w3 = 0, V1 = 0, V2 = 0, V3 = 0;
    if (axial) {
      const float *m0 = A.m0 + (size_t)ax * 50;
      w1 = m0[0 * 5 + j]; w2 = m0[1 * 5 + j]; w3 = m0[2 * 5 + j];
      V1 = vmul(G.g0, U0, j);        // vxm(G0, us)
      V2 = mulb(U1, G.g2, 0, j);       // vxm(uz(0,:), G2)
      V3 = vmul(G.g0, U1, j);        // vxm(G0, uz)
      if (i == 0) sV[le][j] = w3 * V1;
    }
    __syncthreads();
    if (axial) {
      float V4 = w1 * V1 + w3 * V2;
      V4 = V4 + w2 * V3;
      float X2a = G.g0i * (w2 * V1);
      const float V1b = vmul(&sV[le][0], G.g2T, j);  // vxm(w3*V1, G2T)
      if (i == 0) X2a = X2a + V1b;
      ls = ls + G.g0i * V4;
      lz = X2a + lz;
    }
    acc[0] = ls;
    acc[1] = 0.0f;
    acc[2] = lz;
  } else if (SRC == AXH_DIPOLE) {
    const float *U0 = &sU[0][le * 25], *U1 = &sU[1][le * 25], *U2 = &sU[2][le * 25];
    const float u1 = U0[k], u2 = U1[k], u3 = U2[k];
    const float X4 = mulb(U0, G.g2, i, j), X5 = mulb(U1, G.g2, i, j),
                X6 = mulb(U2, G.g2, i, j);
    const float X1 = amul(G.gT, U0, i, j), X2 = amul(G.gT, U1, i, j),
                X3 = amul(G.gT, U2, i, j);
    const float X7 = X1 + X2, X8 = X4 + X5;
    const float m1 = coef[0 * 256], m2 = coef[1 * 256], m3 = coef[2 * 256],
                m4 = coef[3 * 256], m5 = coef[4 * 256], m6 = coef[5 * 256],
                m7 = coef[6 * 256], m8 = coef[7 * 256], mw1 = coef[8 * 256],
                mw2 = coef[9 * 256], mw3 = coef[10 * 256], m11s = coef[11 * 256],
                m21s = coef[12 * 256], m41s = coef[13 * 256], m12s = coef[14 * 256],
                m22s = coef[15 * 256], m42s = coef[16 * 256], m13s = coef[17 * 256],
                m23s = coef[18 * 256], m33s = coef[19 * 256], m43s = coef[20 * 256],
                m11z = coef[21 * 256], m21z = coef[22 * 256], m41z = coef[23 * 256];
    const float ls2 = m8 * X6 + m7 * X3 + m1 * X1 + m5 * X2 + m2 * X4 + m6 * X5 +
                      mw1 * u2 + mw2 * u3;
    const float ls3 = m4 * X4 - m4 * X5 + m3 * X1 - m3 * X2 + mw2 * u2 + mw3 * u3;
    float c1 = m13s * X6, c2 = m23s * X3, c3 = m3 * u3;
    sS[0][t] = c1 + c2 + c3 + m11s * X4 + m21s * X1 + m12s * X5 + m22s * X2 + m1 * u2;
    sS[1][t] = c1 + c2 - c3 + m11s * X5 + m21s * X2 + m12s * X4 + m22s * X1 + m5 * u2;
    c1 = m33s * X3;
    c2 = m43s * X6;
    c3 = m4 * u3;
    sS[2][t] = c1 + c2 + c3 + m11s * X1 + m41s * X4 + m12s * X2 + m42s * X5 + m2 * u2;
    sS[3][t] = c1 + c2 - c3 + m11s * X2 + m41s * X5 + m12s * X1 + m42s * X4 + m6 * u2;
    sS[4][t] = m33s * X8 + m23s * X7 + m11z * X6 + m21z * X3 + m7 * u2;
    sS[5][t] = m13s * X7 + m43s * X8 + m11z * X3 + m41z * X6 + m8 * u2;
    __syncthreads();
    const float *S1p = &sS[0][le * 25], *S1m = &sS[1][le * 25],
                *S2p = &sS[2][le * 25], *S2m = &sS[3][le * 25],
                *S1z = &sS[4][le * 25], *S2z = &sS[5][le * 25];
    const float Y1 = amul(G.g, S1p, i, j), Y3 = amul(G.g, S1m, i, j), Y5 = amul(G.g, S1z, i, j);
    const float Y2 = mulb(S2p, G.g2T, i, j), Y4 = mulb(S2m, G.g2T, i, j),
                Y6 = mulb(S2z, G.g2T, i, j);
    float l1 = Y1 + Y2;
    float l2 = Y3 + Y4 + ls2;
    float l3 = Y5 + Y6 + ls3;
    float a = 0, b = 0, c = 0, w2 = 0, w4 = 0, w6 = 0, w8 = 0, V2 = 0, V3 = 0;
    if (axial) {
      const float *m0 = A.m0 + (size_t)ax * 50;
      const float w1 = m0[0 * 5 + j], w3 = m0[2 * 5 + j], w7 = m0[6 * 5 + j],
                  w9 = m0[8 * 5 + j], w10 = m0[9 * 5 + j];
      w2 = m0[1 * 5 + j]; w4 = m0[3 * 5 + j]; w6 = m0[5 * 5 + j]; w8 = m0[7 * 5 + j];
      const float V1 = vmul(G.g0, U0, j);
      V2 = vmul(G.g0, U1, j);
      V3 = vmul(G.g0, U2, j);
      const float V4 = mulb(U0, G.g2, 0, j);   // vxm(u1(0,:), G2)
      a = w1 * V2 + w3 * V3;
      b = w1 * V1 + (w2 + w6) * V4 + w9 * V2 + w10 * V3;
      c = w3 * V1 + (w4 + w8) * V4 + w7 * V3 + w10 * V2;
      if (i == 0) sV[le][j] = (w2 + w6) * V2 + (w4 + w8) * V3;
    }
    __syncthreads();
    if (axial) {
      float S1pa = G.g0i * a;
      const float S1ma = G.g0i * b, S1za = G.g0i * c;
      const float V1b = vmul(&sV[le][0], G.g2T, j);  // vxm(V4, G2T)
      if (i == 0) S1pa = S1pa + V1b;
      l1 = l1 + S1pa;
      l2 = l2 + S1ma;
      l3 = l3 + S1za;
    }
    acc[0] = l1;
    acc[1] = l2;
    acc[2] = l3;
  } else {
    const float *U0 = &sU[0][le * 25], *U1 = &sU[1][le * 25], *U2 = &sU[2][le * 25];
    const float us = U0[k], up = U1[k], uz = U2[k];
    const float X1 = amul(G.gT, U0, i, j), X2 = amul(G.gT, U1, i, j),
                X3 = amul(G.gT, U2, i, j);
    const float X4 = mulb(U0, G.g2, i, j), X5 = mulb(U1, G.g2, i, j),
                X6 = mulb(U2, G.g2, i, j);
    const float m1 = coef[0 * 256], m2 = coef[1 * 256], m3 = coef[2 * 256],
                m4 = coef[3 * 256], m5 = coef[4 * 256], m6 = coef[5 * 256],
                m7 = coef[6 * 256], m8 = coef[7 * 256], mw1 = coef[8 * 256],
                mw2 = coef[9 * 256], mw3 = coef[10 * 256], mw4 = coef[11 * 256],
                mw5 = coef[12 * 256], m11s = coef[13 * 256], m21s = coef[14 * 256],
                m41s = coef[15 * 256], m12s = coef[16 * 256], m22s = coef[17 * 256],
                m32s = coef[18 * 256], m42s = coef[19 * 256], m11z = coef[20 * 256],
                m21z = coef[21 * 256], m41z = coef[22 * 256], m1phi = coef[23 * 256],
                m2phi = coef[24 * 256], m4phi = coef[25 * 256];
    const float c1 = m2 * X4, c2 = m1 * X1, c3 = m6 * X5, c4 = m5 * X2,
                c5 = m4 * X6, c6 = m3 * X3;
    float ls = c1 + c2 + 2.0f * (c3 + c4) + c5 + c6 + mw1 * us + mw2 * up +
               2.0f * mw3 * uz;
    float lp = -2.0f * (c1 + c2 + c5 + c6) - (c3 + c4) + mw2 * us + mw4 * up -
               mw3 * uz;
    float lz = 2.0f * (m8 * X5 + m7 * X2) + mw3 * (2.0f * us - up) + mw5 * uz;
    const float usp = us - 2.0f * up, tus = 2.0f * us - up;
    sS[0][t] = m11s * X4 + m21s * X1 + m12s * X6 + m22s * X3 + m1 * usp;   // S1s
    sS[1][t] = m11s * X1 + m41s * X4 + m32s * X3 + m42s * X6 + m2 * usp;   // S2s
    sS[2][t] = m11z * X6 + m21z * X3 + m32s * X4 + m22s * X1 + m3 * usp;   // S1z
    sS[3][t] = m11z * X3 + m41z * X6 + m12s * X1 + m42s * X4 + m4 * usp;   // S2z
    sS[4][t] = m1phi * X5 + m2phi * X2 + m5 * tus + 2.0f * m7 * uz;        // S1phi
    sS[5][t] = m1phi * X2 + m4phi * X5 + m6 * tus + 2.0f * m8 * uz;        // S2phi
    __syncthreads();
    const float *S1s = &sS[0][le * 25], *S2s = &sS[1][le * 25],
                *S1z = &sS[2][le * 25], *S2z = &sS[3][le * 25],
                *S1p = &sS[4][le * 25], *S2p = &sS[5][le * 25];
    const float Y2 = mulb(S2s, G.g2T, i, j), Y4 = mulb(S2p, G.g2T, i, j),
                Y6 = mulb(S2z, G.g2T, i, j);
    const float Y1 = amul(G.g, S1s, i, j), Y3 = amul(G.g, S1p, i, j), Y5 = amul(G.g, S1z, i, j);
    ls = ls + Y1 + Y2;
    lp = lp + Y3 + Y4;
    lz = lz + Y5 + Y6;
    __syncthreads();  // uniform barrier count with the other source types
    if (axial) {
      const float *m0 = A.m0 + (size_t)ax * 50;
      const float w1 = m0[0 * 5 + j], w2 = m0[1 * 5 + j], w3 = m0[2 * 5 + j],
                  w4 = m0[3 * 5 + j], w5 = m0[4 * 5 + j], w6 = m0[5 * 5 + j];
      const float V1 = vmul(G.g0, U0, j), V2 = vmul(G.g0, U1, j),
                  V3 = vmul(G.g0, U2, j);
      const float a = w1 * V1 + w2 * V2 + w3 * V3;
      const float b = w2 * V1 + w4 * V2 + w5 * V3;
      const float c = w3 * V1 + w5 * V2 + w6 * V3;
      ls = ls + G.g0i * a;
      lp = lp + G.g0i * b;
      lz = lz + G.g0i * c;
    }
    acc[0] = ls;
    acc[1] = lp;
    acc[2] = lz;
  }

  if (ANEL) {
    // sums of the memory variables over the SLS, (c, q) per lane k < 24
    const int n = 24 * A.nsls;
    float *R = &sR[le * 24];
    if (valid && k < 24) {
      if (MODE == 0) {
        float s = 0.0f;
        for (int jj = 0; jj < A.nsls; jj++) s = s + sMv[le * 24 * kMaxSls + jj * 24 + k];
        R[k] = s;
      } else {
        R[k] = A.rsum[(size_t)e * 24 + k];
      }
    }
    const float *cg = A.cg4 + (size_t)(valid ? e : 0) * (kCg4Terms * 4);
    __syncthreads();
    float *Sa = &sSa[le * 24];
    if (valid && k < 4) {
      const int q = k;
      const float vse = cg[1 * 4 + q], vsx = cg[2 * 4 + q], vze = cg[3 * 4 + q],
                  vzx = cg[4 * 4 + q];
      const float r1 = R[0 * 4 + q], r3 = R[2 * 4 + q], r4 = R[3 * 4 + q],
                  r5 = R[4 * 4 + q], r6 = R[5 * 4 + q];
      if (SRC == AXH_MONOPOLE) {
        Sa[0 * 4 + q] = vze * r1 + vse * r5;
        Sa[1 * 4 + q] = vzx * r1 + vsx * r5;
        Sa[4 * 4 + q] = vze * r5 + vse * r3;
        Sa[5 * 4 + q] = vzx * r5 + vsx * r3;
      } else if (SRC == AXH_DIPOLE) {
        Sa[0 * 4 + q] = vze * (r1 - r6) + vse * (r5 - r4);
        Sa[1 * 4 + q] = vzx * (r1 - r6) + vsx * (r5 - r4);
        Sa[2 * 4 + q] = vze * (r1 + r6) + vse * (r5 + r4);
        Sa[3 * 4 + q] = vzx * (r1 + r6) + vsx * (r5 + r4);
        Sa[4 * 4 + q] = vze * r5 + vse * r3;
        Sa[5 * 4 + q] = vzx * r5 + vsx * r3;
      } else {
        Sa[0 * 4 + q] = vze * r1 + vse * r5;
        Sa[1 * 4 + q] = vzx * r1 + vsx * r5;
        Sa[2 * 4 + q] = vze * r6 + vse * r4;
        Sa[3 * 4 + q] = vzx * r6 + vsx * r4;
        Sa[4 * 4 + q] = vze * r5 + vse * r3;
        Sa[5 * 4 + q] = vzx * r5 + vsx * r3;
      }
    }
    __syncthreads();
    if (valid) {
      // mxm_cg4_sparse_b(G, S1) has columns 1,3; mxm_cg4_sparse_a(S2, G2T)
      // rows 1,3 (SOLVER/unrolled_loops.f90:76-121)
      auto spb = [&](int s1) -> float {
        if (j == 1) return G.g[i + 5] * Sa[s1 * 4 + 0] + G.g[i + 15] * Sa[s1 * 4 + 2];
        if (j == 3) return G.g[i + 5] * Sa[s1 * 4 + 1] + G.g[i + 15] * Sa[s1 * 4 + 3];
        return 0.0f;
      };
      auto spa = [&](int s2) -> float {
        if (i == 1) return Sa[s2 * 4 + 0] * G.g2T[1 + 5 * j] + Sa[s2 * 4 + 1] * G.g2T[3 + 5 * j];
        if (i == 3) return Sa[s2 * 4 + 2] * G.g2T[1 + 5 * j] + Sa[s2 * 4 + 3] * G.g2T[3 + 5 * j];
        return 0.0f;
      };
      const bool cgp = (i == 1 || i == 3) && (j == 1 || j == 3);
      const int q = cgp ? ((i == 1 ? 0 : 2) + (j == 1 ? 0 : 1)) : 0;
      const float yl = cg[0 * 4 + q];
      const float r2 = R[1 * 4 + q], r4 = R[3 * 4 + q], r6 = R[5 * 4 + q];
      if (SRC == AXH_MONOPOLE) {
        float l1 = spb(0) + spa(1);
        if (cgp) l1 = l1 + yl * r2;
        const float l3 = spb(4) + spa(5);
        acc[0] = acc[0] - l1;
        acc[2] = acc[2] - l3;
      } else if (SRC == AXH_DIPOLE) {
        const float lp = spb(0) + spa(1);
        float lm = spb(2) + spa(3);
        float lz = spb(4) + spa(5);
        if (cgp) {
          lm = lm + 2.0f * yl * (r2 - r6);
          lz = lz - yl * r4;
        }
        acc[0] = acc[0] - lp;
        acc[1] = acc[1] - lm;
        acc[2] = acc[2] - lz;
      } else {
        float ls = spb(0) + spa(1);
        float lp = -spb(2) - spa(3);
        float lz = spb(4) + spa(5);
        if (cgp) {
          ls = ls + yl * (r2 - 2.0f * r6);
          lp = lp + yl * (r6 - 2.0f * r2);
          lz = lz - 2.0f * yl * r4;
        }
        acc[0] = acc[0] - ls;
        acc[1] = acc[1] - lp;
        acc[2] = acc[2] - lz;
      }
    }
    if (MODE == 0) {
      // time_step_memvars_cg4 with the masked predicted displacement
      float *src = &sSrc[le * 28], *old = &sOld[le * 28];
      if (valid && k < 4) {
        const int q = k;
        float g[6], str, sdev[6];
        strain_cg4<SRC>(A.G, axial, cg, q, &sU[0][le * 25],
                        &sU[NCOMP == 2 ? 0 : 1][le * 25], &sU[NCOMP - 1][le * 25], g);
        memvar_sources<SRC>(cg, q, g, str, sdev);
        old[q] = A.trm1[(size_t)e * 4 + q];
#pragma unroll
        for (int c = 0; c < 6; c++) old[4 + 4 * c + q] = A.devm1[(size_t)e * 24 + 4 * c + q];
        src[q] = str;
#pragma unroll
        for (int c = 0; c < 6; c++) src[4 + 4 * c + q] = sdev[c];
        A.trm1[(size_t)e * 4 + q] = str;
#pragma unroll
        for (int c = 0; c < 6; c++) A.devm1[(size_t)e * 24 + 4 * c + q] = sdev[c];
      }
      __syncthreads();
      if (valid)
        memvar_update_elem<SRC>(A, e, k, &sMv[le * 24 * kMaxSls], src, old, nullptr);
      (void)n;
    }
  }

  if (valid) {
    A.force[slot] = acc[0];
    if (SRC != AXH_MONOPOLE) A.force[(size_t)A.nslot + slot] = acc[1];
    A.force[2 * (size_t)A.nslot + slot] = acc[2];
  }
}

// memory-variable update of the symplectic scheme, once per step after the
// final drift (SOLVER/time_evol_wave.F90:727-731); also refreshes the SLS sums
// read by the stage kernels.
template <int SRC>
__global__ __launch_bounds__(256) void k_memvar(SolidArgs A) {
  constexpr int NCOMP = (SRC == AXH_MONOPOLE) ? 2 : 3;
  __shared__ float sU[NCOMP][256];
  __shared__ float sMv[kElemPerTile * 24 * kMaxSls];
  __shared__ float sSrc[kElemPerTile * 28];
  __shared__ float sOld[kElemPerTile * 28];
  const int t = threadIdx.x, tile = blockIdx.x;
  const int le = t / 25, k = t - 25 * le;
  const int e = tile * kElemPerTile + le;
  const bool valid = (t < 250) && (e < A.nel);
  const size_t slot = (size_t)tile * 256 + t;
  const int n = 24 * A.nsls;
  if (valid) {
    const int g = A.gid[slot];
    if (SRC == AXH_MONOPOLE) {
      sU[0][t] = A.disp[g];
      sU[1][t] = A.disp[2 * (size_t)A.nglob + g];
    } else {
      for (int c = 0; c < 3; c++) sU[c][t] = A.disp[c * (size_t)A.nglob + g];
    }
    for (int idx = k; idx < n; idx += 25)
      sMv[le * 24 * kMaxSls + idx] = A.memvar[(size_t)e * n + idx];
  }
  __syncthreads();
  const bool axial = valid && A.einfo[e] >= 0;
  const float *cg = A.cg4 + (size_t)(valid ? e : 0) * (kCg4Terms * 4);
  float *src = &sSrc[le * 28], *old = &sOld[le * 28];
  if (valid && k < 4) {
    const int q = k;
    float g[6], str, sdev[6];
    strain_cg4<SRC>(A.G, axial, cg, q, &sU[0][le * 25],
                    &sU[NCOMP == 2 ? 0 : 1][le * 25], &sU[NCOMP - 1][le * 25], g);
    memvar_sources<SRC>(cg, q, g, str, sdev);
    old[q] = A.trm1[(size_t)e * 4 + q];
    for (int c = 0; c < 6; c++) old[4 + 4 * c + q] = A.devm1[(size_t)e * 24 + 4 * c + q];
    src[q] = str;
    for (int c = 0; c < 6; c++) src[4 + 4 * c + q] = sdev[c];
    A.trm1[(size_t)e * 4 + q] = str;
    for (int c = 0; c < 6; c++) A.devm1[(size_t)e * 24 + 4 * c + q] = sdev[c];
  }
  __syncthreads();
  float *mv = &sMv[le * 24 * kMaxSls];
  if (valid) memvar_update_elem<SRC>(A, e, k, mv, src, old, mv);
  __syncthreads();
  if (valid && k < 24) {
    float s = 0.0f;
    for (int jj = 0; jj < A.nsls; jj++) s = s + mv[jj * 24 + k];
    A.rsum[(size_t)e * 24 + k] = s;
  }
}

// ---------------------------------------------------------------------------
// fluid stiffness glob_fluid_stiffness_4 (SOLVER/stiffness_fluid.f90:139-216)
// + add_source_fl (SOLVER/time_evol_wave.F90:1062-1076)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fluid(FluidArgs A) {
  __shared__ float sC[256];
  __shared__ float sS[2][256];
  __shared__ float sG[105];
  const int t = threadIdx.x, tile = blockIdx.x;
  const int le = t / 25, k = t - 25 * le;
  const int i = k % 5, j = k / 5;
  const int e = tile * kElemPerTile + le;
  const bool valid = (t < 250) && (e < A.nel);
  const size_t slot = (size_t)tile * 256 + t;
  const int ax = valid ? A.einfo[e] : -1;
  const bool axial = ax >= 0;
  stage_g(A.G, sG);
  sC[t] = valid ? A.chi[A.gid[slot]] : 0.0f;
  __syncthreads();
  GRegs G;
  load_g(sG, axial, i, j, G);
  const int nc = A.multipole ? kNcFluidMulti : kNcFluidMono;
  const float *coef = A.coef + (size_t)tile * nc * 256 + t;
  const float *C = &sC[le * 25];
  const float c0 = C[k];
  const float X1 = amul(G.gT, C, i, j), X2 = mulb(C, G.g2, i, j);
  const float m1 = coef[0], m2 = coef[256], m4 = coef[512];
  sS[0][t] = m1 * X2 + m2 * X1;
  sS[1][t] = m1 * X1 + m4 * X2;
  __syncthreads();
  const float Y1 = amul(G.g, &sS[0][le * 25], i, j);
  const float Y2 = mulb(&sS[1][le * 25], G.g2T, i, j);
  float loc = Y1 + Y2;
  if (A.multipole) {
    loc = loc + coef[768] * c0;
    if (axial) {
      const float w = A.m0[(size_t)ax * 5 + j];
      const float V1 = vmul(G.g0, C, j);
      loc = loc + G.g0i * (w * V1);
    }
  }
  if (valid) {
    if (A.fluid_src && A.stf != 0.0f) {
      const int ks = A.srcel[e];
      if (ks >= 0) loc = loc - A.src_fl[ks * 25 + k] * A.stf;
    }
    A.force[slot] = loc;
  }
}

// ---------------------------------------------------------------------------
// predictor / drift per global point (SOLVER/time_evol_wave.F90:357-367,
// 597-605, 720-725) with the axis masks (SOLVER/apply_masks.f90) and the
// blow-up test of runtime_info (SOLVER/time_evol_wave.F90:1042)
// ---------------------------------------------------------------------------
struct PredArgs {
  int n, ncomp_mask0, ncomp_mask1;  // masked component range [m0, m1]
  int comps;                         // bit c set: component c is stepped
  int newmark, apply_mask, check_blow, iter;
  double dt, hdt2, cd, thr;
  float *u, *v;
  const float *a0;
  const uint8_t *flags;
  int stride;
  int *blow;
  int naxb;
  const int *axb_g;   // masked axis points on the solid-fluid boundary
  float *axb_val;     // their unmasked prediction [naxb][3]
};

__global__ __launch_bounds__(256) void k_predict(PredArgs A) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= A.n) return;
  const uint8_t fl = A.flags[g];
  for (int c = 0; c < 3; c++) {
    if (!(A.comps & (1 << c))) continue;
    const size_t p = (size_t)c * A.stride + g;
    const float u = A.u[p], v = A.v[p];
    if (A.check_blow && (fl & kFlagBlow) && (double)fabsf(u) > A.thr) atomicMin(A.blow, A.iter);
    float un;
    if (A.newmark)
      un = (float)((double)u + A.dt * (double)v + A.hdt2 * (double)A.a0[p]);
    else
      un = (float)((double)u + (double)v * A.cd);
    if (A.apply_mask && (fl & kFlagAxis) && c >= A.ncomp_mask0 && c <= A.ncomp_mask1) {
      if (fl & kFlagAxisBdry) {
        for (int q = 0; q < A.naxb; q++)
          if (A.axb_g[q] == g) A.axb_val[3 * q + c] = un;
      }
      un = 0.0f;
    } else if (fl & kFlagAxisBdry) {
      for (int q = 0; q < A.naxb; q++)
        if (A.axb_g[q] == g) A.axb_val[3 * q + c] = un;
    }
    A.u[p] = un;
  }
}

// ---------------------------------------------------------------------------
// assembly + mass + update, per global point
// ---------------------------------------------------------------------------
struct AsmSolidArgs {
  int n, src, newmark, nslot, comps, mask0, mask1;
  const int *gstart, *gslot;
  const uint8_t *flags;
  const float *force, *minv;
  float *velo, *acc0;
  double hdt, cv;
  // source
  int nsrc;
  const int *src_g;
  const float *src_val;
  float stf;
};

__global__ __launch_bounds__(256) void k_asm_solid(AsmSolidArgs A) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= A.n) return;
  const uint8_t fl = A.flags[g];
  const int b = A.gstart[g], en = A.gstart[g + 1];
  const float mi = A.minv[g];
  int si = -1;
  if ((fl & kFlagSource) && A.stf != 0.0f) {
    for (int s = 0; s < A.nsrc; s++)
      if (A.src_g[s] == g) { si = s; break; }
  }
  const int order[3] = {0, 2, 1};
  for (int kk = 0; kk < 3; kk++) {
    const int c = order[kk];
    if (!(A.comps & (1 << c))) continue;
    float s = 0.0f;
    if ((fl & kFlagAxis) && c >= A.mask0 && c <= A.mask1) {
      s = 0.0f;  // masked copies are zero before the gather
    } else {
      const float *f = A.force + (size_t)c * A.nslot;
      for (int q = b; q < en; q++) s = s + f[A.gslot[q]];
    }
    if (si >= 0) s = s - A.src_val[3 * si + c] * A.stf;
    const size_t p = (size_t)c * A.n + g;
    if (A.newmark) {
      float a;
      if (c == 2 && A.src == AXH_DIPOLE)
        a = (float)(-2.0 * (double)mi * (double)s);
      else
        a = -mi * s;
      const float a0 = A.acc0[p];
      A.velo[p] = (float)((double)A.velo[p] + A.hdt * (double)(a0 + a));
      A.acc0[p] = a;
    } else {
      const float a = -mi * s;
      if (c == 2 && A.src == AXH_DIPOLE)
        A.velo[p] = (float)((double)A.velo[p] + 2.0 * (double)a * A.cv);
      else
        A.velo[p] = (float)((double)A.velo[p] + (double)a * A.cv);
    }
  }
}

struct AsmFluidArgs {
  int n, newmark, mask_axis, mask_surface;
  const int *gstart, *gslot;
  const uint8_t *flags;
  const float *force, *minv;
  float *dchi, *ddchi;
  double hdt, cv;
};

__global__ __launch_bounds__(256) void k_asm_fluid(AsmFluidArgs A) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= A.n) return;
  const uint8_t fl = A.flags[g];
  float s = 0.0f;
  const bool masked = (A.mask_axis && (fl & kFlagAxis)) ||
                      (A.mask_surface && (fl & kFlagSurface));
  if (!masked)
    for (int q = A.gstart[g]; q < A.gstart[g + 1]; q++) s = s + A.force[A.gslot[q]];
  const float mi = A.minv[g];
  if (A.newmark) {
    const float d1 = -mi * s;
    const float d0 = A.ddchi[g];
    A.dchi[g] = (float)((double)A.dchi[g] + A.hdt * (double)(d0 + d1));
    A.ddchi[g] = d1;
  } else {
    const float d = -s * mi;
    A.ddchi[g] = d;
    A.dchi[g] = (float)((double)A.dchi[g] + A.cv * (double)d);
  }
}

// ---------------------------------------------------------------------------
// solid-fluid coupling (SOLVER/time_evol_wave.F90:1532-1611), one lane per
// boundary GLL point
// ---------------------------------------------------------------------------
struct CoupleArgs {
  int n, dipole, nslot_s, nglob_s;
  const int *ss, *gs, *sf, *gf;
  const float *bm;     // [n][2]
  const float *disp;   // [3][nglob_s]
  const float *ddchi;  // [nglob_f]
  float *force_f;      // [nslot_f]
  float *force_s;      // [3][nslot_s]
  const int *axb;      // side-buffer slot per boundary point or -1
  const float *axb_val;
};

__global__ void k_copy2fluid(CoupleArgs A) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= A.n) return;
  const int g = A.gs[b], k = A.axb[b];
  float s1, s2, s3;
  if (k >= 0) {  // prediction before the axis mask
    s1 = A.axb_val[3 * k];
    s2 = A.axb_val[3 * k + 1];
    s3 = A.axb_val[3 * k + 2];
  } else {
    s1 = A.disp[g];
    s2 = A.disp[(size_t)A.nglob_s + g];
    s3 = A.disp[2 * (size_t)A.nglob_s + g];
  }
  const float bm1 = A.bm[2 * b], bm2 = A.bm[2 * b + 1];
  float f = A.force_f[A.sf[b]];
  if (A.dipole)
    f = f - bm1 * (s1 + s2) - bm2 * s3;
  else
    f = f - bm1 * s1 - bm2 * s3;
  A.force_f[A.sf[b]] = f;
}

__global__ void k_copy2solid(CoupleArgs A) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= A.n) return;
  const float fl = A.ddchi[A.gf[b]];
  const float bm1 = A.bm[2 * b], bm2 = A.bm[2 * b + 1];
  const size_t s = A.ss[b];
  A.force_s[s] = A.force_s[s] + bm1 * fl;
  if (A.dipole)
    A.force_s[(size_t)A.nslot_s + s] = A.force_s[(size_t)A.nslot_s + s] + bm1 * fl;
  A.force_s[2 * (size_t)A.nslot_s + s] = A.force_s[2 * (size_t)A.nslot_s + s] + bm2 * fl;
}

// receiver samples of compute_recfile_seis_bare (SOLVER/seismograms.f90:742)
__global__ void k_receivers(int n, int src, const int *rec_g, const float *disp,
                            int nglob, float *out) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const int g = rec_g[r];
  const float d1 = disp[g], d2 = disp[(size_t)nglob + g], d3 = disp[2 * (size_t)nglob + g];
  float *o = out + 3 * (size_t)r;
  if (src == AXH_MONOPOLE) { o[0] = d1; o[1] = d3; o[2] = 0.0f; }
  else if (src == AXH_DIPOLE) { o[0] = d1 + d2; o[1] = d1 - d2; o[2] = d3; }
  else { o[0] = d1; o[1] = d2; o[2] = d3; }
}

// ---------------------------------------------------------------------------
// host side: building the device layout
// ---------------------------------------------------------------------------
static inline size_t slot_of(int e, int k) {
  return (size_t)(e / kElemPerTile) * kTile + (e % kElemPerTile) * kNpp + k;
}

struct Numbering {
  std::vector<int> gid;      // per slot, -1 for padding
  std::vector<int> gstart;   // CSR over global points
  std::vector<int> gslot;
  std::vector<int> first;    // per global point: first (element, k) as e*25+k
  int nglob = 0;
};

// Renumber the reference's global points by first appearance in element
// order and list the element copies of every point in ascending element order
// (the order pdistsum accumulates them in, SOLVER/commun.F90:101-128).
static Numbering build_numbering(const int32_t *igloc, int nel, int nglob_ref,
                                 const char *what) {
  Numbering N;
  const int ntile = (nel + kElemPerTile - 1) / kElemPerTile;
  N.gid.assign((size_t)ntile * kTile, -1);
  std::vector<int> map(nglob_ref + 1, -1);
  std::vector<int> count;
  for (int e = 0; e < nel; e++) {
    for (int k = 0; k < kNpp; k++) {
      int r = igloc[(size_t)e * kNpp + k];
      if (r < 1 || r > nglob_ref)
        fail(AXH_ERR_ARG, std::string(what) + ": igloc out of range");
      int &m = map[r];
      if (m < 0) {
        m = N.nglob++;
        count.push_back(0);
        N.first.push_back(e * kNpp + k);
      }
      count[m]++;
      N.gid[slot_of(e, k)] = m;
    }
    // a point may appear only once per element
    for (int k = 0; k < kNpp; k++)
      for (int k2 = k + 1; k2 < kNpp; k2++)
        if (igloc[(size_t)e * kNpp + k] == igloc[(size_t)e * kNpp + k2])
          fail(AXH_ERR_ARG, std::string(what) + ": repeated point in an element");
  }
  N.gstart.assign(N.nglob + 1, 0);
  for (int g = 0; g < N.nglob; g++) N.gstart[g + 1] = N.gstart[g] + count[g];
  N.gslot.assign(N.gstart[N.nglob], 0);
  std::vector<int> fill(N.gstart.begin(), N.gstart.end() - 1);
  for (int e = 0; e < nel; e++)
    for (int k = 0; k < kNpp; k++) {
      int g = N.gid[slot_of(e, k)];
      N.gslot[fill[g]++] = (int)slot_of(e, k);
    }
  return N;
}

// per-global-point value of an element-local array; all copies must agree
static std::vector<float> per_point(const Numbering &N, const float *a, int nel,
                                    const char *what) {
  std::vector<float> v(N.nglob, 0.0f);
  std::vector<char> seen(N.nglob, 0);
  for (int e = 0; e < nel; e++)
    for (int k = 0; k < kNpp; k++) {
      int g = N.gid[slot_of(e, k)];
      float x = a[(size_t)e * kNpp + k];
      if (!seen[g]) { v[g] = x; seen[g] = 1; }
      else if (std::memcmp(&v[g], &x, 4) != 0)
        fail(AXH_ERR_UNSUPPORTED, std::string(what) + " differs between copies of a point");
    }
  return v;
}

static void pack_terms(const axh_ctx *ctx, const std::vector<std::string> &names,
                       int nel, std::vector<float> &out) {
  const int ntile = (nel + kElemPerTile - 1) / kElemPerTile;
  const int nc = (int)names.size();
  out.assign((size_t)ntile * nc * kTile, 0.0f);
  for (int a = 0; a < nc; a++) {
    const float *src = ctx->need(names[a]).as<float>();
    for (int e = 0; e < nel; e++) {
      const int tile = e / kElemPerTile, le = e % kElemPerTile;
      float *dst = &out[((size_t)tile * nc + a) * kTile + le * kNpp];
      std::memcpy(dst, src + (size_t)e * kNpp, sizeof(float) * kNpp);
    }
  }
}

// Host-side layout of one context, computed without touching the GPU: global
// numbering and gather lists, per-point values, flags, packed terms, coupling
// and receiver maps.  Every consistency requirement of the per-point storage
// is checked here.
struct HostLayout {
  int nel_s = 0, nel_f = 0, ntile_s = 0, ntile_f = 0, nglob_s = 0, nglob_f = 0;
  int nstages = 1, nsrcpts = 0, naxb = 0;
  double coefd[16] = {0}, coefv[16] = {0};
  std::vector<float> G, coef_s, m0_s, minv_s, src_val, coef_f, m0_f, minv_f, src_fl, cpl_bm,
      cg4, stf;
  std::vector<int> einfo_s, gid_s, gstart_s, gslot_s, src_g, einfo_f, gid_f, gstart_f,
      gslot_f, srcel_f, cpl_ss, cpl_gs, cpl_sf, cpl_gf, cpl_axb, axb_g, rec_g;
  std::vector<uint8_t> flags_s, flags_f;
  std::vector<double> aj_mu, aj_k;
  double expw[kMaxSls] = {0}, tst[kMaxSls] = {0}, tstm1[kMaxSls] = {0};
};

static void build_layout(const axh_ctx *ctx, HostLayout &H) {
  const axh_dims &d = ctx->dims;
  const int src = d.src_type;
  const bool multipole = src != AXH_MONOPOLE;
  H.nel_s = d.nel_solid;
  H.nel_f = d.nel_fluid;
  H.ntile_s = (H.nel_s + kElemPerTile - 1) / kElemPerTile;
  H.ntile_f = (H.nel_f + kElemPerTile - 1) / kElemPerTile;
  if (d.scheme != AXH_NEWMARK2) {
    double cf[16];
    H.nstages = symplectic_coefficients(d.scheme, d.deltat, H.coefd, H.coefv, cf);
    if (H.nstages <= 0) fail(AXH_ERR_UNSUPPORTED, "time scheme");
  }
  // spectral constants
  H.G.resize(105);
  std::memcpy(&H.G[0], ctx->need("G0").as<float>(), 5 * 4);
  std::memcpy(&H.G[5], ctx->need("G1").as<float>(), 25 * 4);
  std::memcpy(&H.G[30], ctx->need("G1T").as<float>(), 25 * 4);
  std::memcpy(&H.G[55], ctx->need("G2").as<float>(), 25 * 4);
  std::memcpy(&H.G[80], ctx->need("G2T").as<float>(), 25 * 4);

  // ---- solid ----
  const int32_t *axis_s = ctx->need("axis_solid").as<int32_t>();
  Numbering NS = build_numbering(ctx->need("igloc_solid").as<int32_t>(), H.nel_s,
                                 d.nglob_solid, "solid");
  H.nglob_s = NS.nglob;
  H.minv_s = per_point(NS, ctx->need("inv_mass_rho").as<float>(), H.nel_s, "inv_mass_rho");
  H.flags_s.assign(NS.nglob, 0);
  std::vector<char> nonaxis_copy(NS.nglob, 0);
  for (int e = 0; e < H.nel_s; e++)
    for (int k = 0; k < kNpp; k++) {
      int g = NS.gid[slot_of(e, k)];
      if (axis_s[e] && (k % 5) == 0) H.flags_s[g] |= kFlagAxis;
      else nonaxis_copy[g] = 1;
      if (k == 1 + 5 * 1) H.flags_s[g] |= kFlagBlow;
    }
  for (int g = 0; g < NS.nglob; g++)
    if ((H.flags_s[g] & kFlagAxis) && nonaxis_copy[g])
      fail(AXH_ERR_UNSUPPORTED, "axis point shared with a non-axial copy");

  // source terms (solid), added after assembly to every copy in a source
  // element (SOLVER/time_evol_wave.F90:1082-1097): copies must agree
  if (!d.fluid_src && d.nelsrc > 0) {
    const int32_t *iels = ctx->need("ielsrc").as<int32_t>();
    const float *st = ctx->need("source_term").as<float>();
    std::vector<int> srcidx(H.nel_s, -1);
    for (int kk = 0; kk < d.nelsrc; kk++) {
      if (iels[kk] < 1 || iels[kk] > H.nel_s) fail(AXH_ERR_ARG, "ielsrc out of range");
      srcidx[iels[kk] - 1] = kk;
    }
    std::vector<float> val(3 * (size_t)NS.nglob, 0.0f);
    std::vector<char> seen(NS.nglob, 0);
    for (int e = 0; e < H.nel_s; e++)
      for (int k = 0; k < kNpp; k++) {
        int g = NS.gid[slot_of(e, k)];
        float x[3];
        for (int c = 0; c < 3; c++)
          x[c] = srcidx[e] >= 0 ? st[k + kNpp * ((size_t)srcidx[e] + 8 * c)] : 0.0f;
        if (!seen[g]) {
          seen[g] = 1;
          for (int c = 0; c < 3; c++) val[3 * g + c] = x[c];
        } else {
          for (int c = 0; c < 3; c++)
            if (val[3 * g + c] != x[c])
              fail(AXH_ERR_UNSUPPORTED, "source term differs between copies of a point");
        }
      }
    for (int g = 0; g < NS.nglob; g++)
      if (val[3 * g] != 0.0f || val[3 * g + 1] != 0.0f || val[3 * g + 2] != 0.0f) {
        H.flags_s[g] |= kFlagSource;
        H.src_g.push_back(g);
        for (int c = 0; c < 3; c++) H.src_val.push_back(val[3 * g + c]);
      }
  }
  H.nsrcpts = (int)H.src_g.size();

  // packed stiffness terms
  std::vector<std::string> names;
  auto S = [&](int term) { names.push_back(kSolid3dNames[term]); };
  if (src == AXH_MONOPOLE) {
    for (int t : {AXH_M_1, AXH_M_2, AXH_M_3, AXH_M_4, AXH_M_W1, AXH_M11S, AXH_M21S,
                  AXH_M41S, AXH_M12S, AXH_M22S, AXH_M32S, AXH_M42S, AXH_M11Z, AXH_M21Z,
                  AXH_M41Z})
      S(t);
  } else if (src == AXH_DIPOLE) {
    for (int t : {AXH_M_1, AXH_M_2, AXH_M_3, AXH_M_4, AXH_M_5, AXH_M_6, AXH_M_7, AXH_M_8,
                  AXH_M_W1, AXH_M_W2, AXH_M_W3, AXH_M11S, AXH_M21S, AXH_M41S, AXH_M12S,
                  AXH_M22S, AXH_M42S, AXH_M13S, AXH_M32S, AXH_M33S, AXH_M43S, AXH_M11Z,
                  AXH_M21Z, AXH_M41Z})
      S(t);
  } else {
    for (int t : {AXH_M_1, AXH_M_2, AXH_M_3, AXH_M_4, AXH_M_5, AXH_M_6, AXH_M_7, AXH_M_8,
                  AXH_M_W1, AXH_M_W2, AXH_M_W3, AXH_M_W4, AXH_M_W5, AXH_M11S, AXH_M21S,
                  AXH_M41S, AXH_M12S, AXH_M22S, AXH_M32S, AXH_M42S, AXH_M11Z, AXH_M21Z,
                  AXH_M41Z, AXH_M1PHI, AXH_M2PHI, AXH_M4PHI})
      S(t);
  }
  pack_terms(ctx, names, H.nel_s, H.coef_s);

  // axial terms M0_w1..10
  H.einfo_s.assign(H.nel_s, -1);
  int nax = 0;
  for (int e = 0; e < H.nel_s; e++)
    if (axis_s[e]) {
      H.einfo_s[e] = nax++;
      for (int w = 1; w <= 10; w++) {
        const Array *a = ctx->get("M0_w" + std::to_string(w));
        for (int j = 0; j < 5; j++)
          H.m0_s.push_back(a ? a->as<float>()[(size_t)e * 5 + j] : 0.0f);
      }
    }
  H.gid_s = std::move(NS.gid);
  H.gstart_s = std::move(NS.gstart);
  H.gslot_s = std::move(NS.gslot);

  // ---- fluid ----
  if (H.nel_f > 0) {
    const int32_t *axis_f = ctx->need("axis_fluid").as<int32_t>();
    Numbering NF = build_numbering(ctx->need("igloc_fluid").as<int32_t>(), H.nel_f,
                                   d.nglob_fluid, "fluid");
    H.nglob_f = NF.nglob;
    H.minv_f = per_point(NF, ctx->need("inv_mass_fluid").as<float>(), H.nel_f,
                         "inv_mass_fluid");
    std::vector<float> surf = per_point(NF, ctx->need("fluid_free_surface_mask").as<float>(),
                                        H.nel_f, "fluid_free_surface_mask");
    H.flags_f.assign(NF.nglob, 0);
    std::vector<char> nonaxis(NF.nglob, 0);
    for (int e = 0; e < H.nel_f; e++)
      for (int k = 0; k < kNpp; k++) {
        int g = NF.gid[slot_of(e, k)];
        if (axis_f[e] && (k % 5) == 0) H.flags_f[g] |= kFlagAxis;
        else nonaxis[g] = 1;
      }
    for (int g = 0; g < NF.nglob; g++) {
      if ((H.flags_f[g] & kFlagAxis) && nonaxis[g] && multipole)
        fail(AXH_ERR_UNSUPPORTED, "fluid axis point shared with a non-axial copy");
      if (surf[g] == 0.0f) H.flags_f[g] |= kFlagSurface;
      else if (surf[g] != 1.0f) fail(AXH_ERR_UNSUPPORTED, "fluid surface mask must be 0 or 1");
    }
    std::vector<std::string> fn = {kFluid3dNames[0], kFluid3dNames[1], kFluid3dNames[2]};
    if (multipole) fn.push_back(kFluid3dNames[3]);
    pack_terms(ctx, fn, H.nel_f, H.coef_f);
    H.einfo_f.assign(H.nel_f, -1);
    int naxf = 0;
    const Array *m0a = ctx->get("M0_w_fl");
    for (int e = 0; e < H.nel_f; e++)
      if (axis_f[e]) {
        H.einfo_f[e] = naxf++;
        for (int j = 0; j < 5; j++) H.m0_f.push_back(m0a ? m0a->as<float>()[(size_t)e * 5 + j] : 0.0f);
      }
    if (multipole && naxf > 0 && !m0a) fail(AXH_ERR_STATE, "missing upload: M0_w_fl");
    H.srcel_f.assign(H.nel_f, -1);
    H.src_fl.assign(25 * 8, 0.0f);
    if (d.fluid_src && d.nelsrc > 0) {
      const int32_t *iels = ctx->need("ielsrc").as<int32_t>();
      for (int kk = 0; kk < d.nelsrc; kk++) {
        if (iels[kk] < 1 || iels[kk] > H.nel_f) fail(AXH_ERR_ARG, "ielsrc out of range");
        H.srcel_f[iels[kk] - 1] = kk;
      }
      std::memcpy(H.src_fl.data(), ctx->need("source_term").as<float>(), 25 * 8 * 4);
    }
    H.gid_f = std::move(NF.gid);
    H.gstart_f = std::move(NF.gstart);
    H.gslot_f = std::move(NF.gslot);
  }

  // ---- coupling ----
  if (d.nel_bdry > 0) {
    const int32_t *bse = ctx->need("bdry_solid_el").as<int32_t>();
    const int32_t *bfe = ctx->need("bdry_fluid_el").as<int32_t>();
    const int32_t *bjs = ctx->need("bdry_jpol_solid").as<int32_t>();
    const int32_t *bjf = ctx->need("bdry_jpol_fluid").as<int32_t>();
    const float *bm = ctx->need("bdry_matr").as<float>();
    std::vector<char> used_s((size_t)H.ntile_s * kTile, 0), used_f((size_t)H.ntile_f * kTile, 0);
    for (int b = 0; b < d.nel_bdry; b++) {
      int es = bse[b] - 1, ef = bfe[b] - 1;
      if (es < 0 || es >= H.nel_s || ef < 0 || ef >= H.nel_f || bjs[b] < 0 || bjs[b] > 4 ||
          bjf[b] < 0 || bjf[b] > 4)
        fail(AXH_ERR_ARG, "boundary element index out of range");
      for (int i = 0; i < 5; i++) {
        size_t s1 = slot_of(es, i + 5 * bjs[b]), s2 = slot_of(ef, i + 5 * bjf[b]);
        if (used_s[s1] || used_f[s2])
          fail(AXH_ERR_UNSUPPORTED, "a GLL point appears twice on the boundary");
        used_s[s1] = used_f[s2] = 1;
        const int g = H.gid_s[s1];
        // bdry_copy2fluid reads the predicted displacement before the axis
        // mask (SOLVER/time_evol_wave.F90:377,389): masked boundary points
        // keep their unmasked prediction in a side buffer
        int k = -1;
        if (H.flags_s[g] & kFlagAxis) {
          for (size_t q = 0; q < H.axb_g.size(); q++)
            if (H.axb_g[q] == g) k = (int)q;
          if (k < 0) {
            k = (int)H.axb_g.size();
            H.axb_g.push_back(g);
            H.flags_s[g] |= kFlagAxisBdry;
          }
        }
        H.cpl_ss.push_back((int)s1);
        H.cpl_gs.push_back(g);
        H.cpl_sf.push_back((int)s2);
        H.cpl_gf.push_back(H.gid_f[s2]);
        H.cpl_axb.push_back(k);
        H.cpl_bm.push_back(bm[i + 5 * (size_t)b]);
        H.cpl_bm.push_back(bm[i + 5 * ((size_t)b + d.nel_bdry)]);
      }
    }
    H.naxb = (int)H.axb_g.size();
    if (H.naxb > kMaxAxisBdry) fail(AXH_ERR_UNSUPPORTED, "too many axis points on the boundary");
  }

  // ---- attenuation ----
  if (d.anel) {
    const char *cgn[kCg4Terms] = {"Y_cg4", "V_s_eta_cg4", "V_s_xi_cg4", "V_z_eta_cg4",
                                  "V_z_xi_cg4", "DsDeta_over_J_sol_cg4",
                                  "DzDeta_over_J_sol_cg4", "DsDxi_over_J_sol_cg4",
                                  "DzDxi_over_J_sol_cg4", "inv_s_solid_cg4", "delta_mu_cg4",
                                  "delta_kappa_cg4"};
    H.cg4.resize((size_t)H.nel_s * kCg4Terms * 4);
    for (int a = 0; a < kCg4Terms; a++) {
      const float *p = ctx->need(cgn[a]).as<float>();
      for (int e = 0; e < H.nel_s; e++)
        for (int q = 0; q < 4; q++) H.cg4[((size_t)e * kCg4Terms + a) * 4 + q] = p[(size_t)e * 4 + q];
    }
    const double *ew = ctx->need("exp_w_j_deltat").as<double>();
    const double *tt = ctx->need("ts_fac_t").as<double>();
    const double *tm = ctx->need("ts_fac_tm1").as<double>();
    for (int s = 0; s < d.n_sls; s++) {
      H.expw[s] = ew[s];
      H.tst[s] = tt[s];
      H.tstm1[s] = tm[s];
    }
    const Array &amu = ctx->need("a_j_mu"), &ak = ctx->need("a_j_kappa");
    H.aj_mu.assign(amu.as<double>(), amu.as<double>() + amu.count());
    H.aj_k.assign(ak.as<double>(), ak.as<double>() + ak.count());
  }

  // ---- receivers ----
  if (d.num_rec > 0) {
    const int32_t *rf = ctx->need("recfile_el").as<int32_t>();
    H.rec_g.resize(d.num_rec);
    for (int r = 0; r < d.num_rec; r++) {
      int e = rf[r] - 1, i = rf[r + d.num_rec], j = rf[r + 2 * d.num_rec];
      if (e < 0 || e >= H.nel_s || i < 0 || i > 4 || j < 0 || j > 4)
        fail(AXH_ERR_ARG, "receiver index out of range");
      H.rec_g[r] = H.gid_s[slot_of(e, i + 5 * j)];
    }
  }
  const Array &stf = ctx->need("stf");
  H.stf.assign(stf.as<float>(), stf.as<float>() + stf.count());
}

void check_inputs(const axh_ctx *ctx) {
  HostLayout H;
  build_layout(ctx, H);
}

void global_numbering(const axh_ctx *ctx, int domain, int32_t *gid, int32_t *nglob) {
  const bool solid = domain == 0;
  const int nel = solid ? ctx->dims.nel_solid : ctx->dims.nel_fluid;
  Numbering N = build_numbering(ctx->need(solid ? "igloc_solid" : "igloc_fluid").as<int32_t>(),
                                nel, solid ? ctx->dims.nglob_solid : ctx->dims.nglob_fluid,
                                solid ? "solid" : "fluid");
  for (int e = 0; e < nel; e++)
    for (int k = 0; k < kNpp; k++) gid[(size_t)e * kNpp + k] = N.gid[slot_of(e, k)];
  *nglob = N.nglob;
}

void device_build(axh_ctx *ctx) {
  const axh_dims &d = ctx->dims;
  HostLayout H;
  build_layout(ctx, H);
  std::unique_ptr<Device> D(new Device);
  Device &V = *D;
  try {
    V.src = d.src_type;
    V.scheme = d.scheme;
    V.anel = d.anel;
    V.nsls = d.n_sls;
    V.nel_s = H.nel_s;
    V.nel_f = H.nel_f;
    V.ntile_s = H.ntile_s;
    V.ntile_f = H.ntile_f;
    V.nbdry = d.nel_bdry;
    V.nrec = d.num_rec;
    V.nstages = H.nstages;
    V.nglob_s = H.nglob_s;
    V.nglob_f = H.nglob_f;
    V.nsrcpts = H.nsrcpts;
    V.naxb = H.naxb;
    std::memcpy(V.coefd, H.coefd, sizeof V.coefd);
    std::memcpy(V.coefv, H.coefv, sizeof V.coefv);
    std::memcpy(V.expw, H.expw, sizeof V.expw);
    std::memcpy(V.tst, H.tst, sizeof V.tst);
    std::memcpy(V.tstm1, H.tstm1, sizeof V.tstm1);
    V.stf = H.stf;

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
      fail(AXH_ERR_HIP, "no HIP device available");
    HIPCHECK(hipStreamCreateWithFlags(&V.stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreate(&V.ev0));
    HIPCHECK(hipEventCreate(&V.ev1));

    V.G = V.upload(H.G);
    const size_t nslot_s = (size_t)V.ntile_s * kTile;
    V.coef_s = V.upload(H.coef_s);
    V.einfo_s = V.upload(H.einfo_s);
    V.m0_s = V.upload(H.m0_s);
    V.gid_s = V.upload(H.gid_s);
    V.gstart_s = V.upload(H.gstart_s);
    V.gslot_s = V.upload(H.gslot_s);
    V.minv_s = V.upload(H.minv_s);
    V.flags_s = V.upload(H.flags_s);
    V.src_g = V.upload(H.src_g);
    V.src_val = V.upload(H.src_val);
    V.disp = V.zeros<float>(3 * (size_t)V.nglob_s);
    V.velo = V.zeros<float>(3 * (size_t)V.nglob_s);
    V.acc0 = V.zeros<float>(3 * (size_t)V.nglob_s);
    V.force_s = V.zeros<float>(3 * nslot_s);
    if (V.nel_f > 0) {
      V.coef_f = V.upload(H.coef_f);
      V.einfo_f = V.upload(H.einfo_f);
      V.m0_f = V.upload(H.m0_f);
      V.gid_f = V.upload(H.gid_f);
      V.gstart_f = V.upload(H.gstart_f);
      V.gslot_f = V.upload(H.gslot_f);
      V.minv_f = V.upload(H.minv_f);
      V.flags_f = V.upload(H.flags_f);
      V.srcel_f = V.upload(H.srcel_f);
      V.src_fl = V.upload(H.src_fl);
      V.chi = V.zeros<float>(V.nglob_f);
      V.dchi = V.zeros<float>(V.nglob_f);
      V.ddchi = V.zeros<float>(V.nglob_f);
      V.force_f = V.zeros<float>((size_t)V.ntile_f * kTile);
    }
    if (V.nbdry > 0) {
      V.cpl_ss = V.upload(H.cpl_ss);
      V.cpl_gs = V.upload(H.cpl_gs);
      V.cpl_sf = V.upload(H.cpl_sf);
      V.cpl_gf = V.upload(H.cpl_gf);
      V.cpl_axb = V.upload(H.cpl_axb);
      V.cpl_bm = V.upload(H.cpl_bm);
      V.axb_g = V.upload(H.axb_g);
      V.axb_val = V.zeros<float>(3 * (size_t)std::max(V.naxb, 1));
    }
    if (V.anel) {
      V.cg4 = V.upload(H.cg4);
      V.aj_mu = V.upload(H.aj_mu);
      V.aj_k = V.upload(H.aj_k);
      V.memvar = V.zeros<float>((size_t)V.nel_s * 24 * V.nsls);
      V.rsum = V.zeros<float>((size_t)V.nel_s * 24);
      V.trm1 = V.zeros<float>((size_t)V.nel_s * 4);
      V.devm1 = V.zeros<float>((size_t)V.nel_s * 24);
    }
    if (V.nrec > 0) V.rec_g = V.upload(H.rec_g);
    V.blow = V.alloc<int>(1);
    HIPCHECK(hipDeviceSynchronize());
  } catch (...) {
    device_destroy(D.release());
    throw;
  }
  ctx->dev.reset(D.release());
}


void DeviceDeleter::operator()(Device *d) const { device_destroy(d); }

void device_destroy(Device *d) {
  if (!d) return;
  // teardown: errors are not actionable here
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  for (auto e : d->kev) (void)hipEventDestroy(e);
  if (d->ev0) (void)hipEventDestroy(d->ev0);
  if (d->ev1) (void)hipEventDestroy(d->ev1);
  for (void *p : d->allocs) (void)hipFree(p);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
}

size_t device_bytes(const Device *d) { return d ? d->bytes : 0; }

const char *kernel_class_name(int k) { return (k >= 0 && k < KC_N) ? kClassName[k] : nullptr; }

bool kernel_stats(const Device *d, int k, double *ms, int64_t *n) {
  if (k < 0 || k >= KC_N) return false;
  *ms = d ? d->kms[k] : 0.0;
  *n = d ? d->kcount[k] : 0;
  return true;
}

// ---------------------------------------------------------------------------
// the time loop
// ---------------------------------------------------------------------------
static inline int nblk(long n) { return (int)((n + 255) / 256); }

struct Launcher {
  Device &V;
  template <class F> void operator()(int kc, F &&launch) {
    if (V.profile) {
      hipEvent_t a, b;
      HIPCHECK(hipEventCreate(&a));
      HIPCHECK(hipEventCreate(&b));
      HIPCHECK(hipEventRecord(a, V.stream));
      launch();
      HIPCHECK(hipEventRecord(b, V.stream));
      V.kev.push_back(a);
      V.kev.push_back(b);
      V.kev_class.push_back(kc);
    } else {
      launch();
    }
  }
};

template <int SRC, int MODE, bool ANEL>
static void launch_solid(Device &V, const SolidArgs &A) {
  hipLaunchKernelGGL((k_solid<SRC, MODE, ANEL>), dim3(V.ntile_s), dim3(256), 0,
                     V.stream, A);
}

static void dispatch_solid(Device &V, const SolidArgs &A, int mode) {
  const bool an = V.anel != 0;
#define AXH_SOLID_CASE(S)                                                  \
  if (mode == 0) {                                                         \
    if (an) launch_solid<S, 0, true>(V, A); else launch_solid<S, 0, false>(V, A); \
  } else {                                                                 \
    if (an) launch_solid<S, 1, true>(V, A); else launch_solid<S, 1, false>(V, A); \
  }
  if (V.src == AXH_MONOPOLE) { AXH_SOLID_CASE(AXH_MONOPOLE) }
  else if (V.src == AXH_DIPOLE) { AXH_SOLID_CASE(AXH_DIPOLE) }
  else { AXH_SOLID_CASE(AXH_QUADPOLE) }
#undef AXH_SOLID_CASE
}

static void dispatch_memvar(Device &V, const SolidArgs &A) {
  if (V.src == AXH_MONOPOLE)
    hipLaunchKernelGGL(k_memvar<AXH_MONOPOLE>, dim3(V.ntile_s), dim3(256), 0, V.stream, A);
  else if (V.src == AXH_DIPOLE)
    hipLaunchKernelGGL(k_memvar<AXH_DIPOLE>, dim3(V.ntile_s), dim3(256), 0, V.stream, A);
  else
    hipLaunchKernelGGL(k_memvar<AXH_QUADPOLE>, dim3(V.ntile_s), dim3(256), 0, V.stream, A);
}

void device_run(axh_ctx *ctx, int32_t it0, int32_t it1, float *seis) {
  Device &V = *ctx->dev;
  const axh_dims &d = ctx->dims;
  const bool multipole = V.src != AXH_MONOPOLE;
  const bool newmark = V.scheme == AXH_NEWMARK2;
  const double dt = d.deltat, hdt = 0.5 * dt, hdt2 = 0.5 * (dt * dt);
  const size_t nslot_s = (size_t)V.ntile_s * kTile;
  V.profile = ctx->profile;
  Launcher L{V};

  const int nsamp = axh_seis_samples(ctx, it0, it1);
  if (nsamp > 0 && (size_t)nsamp * 3 * V.nrec > V.seis_cap) {
    V.seis_cap = (size_t)nsamp * 3 * V.nrec;
    V.seis_d = V.alloc<float>(V.seis_cap);
  }
  const int intmax = INT_MAX;
  HIPCHECK(hipMemcpyAsync(V.blow, &intmax, sizeof(int), hipMemcpyHostToDevice, V.stream));

  SolidArgs SA{};
  SA.nel = V.nel_s; SA.nsls = V.nsls; SA.G = V.G; SA.coef = V.coef_s; SA.m0 = V.m0_s;
  SA.einfo = V.einfo_s; SA.gid = V.gid_s; SA.disp = V.disp; SA.nglob = V.nglob_s;
  SA.force = V.force_s; SA.nslot = (int)nslot_s; SA.cg4 = V.cg4; SA.memvar = V.memvar;
  SA.rsum = V.rsum; SA.trm1 = V.trm1; SA.devm1 = V.devm1; SA.aj_mu = V.aj_mu;
  SA.aj_k = V.aj_k;
  for (int s = 0; s < kMaxSls; s++) {
    SA.expw[s] = V.expw[s]; SA.tst[s] = V.tst[s]; SA.tstm1[s] = V.tstm1[s];
  }
  FluidArgs FA{};
  FA.nel = V.nel_f; FA.multipole = multipole; FA.fluid_src = d.fluid_src; FA.G = V.G;
  FA.coef = V.coef_f; FA.m0 = V.m0_f; FA.einfo = V.einfo_f; FA.gid = V.gid_f;
  FA.srcel = V.srcel_f; FA.src_fl = V.src_fl; FA.chi = V.chi; FA.force = V.force_f;
  CoupleArgs CA{};
  CA.n = 5 * V.nbdry; CA.dipole = V.src == AXH_DIPOLE; CA.nslot_s = (int)nslot_s;
  CA.nglob_s = V.nglob_s; CA.ss = V.cpl_ss; CA.gs = V.cpl_gs; CA.sf = V.cpl_sf;
  CA.gf = V.cpl_gf; CA.bm = V.cpl_bm; CA.disp = V.disp; CA.ddchi = V.ddchi;
  CA.force_f = V.force_f; CA.force_s = V.force_s; CA.axb = V.cpl_axb;
  CA.axb_val = V.axb_val;
  // solid axis masks: monopole s; dipole +,z (2,3); quadrupole all
  const int mask0 = (V.src == AXH_DIPOLE) ? 1 : 0;
  const int mask1 = (V.src == AXH_MONOPOLE) ? 0 : 2;
  const int comps = multipole ? 7 : 5;
  PredArgs PS{};
  PS.n = V.nglob_s; PS.ncomp_mask0 = mask0; PS.ncomp_mask1 = mask1; PS.comps = comps;
  PS.dt = dt; PS.hdt2 = hdt2; PS.thr = 0.0; PS.u = V.disp; PS.v = V.velo;
  PS.a0 = V.acc0; PS.flags = V.flags_s; PS.stride = V.nglob_s; PS.blow = V.blow;
  PS.naxb = V.naxb; PS.axb_g = V.axb_g; PS.axb_val = V.axb_val;
  PredArgs PF = PS;
  PF.n = V.nglob_f; PF.ncomp_mask0 = 0; PF.ncomp_mask1 = 0; PF.comps = 1;
  PF.u = V.chi; PF.v = V.dchi; PF.a0 = V.ddchi; PF.flags = V.flags_f;
  PF.stride = V.nglob_f; PF.check_blow = 0; PF.apply_mask = multipole; PF.naxb = 0;
  AsmSolidArgs AS{};
  AS.n = V.nglob_s; AS.src = V.src; AS.newmark = newmark; AS.nslot = (int)nslot_s;
  AS.comps = comps; AS.mask0 = mask0; AS.mask1 = mask1; AS.gstart = V.gstart_s;
  AS.gslot = V.gslot_s; AS.flags = V.flags_s; AS.force = V.force_s; AS.minv = V.minv_s;
  AS.velo = V.velo; AS.acc0 = V.acc0; AS.hdt = hdt; AS.nsrc = V.nsrcpts;
  AS.src_g = V.src_g; AS.src_val = V.src_val;
  AsmFluidArgs AF{};
  AF.n = V.nglob_f; AF.newmark = newmark; AF.mask_axis = multipole;
  AF.mask_surface = newmark; AF.gstart = V.gstart_f; AF.gslot = V.gslot_f;
  AF.flags = V.flags_f; AF.force = V.force_f; AF.minv = V.minv_f; AF.dchi = V.dchi;
  AF.ddchi = V.ddchi; AF.hdt = hdt;
  // blow-up threshold of runtime_info: maxval(abs(disp(1,1,:,:))) >
  // 10*abs(magnitude), the real(4) maximum promoted to double
  PS.thr = 10.0 * std::fabs(d.magnitude);

  HIPCHECK(hipEventRecord(V.ev0, V.stream));
  int isamp = 0;
  for (int it = it0; it <= it1; it++) {
    for (int st = 0; st < V.nstages; st++) {
      const float stf = V.stf[(size_t)(it - 1) * V.nstages + st];
      // predictor / drift
      PS.newmark = PF.newmark = newmark;
      PS.cd = PF.cd = newmark ? 0.0 : V.coefd[st];
      PS.apply_mask = 1;
      PS.check_blow = (st == 0);
      PS.iter = it;
      L(KC_PRED, [&] {
        hipLaunchKernelGGL(k_predict, dim3(nblk(PS.n)), dim3(256), 0, V.stream, PS);
        if (V.nel_f > 0)
          hipLaunchKernelGGL(k_predict, dim3(nblk(PF.n)), dim3(256), 0, V.stream, PF);
      });
      if (V.nel_f > 0) {
        FA.stf = newmark ? stf : 0.0f;
        L(KC_FLUID, [&] {
          hipLaunchKernelGGL(k_fluid, dim3(V.ntile_f), dim3(256), 0, V.stream, FA);
        });
        if (V.nbdry > 0)
          L(KC_COUPLE_F, [&] {
            hipLaunchKernelGGL(k_copy2fluid, dim3(nblk(CA.n)), dim3(256), 0, V.stream, CA);
          });
      }
      L(KC_SOLID, [&] { dispatch_solid(V, SA, newmark ? 0 : 1); });
      if (V.nel_f > 0) {
        AF.cv = newmark ? 0.0 : V.coefv[st];
        L(KC_ASM_F, [&] {
          hipLaunchKernelGGL(k_asm_fluid, dim3(nblk(AF.n)), dim3(256), 0, V.stream, AF);
        });
        if (V.nbdry > 0)
          L(KC_COUPLE_S, [&] {
            hipLaunchKernelGGL(k_copy2solid, dim3(nblk(CA.n)), dim3(256), 0, V.stream, CA);
          });
      }
      AS.cv = newmark ? 0.0 : V.coefv[st];
      AS.stf = (newmark && d.fluid_src) ? 0.0f : stf;
      L(KC_ASM_S, [&] {
        hipLaunchKernelGGL(k_asm_solid, dim3(nblk(AS.n)), dim3(256), 0, V.stream, AS);
      });
    }
    if (!newmark) {
      // final drift with coefd(nstages+1), no mask (SOLVER/time_evol_wave.F90:720-725)
      PredArgs PD = PS, PDF = PF;
      PD.cd = PDF.cd = V.coefd[V.nstages];
      PD.newmark = PDF.newmark = 0;
      PD.apply_mask = PDF.apply_mask = 0;
      PD.check_blow = 0;
      L(KC_DRIFT, [&] {
        hipLaunchKernelGGL(k_predict, dim3(nblk(PD.n)), dim3(256), 0, V.stream, PD);
        if (V.nel_f > 0)
          hipLaunchKernelGGL(k_predict, dim3(nblk(PDF.n)), dim3(256), 0, V.stream, PDF);
      });
      if (V.anel) L(KC_MEMVAR, [&] { dispatch_memvar(V, SA); });
    }
    if (V.nrec > 0 && it % d.seis_it == 0) {
      float *out = V.seis_d + (size_t)isamp * 3 * V.nrec;
      L(KC_REC, [&] {
        hipLaunchKernelGGL(k_receivers, dim3(nblk(V.nrec)), dim3(256), 0, V.stream,
                           V.nrec, V.src, (const int *)V.rec_g, (const float *)V.disp,
                           V.nglob_s, out);
      });
      isamp++;
    }
  }
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipEventRecord(V.ev1, V.stream));
  int blow = INT_MAX;
  HIPCHECK(hipMemcpyAsync(&blow, V.blow, sizeof(int), hipMemcpyDeviceToHost, V.stream));
  if (seis && isamp > 0)
    HIPCHECK(hipMemcpyAsync(seis, V.seis_d, (size_t)isamp * 3 * V.nrec * sizeof(float),
                            hipMemcpyDeviceToHost, V.stream));
  HIPCHECK(hipStreamSynchronize(V.stream));
  float ms = 0.0f;
  HIPCHECK(hipEventElapsedTime(&ms, V.ev0, V.ev1));
  ctx->last_run_ms = ms;
  if (V.profile) {
    for (size_t q = 0; q < V.kev_class.size(); q++) {
      float kms = 0.0f;
      HIPCHECK(hipEventElapsedTime(&kms, V.kev[2 * q], V.kev[2 * q + 1]));
      V.kms[V.kev_class[q]] += kms;
      V.kcount[V.kev_class[q]] += 1;
    }
    for (auto e : V.kev) (void)hipEventDestroy(e);
    V.kev.clear();
    V.kev_class.clear();
  }
  ctx->iter_done = it1;
  if (blow != INT_MAX)
    fail(AXH_ERR_BLOWUP, "displacements blew up at time step " + std::to_string(blow));
}

// element-local fields from the per-point device arrays
void device_download(axh_ctx *ctx, float *disp, float *velo, float *chi,
                     float *dchi, float *ddchi) {
  Device &V = *ctx->dev;
  HIPCHECK(hipStreamSynchronize(V.stream));
  std::vector<int> gid_s((size_t)V.ntile_s * kTile);
  HIPCHECK(hipMemcpy(gid_s.data(), V.gid_s, gid_s.size() * 4, hipMemcpyDeviceToHost));
  auto scatter_s = [&](const float *dsrc, float *out) {
    std::vector<float> h(3 * (size_t)V.nglob_s);
    HIPCHECK(hipMemcpy(h.data(), dsrc, h.size() * 4, hipMemcpyDeviceToHost));
    for (int c = 0; c < 3; c++)
      for (int e = 0; e < V.nel_s; e++)
        for (int k = 0; k < kNpp; k++)
          out[((size_t)c * V.nel_s + e) * kNpp + k] =
              h[(size_t)c * V.nglob_s + gid_s[slot_of(e, k)]];
  };
  if (disp) scatter_s(V.disp, disp);
  if (velo) scatter_s(V.velo, velo);
  if (V.nel_f > 0 && (chi || dchi || ddchi)) {
    std::vector<int> gid_f((size_t)V.ntile_f * kTile);
    HIPCHECK(hipMemcpy(gid_f.data(), V.gid_f, gid_f.size() * 4, hipMemcpyDeviceToHost));
    auto scatter_f = [&](const float *dsrc, float *out) {
      std::vector<float> h(V.nglob_f);
      HIPCHECK(hipMemcpy(h.data(), dsrc, h.size() * 4, hipMemcpyDeviceToHost));
      for (int e = 0; e < V.nel_f; e++)
        for (int k = 0; k < kNpp; k++)
          out[(size_t)e * kNpp + k] = h[gid_f[slot_of(e, k)]];
    };
    if (chi) scatter_f(V.chi, chi);
    if (dchi) scatter_f(V.dchi, dchi);
    if (ddchi) scatter_f(V.ddchi, ddchi);
  }
}

void device_upload_fields(axh_ctx *ctx, const float *disp, const float *velo,
                          const float *chi, const float *dchi, const float *ddchi) {
  Device &V = *ctx->dev;
  HIPCHECK(hipStreamSynchronize(V.stream));
  std::vector<int> gid_s((size_t)V.ntile_s * kTile);
  HIPCHECK(hipMemcpy(gid_s.data(), V.gid_s, gid_s.size() * 4, hipMemcpyDeviceToHost));
  auto gather_s = [&](const float *in, float *ddst, const char *what) {
    std::vector<float> h(3 * (size_t)V.nglob_s, 0.0f);
    std::vector<char> seen(h.size(), 0);
    for (int c = 0; c < 3; c++)
      for (int e = 0; e < V.nel_s; e++)
        for (int k = 0; k < kNpp; k++) {
          size_t g = (size_t)c * V.nglob_s + gid_s[slot_of(e, k)];
          float x = in[((size_t)c * V.nel_s + e) * kNpp + k];
          if (!seen[g]) { h[g] = x; seen[g] = 1; }
          else if (h[g] != x) fail(AXH_ERR_ARG, std::string(what) + ": copies of a point differ");
        }
    HIPCHECK(hipMemcpy(ddst, h.data(), h.size() * 4, hipMemcpyHostToDevice));
  };
  if (disp) gather_s(disp, V.disp, "disp");
  if (velo) gather_s(velo, V.velo, "velo");
  if (V.nel_f > 0 && (chi || dchi || ddchi)) {
    std::vector<int> gid_f((size_t)V.ntile_f * kTile);
    HIPCHECK(hipMemcpy(gid_f.data(), V.gid_f, gid_f.size() * 4, hipMemcpyDeviceToHost));
    auto gather_f = [&](const float *in, float *ddst, const char *what) {
      std::vector<float> h(V.nglob_f, 0.0f);
      std::vector<char> seen(h.size(), 0);
      for (int e = 0; e < V.nel_f; e++)
        for (int k = 0; k < kNpp; k++) {
          size_t g = gid_f[slot_of(e, k)];
          float x = in[(size_t)e * kNpp + k];
          if (!seen[g]) { h[g] = x; seen[g] = 1; }
          else if (h[g] != x) fail(AXH_ERR_ARG, std::string(what) + ": copies of a point differ");
        }
      HIPCHECK(hipMemcpy(ddst, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    };
    if (chi) gather_f(chi, V.chi, "chi");
    if (dchi) gather_f(dchi, V.dchi, "dchi");
    if (ddchi) gather_f(ddchi, V.ddchi, "ddchi");
  }
}

}  // namespace axh
