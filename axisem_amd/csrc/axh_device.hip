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
//  * The mesh part (stiffness terms, gather lists, masses, coupling maps,
//    attenuation terms) lives once per mesh and is shared by every wavefield
//    stepped on it.  Wavefields are advanced in groups of B = 1..4: every
//    per-wavefield array of a group is interleaved with the wavefield index
//    innermost (element x of wavefield b at x*B + b), so one lane loads,
//    computes and stores the B wavefields of its point together, every mesh
//    term is read once for the group and every barrier is shared by it.
//    B = 1 is the plain layout.
//  * Arithmetic order follows the reference expression by expression; the
//    file is compiled with -ffp-contract=off (no FMA), and dp constants enter
//    in double precision where the Fortran promotes to dp.  A wavefield's
//    arithmetic does not depend on the other members of its group.  A second
//    build (AXH_FUSED_MULTIPLY_ADD, libaxihip_fma.so) contracts a*b+c into
//    fused multiply-adds: one rounding instead of two per pair, so its
//    results are not the reference's bit for bit but within its seismogram
//    misfit budget, for the kernels whose issue rate bounds them (the
//    stiffness of several wavefields per launch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "axh_internal.h"

#ifdef AXH_FUSED_MULTIPLY_ADD
#pragma clang fp contract(fast)
#else
#pragma clang fp contract(off)
#endif

namespace axh {

#define HIPCHECK(x)                                                          \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess)                                                    \
      fail(AXH_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));     \
  } while (0)

// number of packed element terms per source type (order: see pack_terms)
constexpr int kNcMono = 15, kNcDi = 24, kNcQuad = 26;
// the packed term that an isotropic model leaves +0 at every point: M_w2 of
// a dipole, M_w3 of a quadrupole (-1: none)
constexpr int zero_term_of(int src) {
  return src == AXH_DIPOLE ? 9 : (src == AXH_QUADPOLE ? 10 : -1);
}
// tiles of a group whose points include a source point of the group, passed
// to the stiffness kernel by value.  Every element holding a copy of a source
// point is a source element (the wavefield layout rejects a nonzero term
// missing from a copy), so a wavefield's source points lie in at most its
// nelsrc <= 8 elements' tiles and a group's in at most 8 per member.
constexpr int kMaxSrcTiles = 8 * AXH_MAX_GROUP;
constexpr int kNcFluidMono = 3, kNcFluidMulti = 4;
// coarse-grained terms per element, 4 cg points each.  Read by every stage:
// Y, V_s_eta, V_s_xi, V_z_eta, V_z_xi.  Read by the memory-variable update:
// DsDeta, DzDeta, DsDxi, DzDxi over J, inv_s, delta_mu, delta_kappa, pad.
constexpr int kCgStage = 5, kCgUpd = 8;
// wavefields per kernel instance (interleaved group width)
constexpr int kGroup = AXH_MAX_GROUP;

// flag bits per global point
constexpr uint8_t kFlagAxis = 1;      // on the symmetry axis (masked)
constexpr uint8_t kFlagSource = 2;    // source term of some wavefield on the mesh
constexpr uint8_t kFlagBlow = 4;      // GLL point (1,1) of an element
constexpr uint8_t kFlagSurface = 8;   // fluid free surface (masked)
constexpr uint8_t kFlagAxisBdry = 16; // axis point on the solid-fluid boundary
constexpr uint8_t kFlagCoupled = 32;  // owns a slot of the solid-fluid coupling
constexpr uint8_t kFlagDirect = 64;   // one element copy and nothing else to add:
                                      // kicked by the stiffness kernel itself
constexpr int kMaxAxisBdry = 16;

// kernel classes for timing
// (the fluid stiffness is part of the stiffness launch; the solid-fluid
// coupling and the symplectic drifts after the first are part of the
// assemble_update kernels)
enum KClass { KC_PRED = 0, KC_SOLID, KC_ASM_F, KC_ASM_S, KC_MEMVAR, KC_REC, KC_N };
static const char *const kClassName[KC_N] = {
    "predict", "stiffness", "fluid_assemble_update", "solid_assemble_update",
    "memvar_update", "receivers"};

// Device allocations owned by one object and freed with it.
struct DevAlloc {
  std::vector<void *> ptrs;
  size_t bytes = 0;
  DevAlloc() = default;
  DevAlloc(const DevAlloc &) = delete;
  DevAlloc &operator=(const DevAlloc &) = delete;
  template <class T> T *alloc(size_t n) {
    if (n == 0) n = 1;
    void *p = nullptr;
    HIPCHECK(hipMalloc(&p, n * sizeof(T)));
    ptrs.push_back(p);
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
  ~DevAlloc() {
    for (void *p : ptrs) (void)hipFree(p);  // teardown: not actionable
  }
};

// Mesh part: depends on the mesh, the model, the source type and the time
// scheme only.
struct MeshDev {
  DevAlloc mem;
  int src = 0, scheme = 0, anel = 0, nsls = 0, nstages = 1;
  int anel_full = 0;  // memory variables at all 25 points (anel is then 0)
  int nel_s = 0, nel_f = 0, ntile_s = 0, ntile_f = 0;
  int nglob_s = 0, nglob_f = 0, nbdry = 0, naxb = 0;
  double coefd[16] = {0}, coefv[16] = {0};
  double expw[kMaxSls] = {0}, tst[kMaxSls] = {0}, tstm1[kMaxSls] = {0};
  // host copies used by wavefield builds and field transfers
  std::vector<int> gid_s_h, gid_f_h;
  std::vector<int> first_order_s_h;  // solid points in first-appearance order
  int first_direct = 0;  // single-copy points are [first_direct, nglob_s)
  std::vector<uint8_t> flags_s_h;  // solid point flags without source bits (a
                                   // group's device copy adds its sources)
  float *G = nullptr;  // G0 | G1 | G1T | G2 | G2T
  // solid
  float *coef_s = nullptr, *m0_s = nullptr, *minv_s = nullptr;
  bool zero_term = false;  // the term zero_term_of(src) is +0 at every point
  int *einfo_s = nullptr, *gid_s = nullptr, *gstart_s = nullptr, *gslot_s = nullptr;
  // fluid
  float *coef_f = nullptr, *m0_f = nullptr, *minv_f = nullptr;
  float *gamma_s = nullptr, *gamma_f = nullptr;  // absorbing boundaries, or NULL
  int *einfo_f = nullptr, *gid_f = nullptr, *gstart_f = nullptr, *gslot_f = nullptr;
  uint8_t *flags_f = nullptr;
  // coupling
  int *cpl_ss = nullptr, *cpl_gs = nullptr, *cpl_sf = nullptr, *cpl_gf = nullptr;
  float *cpl_bm = nullptr;    // [nbdry*5][2]
  int *cpl_axb = nullptr;     // side-buffer slot of a masked axis point or -1
  int *axb_g = nullptr;       // global ids of the masked axis boundary points
  int *cpl_slot_s = nullptr;  // per solid slot: coupling entry or -1
  int *cpl_slot_f = nullptr;  // per fluid slot: coupling entry or -1
  // attenuation
  float *cg_stage = nullptr, *cg_upd = nullptr;
  float *af_terms = nullptr, *af_axis = nullptr;  // anel_full
  double *aj_mu = nullptr, *aj_k = nullptr;
};

// ---------------------------------------------------------------------------
// XCD-aware block order.  Workgroups are dispatched round-robin over the 8
// XCDs (block b to XCD b % 8), each with its own L2.  Neighbouring points
// gather from the same element slots, so the assembly kernels give each XCD
// one contiguous range of the n blocks: block b -> 8-way range b % 8,
// position b / 8 in it.  A bijection on [0, n).  (The element kernels keep
// the plain order: measured, the contiguous ranges did not help them.)
// ---------------------------------------------------------------------------
constexpr int kXcd = 8;
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int x = b % kXcd, r = b / kXcd;
  const int q = n / kXcd, rem = n % kXcd;
  return x * q + (x < rem ? x : rem) + r;
}

// ---------------------------------------------------------------------------
// B consecutive floats (one point, the wavefields of a group)
// ---------------------------------------------------------------------------
template <int B>
__device__ __forceinline__ void ldv(const float *p, float (&o)[B]) {
  if constexpr (B == 4) {
    const float4 v = *reinterpret_cast<const float4 *>(p);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  } else if constexpr (B == 2) {
    const float2 v = *reinterpret_cast<const float2 *>(p);
    o[0] = v.x; o[1] = v.y;
  } else {
#pragma unroll
    for (int b = 0; b < B; b++) o[b] = p[b];
  }
}
template <int B>
__device__ __forceinline__ void stv(float *p, const float (&o)[B]) {
  if constexpr (B == 4) {
    *reinterpret_cast<float4 *>(p) = make_float4(o[0], o[1], o[2], o[3]);
  } else if constexpr (B == 2) {
    *reinterpret_cast<float2 *>(p) = make_float2(o[0], o[1]);
  } else {
#pragma unroll
    for (int b = 0; b < B; b++) p[b] = o[b];
  }
}

// the same for data read or written once per pass (non-temporal)
template <int B>
__device__ __forceinline__ void ldv_nt(const float *p, float (&o)[B]) {
  if constexpr (B == 4) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(p));
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  } else if constexpr (B == 2) {
    typedef float v2f __attribute__((ext_vector_type(2)));
    const v2f v = __builtin_nontemporal_load(reinterpret_cast<const v2f *>(p));
    o[0] = v.x; o[1] = v.y;
  } else {
#pragma unroll
    for (int b = 0; b < B; b++) o[b] = __builtin_nontemporal_load(p + b);
  }
}
template <int B>
__device__ __forceinline__ void stv_nt(float *p, const float (&o)[B]) {
  if constexpr (B == 4) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    v4f v;
    v.x = o[0]; v.y = o[1]; v.z = o[2]; v.w = o[3];
    __builtin_nontemporal_store(v, reinterpret_cast<v4f *>(p));
  } else if constexpr (B == 2) {
    typedef float v2f __attribute__((ext_vector_type(2)));
    v2f v;
    v.x = o[0]; v.y = o[1];
    __builtin_nontemporal_store(v, reinterpret_cast<v2f *>(p));
  } else {
#pragma unroll
    for (int b = 0; b < B; b++) __builtin_nontemporal_store(o[b], p + b);
  }
}

// ---------------------------------------------------------------------------
// kernel argument blocks: mesh part, the group's interleaved fields, and the
// per-wavefield source / receiver data of its members
// ---------------------------------------------------------------------------
// The time step a launch belongs to, read on the device: wavefield b runs
// step base[b] + koff.  The host sets base once per run (k_step_set) and the
// launches of a step differ only in koff, so a captured graph of steps
// replays with base advanced on the device (k_step_advance).
struct StepRef {
  const int *base;  // [B]
  int koff;
};
__device__ __forceinline__ int step_of(const StepRef &r, int b) { return r.base[b] + r.koff; }

// Source time function of wavefield b at the step of r and stage st (table
// [niter][nstages], NULL: no source term in this kernel), SOLVER/time_evol_wave.F90:466,
// 1082-1097.
__device__ __forceinline__ float stf_at(const float *tab, const StepRef &r, int b, int nst,
                                        int st) {
  return tab ? tab[(size_t)(step_of(r, b) - 1) * nst + st] : 0.0f;
}

struct SolidArgs {
  int nel, nsls, nglob, nslot, ntile;
  const float *G, *coef, *m0;
  const int *einfo, *gid;
  const float *cg_stage;  // [nel][5][4]
  const float *cg_upd;    // [nel][8][4]
  const double *aj_mu, *aj_k;  // [nel][nsls]
  double expw[kMaxSls], tst[kMaxSls], tstm1[kMaxSls];
  const float *disp;      // [3][nglob][B]
  float *force;           // [3][nslot][B]
  float *memvar;          // [nel][nsls][6][4][B]
  float *rsum;            // [nel][6][4][B]
  float *trm1;            // [nel][4][B]
  float *devm1;           // [nel][6][4][B]
  // points of kFlagDirect: the stiffness kernel does their assembly (one
  // term), kick and, for symplectic stages, the drift that follows
  int direct;
  // the mesh's single-copy points are the range [first_direct, nglob); a
  // group's source points among them are not direct, so the tiles holding
  // source points (src_tile) read the group's flags instead
  int first_direct, nsrc_tile;
  int src_tile[kMaxSrcTiles];
  const uint8_t *flags;
  const float *minv;
  float *velo, *acc0;     // [3][nglob][B]
  float *disp_w;          // disp, written by the drift
  double hdt, cv, cd_next;
};

struct FluidArgs {
  int nel, multipole;
  const float *G, *coef, *m0;
  const int *einfo, *gid;
  const float *chi;       // [nglob_f][B]
  float *force;           // [nslot_f][B]
  const int *srcel[kGroup];
  const float *src_fl[kGroup];
  const float *stf[kGroup];  // NULL when the wavefield has no fluid source
  StepRef step;
  int nstages, stage;
};

// ---------------------------------------------------------------------------
// 5x5 products on LDS tiles (mxm_4 / vxm_4, SOLVER/unrolled_loops.f90:164-206)
// ---------------------------------------------------------------------------
// Spectral operators in LDS (105 floats, copied once per workgroup); lanes
// index them with their own (i, j), which keeps the register footprint small.
// Field tiles in LDS hold the B wavefields of a point together: entry k of a
// wavefield's 5x5 tile is at k*S (S = B).
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

// (A M)(i,j) = sum_k A(i,k) M(k,j)  for a 5x5 operator A
template <int S>
__device__ __forceinline__ float amul(const float *A, const float *M, int i, int j) {
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) s = s + A[i + 5 * k] * M[(k + 5 * j) * S];
  return s;
}
// (M B)(i,j) = sum_k M(i,k) B(k,j)  for a 5x5 operator B
template <int S>
__device__ __forceinline__ float mulb(const float *M, const float *Bop, int i, int j) {
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) s = s + M[(i + 5 * k) * S] * Bop[k + 5 * j];
  return s;
}
// sum_k g(k) M(k,j) for an operator vector g
template <int S>
__device__ __forceinline__ float vgm(const float *g, const float *M, int j) {
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) s = s + g[k] * M[(k + 5 * j) * S];
  return s;
}
// sum_k v(k) B(k,j) for a field vector v and an operator B
template <int S>
__device__ __forceinline__ float vmb(const float *v, const float *Bop, int j) {
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) s = s + v[k * S] * Bop[k + 5 * j];
  return s;
}

// Keeps v computed at this point: the compiler would otherwise sink the
// arithmetic producing it below the next barrier while the LDS reads feeding
// it stay above, keeping all their registers live across the barrier.
__device__ __forceinline__ void pin(float v) { asm volatile("" ::"v"(v)); }

// The same products for the B wavefields of a field tile at once: each step
// of the k sum reads the B values of one entry together (one LDS access) and
// the operator entry once; every wavefield's sum runs over k in order.
template <int B>
__device__ __forceinline__ void amulv(const float *A, const float *M, int i, int j, float (&o)[B]) {
#pragma unroll
  for (int b = 0; b < B; b++) o[b] = 0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const float a = A[i + 5 * k];
    float m[B];
    ldv<B>(M + (k + 5 * j) * B, m);
#pragma unroll
    for (int b = 0; b < B; b++) o[b] = o[b] + a * m[b];
  }
}
template <int B>
__device__ __forceinline__ void mulbv(const float *M, const float *Bop, int i, int j, float (&o)[B]) {
#pragma unroll
  for (int b = 0; b < B; b++) o[b] = 0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const float c = Bop[k + 5 * j];
    float m[B];
    ldv<B>(M + (i + 5 * k) * B, m);
#pragma unroll
    for (int b = 0; b < B; b++) o[b] = o[b] + m[b] * c;
  }
}
// vgm and vmb for the B wavefields of a field tile (entry k of a wavefield at
// k*B + b); every sum runs over k in order from zero
template <int B>
__device__ __forceinline__ void vgmv(const float *g, const float *M, int j, float (&o)[B]) {
#pragma unroll
  for (int b = 0; b < B; b++) o[b] = 0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const float a = g[k];
    float m[B];
    ldv<B>(M + (k + 5 * j) * B, m);
#pragma unroll
    for (int b = 0; b < B; b++) o[b] = o[b] + a * m[b];
  }
}
template <int B>
__device__ __forceinline__ void vmbv(const float *v, const float *Bop, int j, float (&o)[B]) {
#pragma unroll
  for (int b = 0; b < B; b++) o[b] = 0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const float c = Bop[k + 5 * j];
    float m[B];
    ldv<B>(v + k * B, m);
#pragma unroll
    for (int b = 0; b < B; b++) o[b] = o[b] + m[b] * c;
  }
}

// ---------------------------------------------------------------------------
// coarse-grained strain and memory variables (SOLVER/attenuation.f90:81-202,
// 472-537; SOLVER/pointwise_derivatives.f90:55-75,216-246)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int cg_r(int q) { return (q < 2) ? 1 : 3; }
__device__ __forceinline__ int cg_c(int q) { return (q & 1) ? 3 : 1; }

// gradient of f at cg point q: returns (dsdf, dzdf); cu = the element's
// update terms [8][4]: 0 DsDeta, 1 DzDeta, 2 DsDxi, 3 DzDxi (over J)
template <class F>
__device__ __forceinline__ void grad_cg4(const float *G, bool axial, const float *cu, int q,
                                         F f, float &gs, float &gz) {
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
  gs = cu[1 * 4 + q] * m1 + cu[3 * 4 + q] * m2;
  gz = cu[0 * 4 + q] * m1 + cu[2 * 4 + q] * m2;
}

// Voigt strain E1..E6 at cg point q from a wavefield's displacement tile in
// LDS (entry k at k*S).
template <int SRC, int S>
__device__ void strain_cg4(const float *G, bool axial, const float *cu, int q,
                           const float *U1, const float *U2, const float *U3,
                           float g[6]) {
  const float *invs = cu + 4 * 4;
  const int p = (cg_r(q) + 5 * cg_c(q)) * S;
  float b1s, b1z, b2s, b2z;
#pragma unroll
  for (int c = 0; c < 6; c++) g[c] = 0.0f;
  if (SRC == AXH_DIPOLE)
    grad_cg4(G, axial, cu, q, [&](int k) { return U1[k * S] + U2[k * S]; }, b1s, b1z);
  else
    grad_cg4(G, axial, cu, q, [&](int k) { return U1[k * S]; }, b1s, b1z);
  grad_cg4(G, axial, cu, q, [&](int k) { return U3[k * S]; }, b2s, b2z);
  g[0] = b1s;
  g[2] = b2z;
  g[4] = b1z + b2s;
  if (SRC == AXH_MONOPOLE) {
    g[1] = invs[q] * U1[p];
  } else if (SRC == AXH_DIPOLE) {
    g[1] = 2.0f * (invs[q] * U2[p]);
    grad_cg4(G, axial, cu, q, [&](int k) { return U1[k * S] - U2[k * S]; }, b1s, b1z);
    float f3 = invs[q] * U3[p];
    g[3] = -f3 - b1z;
    g[5] = -g[1] - b1s;
  } else {
    g[1] = invs[q] * (U1[p] - 2.0f * U2[p]);
    grad_cg4(G, axial, cu, q, [&](int k) { return U2[k * S]; }, b1s, b1z);
    float f3 = invs[q] * U3[p];
    float f6 = invs[q] * (U2[p] - 2.0f * U1[p]);
    g[3] = -(2.0f * f3) - b1z;
    g[5] = f6 - b1s;
  }
}

// source terms of the memory-variable update at cg point q
template <int SRC>
__device__ __forceinline__ void memvar_sources(const float *cu, int q, const float g[6],
                                               float &str, float sdev[6]) {
  const double third = 1.0 / 3.0;
  const float dm = cu[5 * 4 + q], dk = cu[6 * 4 + q];
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

// Position of deviator component c among the sources of an element: all 6
// (NSRC = 28), or for a monopole 0, 1, 2 and 4 (NSRC = 20; its components 3
// and 5 are not updated and take no slot: -1).
template <int NSRC>
__device__ __forceinline__ constexpr int src_slot(int c) {
  return NSRC == 28 ? c : (c < 3 ? c : (c == 4 ? 3 : -1));
}

// Strain sources of the memory-variable update at cg point q of element e
// for wavefield b of the group (lanes k < 4 of the element, one cg point
// each).  The sources of the previous update go to old, the new ones to src
// and to the t-1 store; src/old are [NSRC][B] per element: the trace at the
// 4 cg points, then the deviator components in src_slot order.
template <int SRC, int B, int NSRC = 28>
__device__ __forceinline__ void memvar_prepare(const SolidArgs &A, const float *sG, bool axial,
                                               int b, int e, int q, const float *U1,
                                               const float *U2, const float *U3, float *src,
                                               float *old) {
  static_assert(NSRC == 28 || SRC == AXH_MONOPOLE, "compact sources are the monopole's");
  const float *cu = A.cg_upd + (size_t)e * (kCgUpd * 4);
  float g[6], str, sdev[6];
  strain_cg4<SRC, B>(sG, axial, cu, q, U1 + b, U2 + b, U3 + b, g);
  memvar_sources<SRC>(cu, q, g, str, sdev);
  float *trm1 = A.trm1 + (size_t)e * 4 * B;
  float *devm1 = A.devm1 + (size_t)e * 24 * B;
  old[q * B + b] = trm1[q * B + b];
#pragma unroll
  for (int c = 0; c < 6; c++)
    if (src_slot<NSRC>(c) >= 0)
      old[(4 + 4 * src_slot<NSRC>(c) + q) * B + b] = devm1[(4 * c + q) * B + b];
  src[q * B + b] = str;
#pragma unroll
  for (int c = 0; c < 6; c++)
    if (src_slot<NSRC>(c) >= 0) src[(4 + 4 * src_slot<NSRC>(c) + q) * B + b] = sdev[c];
  trm1[q * B + b] = str;
#pragma unroll
  for (int c = 0; c < 6; c++) devm1[(4 * c + q) * B + b] = sdev[c];
}

// gather the (predicted, masked) displacement of the tile's points into LDS
// (monopole: components s and z)
template <int NCOMP, int B>
__device__ __forceinline__ void gather_disp(const SolidArgs &A, bool valid, int gp, int t,
                                            float *sU) {
  float u[B];
  if (valid) {
#pragma unroll
    for (int cc = 0; cc < NCOMP; cc++) {
      const int c = (NCOMP == 2 && cc == 1) ? 2 : cc;
      ldv<B>(A.disp + ((size_t)c * A.nglob + gp) * B, u);
      stv<B>(sU + (size_t)cc * 256 * B + t * B, u);
    }
  } else {
#pragma unroll
    for (int b = 0; b < B; b++) u[b] = 0.0f;
#pragma unroll
    for (int cc = 0; cc < NCOMP; cc++) stv<B>(sU + (size_t)cc * 256 * B + t * B, u);
  }
}

// ---------------------------------------------------------------------------
// fluid stiffness glob_fluid_stiffness_4 (SOLVER/stiffness_fluid.f90:139-216)
// + add_source_fl (SOLVER/time_evol_wave.F90:1062-1076) of one fluid tile; run
// by the workgroups of the solid stiffness launch past the solid tiles
// ---------------------------------------------------------------------------
// LDS (the stiffness kernel's own arrays): sC 256*B, sS0/sS1 256*B each,
// sG 105 floats, 16-byte aligned.
template <int B>
__device__ __forceinline__ void fluid_tile(const FluidArgs &A, int tile, float *sC, float *sS0,
                                           float *sS1, float *sG) {
  const int t = threadIdx.x;
  const int le = t / 25, k = t - 25 * le;
  const int i = k % 5, j = k / 5;
  const int e = tile * kElemPerTile + le;
  const bool valid = (t < 250) && (e < A.nel);
  const size_t slot = (size_t)tile * 256 + t;
  const int ax = valid ? A.einfo[e] : -1;
  const bool axial = ax >= 0;
  const int eb = le * 25 * B;
  stage_g(A.G, sG);
  const int nc = A.multipole ? kNcFluidMulti : kNcFluidMono;
  const float *coef = A.coef + (size_t)tile * nc * 256 + t;
  const float m1 = coef[0], m2 = coef[256], m4 = coef[512];
  const float mw = A.multipole ? coef[768] : 0.0f;
  const float w = (A.multipole && axial) ? A.m0[(size_t)ax * 5 + j] : 0.0f;
  {
    float c[B];
    if (valid) {
      ldv<B>(A.chi + (size_t)A.gid[slot] * B, c);
    } else {
#pragma unroll
      for (int b = 0; b < B; b++) c[b] = 0.0f;
    }
    stv<B>(sC + t * B, c);
  }
  __syncthreads();
  GRegs G;
  load_g(sG, axial, i, j, G);
  {
    float X1[B], X2[B], s0[B], s1[B];
    amulv<B>(G.gT, sC + eb, i, j, X1);
    mulbv<B>(sC + eb, G.g2, i, j, X2);
#pragma unroll
    for (int b = 0; b < B; b++) {
      s0[b] = m1 * X2[b] + m2 * X1[b];
      s1[b] = m1 * X1[b] + m4 * X2[b];
    }
    stv<B>(&sS0[t * B], s0);
    stv<B>(&sS1[t * B], s1);
  }
  __syncthreads();
  float out[B], Y1[B], Y2[B], V1[B];
  amulv<B>(G.g, &sS0[eb], i, j, Y1);
  mulbv<B>(&sS1[eb], G.g2T, i, j, Y2);
  if (A.multipole && axial) vgmv<B>(G.g0, sC + eb, j, V1);
#pragma unroll
  for (int b = 0; b < B; b++) {
    const float *C = sC + eb + b;
    float loc = Y1[b] + Y2[b];
    if (A.multipole) {
      loc = loc + mw * C[k * B];
      if (axial) loc = loc + G.g0i * (w * V1[b]);
    }
    if (valid) {
      const float stf = stf_at(A.stf[b], A.step, b, A.nstages, A.stage);
      if (stf != 0.0f) {
        const int ks = A.srcel[b][e];
        if (ks >= 0) loc = loc - A.src_fl[b][ks * 25 + k] * stf;
      }
    }
    out[b] = loc;
  }
  if (valid) stv<B>(A.force + slot * B, out);
}

// ---------------------------------------------------------------------------
// solid stiffness: glob_stiffness_{mono,di,quad}_4 (SOLVER/stiffness_mono.f90:
// 60-157, stiffness_di.f90:60-256, stiffness_quad.f90:238-412) followed by
// glob_anel_stiffness_*_cg4 and, for newmark2, time_step_memvars_cg4.
// MODE 0: newmark2 (memory variables read, used, updated)
// MODE 1: symplectic stage (sums of memory variables read)
// One workgroup = one tile of 10 elements, one lane = one GLL point and the
// B wavefields of the group; the lane's stiffness terms are read once.
// ---------------------------------------------------------------------------
// (occupancy hint: 5 workgroups per CU for a single wavefield, which fits the
// register budget without spilling; wider groups take what their registers
// allow)
// MS: the SLS capacity of the Newmark memory-variable update (5, the
// reference's default number of linear solids, or kMaxSls); only MODE 0
// with attenuation uses it.
// ZT: the term zero_term_of(SRC) is +0 at every point of the mesh and is not
// read (the kernel multiplies by the same +0: the same IEEE operations).
template <int SRC, int MODE, bool ANEL, int B, int MS = kMaxSls, bool ZT = false>
__global__ __launch_bounds__(256, (B == 1 ? 5 : (B == 2 ? 3 : 2)))
void k_solid(SolidArgs A, FluidArgs FA) {
  constexpr int NCOMP = (SRC == AXH_MONOPOLE) ? 2 : 3;
  constexpr int NC = (SRC == AXH_MONOPOLE) ? kNcMono
                     : (SRC == AXH_DIPOLE) ? kNcDi : kNcQuad;
  constexpr int NS = (SRC == AXH_MONOPOLE) ? 4 : 6;
  constexpr bool MV = ANEL && MODE == 0;  // memory variables updated here
  __shared__ __align__(16) float sU[NCOMP * 256 * B];
  __shared__ __align__(16) float sS[NS][256 * B];
  __shared__ __align__(16) float sV[kElemPerTile * 5 * B];
  // the SLS sums and their products with the cg4 terms live in the first two
  // product buffers, which are dead once the second products are summed (a
  // barrier separates the two uses)
  static_assert(kElemPerTile * 24 <= 256, "SLS sums must fit a product buffer");
  float *const sR = &sS[0][0];
  float *const sSa = &sS[1][0];
  // the strain sources of the Newmark memory-variable update go over the
  // product buffers from 2 on, dead once the second products are summed (the
  // SLS sums use buffers 0 and 1): a dipole group of 4 then fits 4 workgroups
  // per CU instead of 3.  The monopole has 4 buffers and keeps NSRC = 20
  // sources per element and wavefield (its components 3 and 5 are not
  // updated: src_slot), which fit the 2 left
  constexpr int NSRC = (SRC == AXH_MONOPOLE) ? 20 : 28;
  static_assert(2 * kElemPerTile * NSRC <= (NS - 2) * 256, "sources must fit the free buffers");
  float *const sSrc = &sS[2][0];
  float *const sOld = &sS[2][0] + kElemPerTile * NSRC * B;
  // ts_fac_t(jj) a_j and ts_fac_tm1(jj) a_j per element (mu, then kappa)
  __shared__ double sCf[MV ? kElemPerTile : 1][MV ? MS : 1][4];
  __shared__ __align__(16) float sG[105];

  if ((int)blockIdx.x >= A.ntile) {  // a fluid tile, in the same LDS
    fluid_tile<B>(FA, blockIdx.x - A.ntile, sU, sS[0], sS[1], sG);
    return;
  }
  const int t = threadIdx.x, tile = blockIdx.x;
  const int le = t / 25, k = t - 25 * le;
  const int i = k % 5, j = k / 5;
  const int e = tile * kElemPerTile + le;
  const bool valid = (t < 250) && (e < A.nel);
  const size_t slot = (size_t)tile * 256 + t;
  const int ax = valid ? A.einfo[e] : -1;
  const bool axial = ax >= 0;
  const int eb = le * 25 * B;  // the element's first entry in an LDS tile
  const int gp = valid ? A.gid[slot] : 0;

  stage_g(A.G, sG);
  // the lane's stiffness terms (the packed rows cover all 256 lanes), read
  // non-temporally: they are streamed once per launch (221 MB at 10 s) and
  // would otherwise evict the fields, forces and index lists that this and
  // the following kernels re-read from the L2 and the Infinity Cache
  float cf[NC];
  {
    const float *coef = A.coef + (size_t)tile * NC * 256 + t;
#pragma unroll
    for (int a = 0; a < NC; a++)
      cf[a] = (ZT && a == zero_term_of(SRC)) ? 0.0f : __builtin_nontemporal_load(coef + a * 256);
  }
  gather_disp<NCOMP, B>(A, valid, gp, t, sU);
  // a single-copy point is kicked here at the end: its velocity (and, for
  // newmark2, acceleration) and inverse mass are requested now
  // (one wavefield only: for a group the extra state traffic at the end of
  // this kernel costs more than the assembly kernel saves)
  bool direct = false;
  if (B == 1 && A.direct) {
    bool src_tile = false;
    for (int q = 0; q < A.nsrc_tile; q++) src_tile = src_tile || A.src_tile[q] == tile;
    direct = valid && (src_tile ? (A.flags[gp] & kFlagDirect) != 0 : gp >= A.first_direct);
  }
  float dv[NCOMP][B], da[MODE == 0 ? NCOMP : 1][B], dmi = 0.0f;
  if (direct) {
    dmi = A.minv[gp];
#pragma unroll
    for (int cc = 0; cc < NCOMP; cc++) {
      const int c = (NCOMP == 2 && cc == 1) ? 2 : cc;
      const size_t p = ((size_t)c * A.nglob + gp) * B;
      ldv<B>(A.velo + p, dv[cc]);
      if (MODE == 0) ldv<B>(A.acc0 + p, da[MODE == 0 ? cc : 0]);
    }
  }
  // newmark2 anelastic: lane k < 24 owns entries (k, b) of every SLS, read
  // now; lane 24 forms the element's SLS coefficient products
  float mvr[MV ? MS : 1][B];
  if (MV && valid && k < 24) {
    const float *mv = A.memvar + (size_t)e * 24 * A.nsls * B;
#pragma unroll
    for (int jj = 0; jj < (MV ? MS : 1); jj++)
      if (jj < A.nsls) ldv<B>(mv + (jj * 24 + k) * B, mvr[jj]);
  }
  if (MV && valid && k == 24) {
    for (int jj = 0; jj < A.nsls; jj++) {
      const double amu = A.aj_mu[(size_t)e * A.nsls + jj], ak = A.aj_k[(size_t)e * A.nsls + jj];
      sCf[MV ? le : 0][MV ? jj : 0][0] = A.tst[jj] * amu;
      sCf[MV ? le : 0][MV ? jj : 0][1] = A.tstm1[jj] * amu;
      sCf[MV ? le : 0][MV ? jj : 0][2] = A.tst[jj] * ak;
      sCf[MV ? le : 0][MV ? jj : 0][3] = A.tstm1[jj] * ak;
    }
  }
  __syncthreads();

  GRegs G;
  load_g(sG, axial, i, j, G);
  float acc[3][B];
  const float *U0 = sU + eb, *U1 = sU + 256 * B + eb, *U2 = sU + 2 * 256 * B + eb;

  if (SRC == AXH_MONOPOLE) {
    const float m1 = cf[0], m2 = cf[1], m3 = cf[2], m4 = cf[3], mw1 = cf[4],
                m11s = cf[5], m21s = cf[6], m41s = cf[7], m12s = cf[8], m22s = cf[9],
                m32s = cf[10], m42s = cf[11], m11z = cf[12], m21z = cf[13], m41z = cf[14];
    float ls[B], lz[B], V1[B], V2[B], V3[B];
    float vX1[B], vX2[B], vX3[B], vX4[B], vus[B];
    amulv<B>(G.gT, U0, i, j, vX1);
    amulv<B>(G.gT, U1, i, j, vX2);
    mulbv<B>(U0, G.g2, i, j, vX3);
    mulbv<B>(U1, G.g2, i, j, vX4);
    ldv<B>(U0 + k * B, vus);
#pragma unroll
    for (int b = 0; b < B; b++) {
      const float us = vus[b];
      const float X1 = vX1[b], X2 = vX2[b], X3 = vX3[b], X4 = vX4[b];
      ls[b] = m4 * X4 + m2 * X3 + m1 * X1 + m3 * X2 + us * mw1;
      sS[0][t * B + b] = m11s * X3 + m21s * X1 + m12s * X4 + m22s * X2 + m1 * us;
      sS[1][t * B + b] = m11s * X1 + m41s * X3 + m32s * X2 + m42s * X4 + m2 * us;
      sS[2][t * B + b] = m11z * X4 + m21z * X2 + m32s * X3 + m22s * X1 + m3 * us;
      sS[3][t * B + b] = m11z * X2 + m41z * X4 + m12s * X1 + m42s * X3 + m4 * us;
    }
    __syncthreads();
    // axial terms with M0_w1..3
    float w1 = 0, w2 = 0, w3 = 0;
    if (axial) {
      const float *m0 = A.m0 + (size_t)ax * 50;
      w1 = m0[0 * 5 + j]; w2 = m0[1 * 5 + j]; w3 = m0[2 * 5 + j];
    }
    // the axial terms first: their registers are free again before the
    // products' LDS reads are issued
    if (axial) {
      vgmv<B>(G.g0, U0, j, V1);          // vxm(G0, us)
      mulbv<B>(U1, G.g2, 0, j, V2);      // vxm(uz(0,:), G2)
      vgmv<B>(G.g0, U1, j, V3);          // vxm(G0, uz)
      if (i == 0) {
        float x[B];
#pragma unroll
        for (int b = 0; b < B; b++) x[b] = w3 * V1[b];
        stv<B>(&sV[(le * 5 + j) * B], x);
      }
    }
    float vY1[B], vY2[B], vY3[B], vY4[B];
    mulbv<B>(&sS[1][eb], G.g2T, i, j, vY2);
    mulbv<B>(&sS[3][eb], G.g2T, i, j, vY4);
    amulv<B>(G.g, &sS[0][eb], i, j, vY1);
    amulv<B>(G.g, &sS[2][eb], i, j, vY3);
#pragma unroll
    for (int b = 0; b < B; b++) {
      ls[b] = ls[b] + vY1[b] + vY2[b];
      lz[b] = vY3[b] + vY4[b];
      pin(ls[b]);
      pin(lz[b]);
    }
    __syncthreads();
    float V1b[B];
    if (axial) vmbv<B>(&sV[le * 5 * B], G.g2T, j, V1b);  // vxm(w3*V1, G2T)
#pragma unroll
    for (int b = 0; b < B; b++) {
      if (axial) {
        float V4 = w1 * V1[b] + w3 * V2[b];
        V4 = V4 + w2 * V3[b];
        float X2a = G.g0i * (w2 * V1[b]);
        if (i == 0) X2a = X2a + V1b[b];
        ls[b] = ls[b] + G.g0i * V4;
        lz[b] = X2a + lz[b];
      }
      acc[0][b] = ls[b];
      acc[1][b] = 0.0f;
      acc[2][b] = lz[b];
    }
  } else if (SRC == AXH_DIPOLE) {
    const float m1 = cf[0], m2 = cf[1], m3 = cf[2], m4 = cf[3], m5 = cf[4], m6 = cf[5],
                m7 = cf[6], m8 = cf[7], mw1 = cf[8], mw2 = cf[9], mw3 = cf[10],
                m11s = cf[11], m21s = cf[12], m41s = cf[13], m12s = cf[14],
                m22s = cf[15], m42s = cf[16], m13s = cf[17], m23s = cf[18],
                m33s = cf[19], m43s = cf[20], m11z = cf[21], m21z = cf[22],
                m41z = cf[23];
    float ls2[B], ls3[B];
    float vX1[B], vX2[B], vX3[B], vX4[B], vX5[B], vX6[B], vu2[B], vu3[B];
    mulbv<B>(U0, G.g2, i, j, vX4);
    mulbv<B>(U1, G.g2, i, j, vX5);
    mulbv<B>(U2, G.g2, i, j, vX6);
    amulv<B>(G.gT, U0, i, j, vX1);
    amulv<B>(G.gT, U1, i, j, vX2);
    amulv<B>(G.gT, U2, i, j, vX3);
    ldv<B>(U1 + k * B, vu2);
    ldv<B>(U2 + k * B, vu3);
#pragma unroll
    for (int b = 0; b < B; b++) {
      const float u2 = vu2[b], u3 = vu3[b];
      const float X1 = vX1[b], X2 = vX2[b], X3 = vX3[b], X4 = vX4[b], X5 = vX5[b], X6 = vX6[b];
      const float X7 = X1 + X2, X8 = X4 + X5;
      ls2[b] = m8 * X6 + m7 * X3 + m1 * X1 + m5 * X2 + m2 * X4 + m6 * X5 + mw1 * u2 +
               mw2 * u3;
      ls3[b] = m4 * X4 - m4 * X5 + m3 * X1 - m3 * X2 + mw2 * u2 + mw3 * u3;
      float c1 = m13s * X6, c2 = m23s * X3, c3 = m3 * u3;
      sS[0][t * B + b] = c1 + c2 + c3 + m11s * X4 + m21s * X1 + m12s * X5 + m22s * X2 + m1 * u2;
      sS[1][t * B + b] = c1 + c2 - c3 + m11s * X5 + m21s * X2 + m12s * X4 + m22s * X1 + m5 * u2;
      c1 = m33s * X3;
      c2 = m43s * X6;
      c3 = m4 * u3;
      sS[2][t * B + b] = c1 + c2 + c3 + m11s * X1 + m41s * X4 + m12s * X2 + m42s * X5 + m2 * u2;
      sS[3][t * B + b] = c1 + c2 - c3 + m11s * X2 + m41s * X5 + m12s * X1 + m42s * X4 + m6 * u2;
      sS[4][t * B + b] = m33s * X8 + m23s * X7 + m11z * X6 + m21z * X3 + m7 * u2;
      sS[5][t * B + b] = m13s * X7 + m43s * X8 + m11z * X3 + m41z * X6 + m8 * u2;
    }
    __syncthreads();
    float w1 = 0, w2 = 0, w3 = 0, w4 = 0, w6 = 0, w7 = 0, w8 = 0, w9 = 0, w10 = 0;
    if (axial) {
      const float *m0 = A.m0 + (size_t)ax * 50;
      w1 = m0[0 * 5 + j]; w2 = m0[1 * 5 + j]; w3 = m0[2 * 5 + j]; w4 = m0[3 * 5 + j];
      w6 = m0[5 * 5 + j]; w7 = m0[6 * 5 + j]; w8 = m0[7 * 5 + j]; w9 = m0[8 * 5 + j];
      w10 = m0[9 * 5 + j];
    }
    float va[B], vb[B], vc[B];
    // the axial terms first: their registers are free again before the
    // products' LDS reads are issued
    if (axial) {
      float V1[B], V2[B], V3[B], V4[B], x[B];
      vgmv<B>(G.g0, U0, j, V1);
      vgmv<B>(G.g0, U1, j, V2);
      vgmv<B>(G.g0, U2, j, V3);
      mulbv<B>(U0, G.g2, 0, j, V4);   // vxm(u1(0,:), G2)
#pragma unroll
      for (int b = 0; b < B; b++) {
        va[b] = w1 * V2[b] + w3 * V3[b];
        vb[b] = w1 * V1[b] + (w2 + w6) * V4[b] + w9 * V2[b] + w10 * V3[b];
        vc[b] = w3 * V1[b] + (w4 + w8) * V4[b] + w7 * V3[b] + w10 * V2[b];
        x[b] = (w2 + w6) * V2[b] + (w4 + w8) * V3[b];
      }
      if (i == 0) stv<B>(&sV[(le * 5 + j) * B], x);
    }
    {
      float Y[6][B];
      amulv<B>(G.g, &sS[0][eb], i, j, Y[0]);
      mulbv<B>(&sS[2][eb], G.g2T, i, j, Y[1]);
      amulv<B>(G.g, &sS[1][eb], i, j, Y[2]);
      mulbv<B>(&sS[3][eb], G.g2T, i, j, Y[3]);
      amulv<B>(G.g, &sS[4][eb], i, j, Y[4]);
      mulbv<B>(&sS[5][eb], G.g2T, i, j, Y[5]);
#pragma unroll
      for (int b = 0; b < B; b++) {
        acc[0][b] = Y[0][b] + Y[1][b];
        acc[1][b] = Y[2][b] + Y[3][b] + ls2[b];
        acc[2][b] = Y[4][b] + Y[5][b] + ls3[b];
      }
    }
    __syncthreads();
    if (axial) {
      float V1b[B];
      vmbv<B>(&sV[le * 5 * B], G.g2T, j, V1b);  // vxm(V4, G2T)
#pragma unroll
      for (int b = 0; b < B; b++) {
        float S1pa = G.g0i * va[b];
        const float S1ma = G.g0i * vb[b], S1za = G.g0i * vc[b];
        if (i == 0) S1pa = S1pa + V1b[b];
        acc[0][b] = acc[0][b] + S1pa;
        acc[1][b] = acc[1][b] + S1ma;
        acc[2][b] = acc[2][b] + S1za;
      }
    }
  } else {
    const float m1 = cf[0], m2 = cf[1], m3 = cf[2], m4 = cf[3], m5 = cf[4], m6 = cf[5],
                m7 = cf[6], m8 = cf[7], mw1 = cf[8], mw2 = cf[9], mw3 = cf[10],
                mw4 = cf[11], mw5 = cf[12], m11s = cf[13], m21s = cf[14],
                m41s = cf[15], m12s = cf[16], m22s = cf[17], m32s = cf[18],
                m42s = cf[19], m11z = cf[20], m21z = cf[21], m41z = cf[22],
                m1phi = cf[23], m2phi = cf[24], m4phi = cf[25];
    float vX1[B], vX2[B], vX3[B], vX4[B], vX5[B], vX6[B], vus[B], vup[B], vuz[B];
    amulv<B>(G.gT, U0, i, j, vX1);
    amulv<B>(G.gT, U1, i, j, vX2);
    amulv<B>(G.gT, U2, i, j, vX3);
    mulbv<B>(U0, G.g2, i, j, vX4);
    mulbv<B>(U1, G.g2, i, j, vX5);
    mulbv<B>(U2, G.g2, i, j, vX6);
    ldv<B>(U0 + k * B, vus);
    ldv<B>(U1 + k * B, vup);
    ldv<B>(U2 + k * B, vuz);
#pragma unroll
    for (int b = 0; b < B; b++) {
      const float us = vus[b], up = vup[b], uz = vuz[b];
      const float X1 = vX1[b], X2 = vX2[b], X3 = vX3[b], X4 = vX4[b], X5 = vX5[b], X6 = vX6[b];
      const float c1 = m2 * X4, c2 = m1 * X1, c3 = m6 * X5, c4 = m5 * X2,
                  c5 = m4 * X6, c6 = m3 * X3;
      acc[0][b] = c1 + c2 + 2.0f * (c3 + c4) + c5 + c6 + mw1 * us + mw2 * up +
                  2.0f * mw3 * uz;
      acc[1][b] = -2.0f * (c1 + c2 + c5 + c6) - (c3 + c4) + mw2 * us + mw4 * up -
                  mw3 * uz;
      acc[2][b] = 2.0f * (m8 * X5 + m7 * X2) + mw3 * (2.0f * us - up) + mw5 * uz;
      const float usp = us - 2.0f * up, tus = 2.0f * us - up;
      sS[0][t * B + b] = m11s * X4 + m21s * X1 + m12s * X6 + m22s * X3 + m1 * usp;   // S1s
      sS[1][t * B + b] = m11s * X1 + m41s * X4 + m32s * X3 + m42s * X6 + m2 * usp;   // S2s
      sS[2][t * B + b] = m11z * X6 + m21z * X3 + m32s * X4 + m22s * X1 + m3 * usp;   // S1z
      sS[3][t * B + b] = m11z * X3 + m41z * X6 + m12s * X1 + m42s * X4 + m4 * usp;   // S2z
      sS[4][t * B + b] = m1phi * X5 + m2phi * X2 + m5 * tus + 2.0f * m7 * uz;        // S1phi
      sS[5][t * B + b] = m1phi * X2 + m4phi * X5 + m6 * tus + 2.0f * m8 * uz;        // S2phi
    }
    __syncthreads();
    {
      float Y[6][B];
      mulbv<B>(&sS[1][eb], G.g2T, i, j, Y[1]);
      mulbv<B>(&sS[5][eb], G.g2T, i, j, Y[3]);
      mulbv<B>(&sS[3][eb], G.g2T, i, j, Y[5]);
      amulv<B>(G.g, &sS[0][eb], i, j, Y[0]);
      amulv<B>(G.g, &sS[4][eb], i, j, Y[2]);
      amulv<B>(G.g, &sS[2][eb], i, j, Y[4]);
#pragma unroll
      for (int b = 0; b < B; b++) {
        acc[0][b] = acc[0][b] + Y[0][b] + Y[1][b];
        acc[1][b] = acc[1][b] + Y[2][b] + Y[3][b];
        acc[2][b] = acc[2][b] + Y[4][b] + Y[5][b];
      }
    }
    __syncthreads();  // uniform barrier count with the other source types
    if (axial) {
      const float *m0 = A.m0 + (size_t)ax * 50;
      const float w1 = m0[0 * 5 + j], w2 = m0[1 * 5 + j], w3 = m0[2 * 5 + j],
                  w4 = m0[3 * 5 + j], w5 = m0[4 * 5 + j], w6 = m0[5 * 5 + j];
      float vV1[B], vV2[B], vV3[B];
      vgmv<B>(G.g0, U0, j, vV1);
      vgmv<B>(G.g0, U1, j, vV2);
      vgmv<B>(G.g0, U2, j, vV3);
#pragma unroll
      for (int b = 0; b < B; b++) {
        const float V1 = vV1[b], V2 = vV2[b], V3 = vV3[b];
        const float a = w1 * V1 + w2 * V2 + w3 * V3;
        const float bq = w2 * V1 + w4 * V2 + w5 * V3;
        const float c = w3 * V1 + w5 * V2 + w6 * V3;
        acc[0][b] = acc[0][b] + G.g0i * a;
        acc[1][b] = acc[1][b] + G.g0i * bq;
        acc[2][b] = acc[2][b] + G.g0i * c;
      }
    }
  }

  if (ANEL) {
    // sums of the memory variables over the SLS, (c, q) per lane k < 24
    float *R = &sR[le * 24 * B];
    if (valid && k < 24) {
      if (MODE == 0) {
#pragma unroll
        for (int b = 0; b < B; b++) {
          float s = 0.0f;
#pragma unroll
          for (int jj = 0; jj < (MV ? MS : 1); jj++)
            if (jj < A.nsls) s = s + mvr[jj][b];
          R[k * B + b] = s;
        }
      } else {
        float r[B];
        ldv<B>(A.rsum + ((size_t)e * 24 + k) * B, r);
        stv<B>(R + k * B, r);
      }
    }
    const float *cs = A.cg_stage + (size_t)(valid ? e : 0) * (kCgStage * 4);
    __syncthreads();
    float *Sa = &sSa[le * 24 * B];
    if (valid && k < 4 * B) {
      // lane k: cg point q = k % 4 of wavefield b = k / 4
      const int q = k & 3, b = k >> 2;
      const float vse = cs[1 * 4 + q], vsx = cs[2 * 4 + q], vze = cs[3 * 4 + q],
                  vzx = cs[4 * 4 + q];
      {
        const float r1 = R[(0 * 4 + q) * B + b], r3 = R[(2 * 4 + q) * B + b],
                    r4 = R[(3 * 4 + q) * B + b], r5 = R[(4 * 4 + q) * B + b],
                    r6 = R[(5 * 4 + q) * B + b];
        if (SRC == AXH_MONOPOLE) {
          Sa[(0 * 4 + q) * B + b] = vze * r1 + vse * r5;
          Sa[(1 * 4 + q) * B + b] = vzx * r1 + vsx * r5;
          Sa[(4 * 4 + q) * B + b] = vze * r5 + vse * r3;
          Sa[(5 * 4 + q) * B + b] = vzx * r5 + vsx * r3;
        } else if (SRC == AXH_DIPOLE) {
          Sa[(0 * 4 + q) * B + b] = vze * (r1 - r6) + vse * (r5 - r4);
          Sa[(1 * 4 + q) * B + b] = vzx * (r1 - r6) + vsx * (r5 - r4);
          Sa[(2 * 4 + q) * B + b] = vze * (r1 + r6) + vse * (r5 + r4);
          Sa[(3 * 4 + q) * B + b] = vzx * (r1 + r6) + vsx * (r5 + r4);
          Sa[(4 * 4 + q) * B + b] = vze * r5 + vse * r3;
          Sa[(5 * 4 + q) * B + b] = vzx * r5 + vsx * r3;
        } else {
          Sa[(0 * 4 + q) * B + b] = vze * r1 + vse * r5;
          Sa[(1 * 4 + q) * B + b] = vzx * r1 + vsx * r5;
          Sa[(2 * 4 + q) * B + b] = vze * r6 + vse * r4;
          Sa[(3 * 4 + q) * B + b] = vzx * r6 + vsx * r4;
          Sa[(4 * 4 + q) * B + b] = vze * r5 + vse * r3;
          Sa[(5 * 4 + q) * B + b] = vzx * r5 + vsx * r3;
        }
      }
    }
    __syncthreads();
    if (valid) {
      // mxm_cg4_sparse_b(G, S1) has columns 1,3; mxm_cg4_sparse_a(S2, G2T)
      // rows 1,3 (SOLVER/unrolled_loops.f90:76-121).  The lane's entries are
      // selected once: lanes outside the sparse pattern take an exact 0.
      const bool cgp = (i == 1 || i == 3) && (j == 1 || j == 3);
      const int q = cgp ? ((i == 1 ? 0 : 2) + (j == 1 ? 0 : 1)) : 0;
      const float yl = cs[0 * 4 + q];
      const bool inb = (j == 1 || j == 3), ina = (i == 1 || i == 3);
      const int qb0 = (j == 3) ? 1 : 0, qa0 = (i == 3) ? 2 : 0;
      const float gb0 = G.g[i + 5], gb1 = G.g[i + 15];
      const float ga0 = G.g2T[1 + 5 * j], ga1 = G.g2T[3 + 5 * j];
      // the B wavefields' entries of Sa and R are adjacent: read together
      auto spbv = [&](int s1, float (&o)[B]) {
        float x0[B], x1[B];
        ldv<B>(Sa + (s1 * 4 + qb0) * B, x0);
        ldv<B>(Sa + (s1 * 4 + qb0 + 2) * B, x1);
#pragma unroll
        for (int b = 0; b < B; b++) {
          const float v = gb0 * x0[b] + gb1 * x1[b];
          o[b] = inb ? v : 0.0f;
        }
      };
      auto spav = [&](int s2, float (&o)[B]) {
        float x0[B], x1[B];
        ldv<B>(Sa + (s2 * 4 + qa0) * B, x0);
        ldv<B>(Sa + (s2 * 4 + qa0 + 1) * B, x1);
#pragma unroll
        for (int b = 0; b < B; b++) {
          const float v = x0[b] * ga0 + x1[b] * ga1;
          o[b] = ina ? v : 0.0f;
        }
      };
      float vr2[B], vr4[B], vr6[B], pb0[B], pa1[B], pb2[B], pa3[B], pb4[B], pa5[B];
      ldv<B>(R + (1 * 4 + q) * B, vr2);
      ldv<B>(R + (3 * 4 + q) * B, vr4);
      ldv<B>(R + (5 * 4 + q) * B, vr6);
      spbv(0, pb0);
      spav(1, pa1);
      if (SRC != AXH_MONOPOLE) {
        spbv(2, pb2);
        spav(3, pa3);
      }
      spbv(4, pb4);
      spav(5, pa5);
#pragma unroll
      for (int b = 0; b < B; b++) {
        const float r2 = vr2[b], r4 = vr4[b], r6 = vr6[b];
        if (SRC == AXH_MONOPOLE) {
          float l1 = pb0[b] + pa1[b];
          l1 = cgp ? l1 + yl * r2 : l1;
          const float l3 = pb4[b] + pa5[b];
          acc[0][b] = acc[0][b] - l1;
          acc[2][b] = acc[2][b] - l3;
        } else if (SRC == AXH_DIPOLE) {
          const float lp = pb0[b] + pa1[b];
          float lm = pb2[b] + pa3[b];
          float lz = pb4[b] + pa5[b];
          lm = cgp ? lm + 2.0f * yl * (r2 - r6) : lm;
          lz = cgp ? lz - yl * r4 : lz;
          acc[0][b] = acc[0][b] - lp;
          acc[1][b] = acc[1][b] - lm;
          acc[2][b] = acc[2][b] - lz;
        } else {
          float ls = pb0[b] + pa1[b];
          float lp = -pb2[b] - pa3[b];
          float lz = pb4[b] + pa5[b];
          ls = cgp ? ls + yl * (r2 - 2.0f * r6) : ls;
          lp = cgp ? lp + yl * (r6 - 2.0f * r2) : lp;
          lz = cgp ? lz - 2.0f * yl * r4 : lz;
          acc[0][b] = acc[0][b] - ls;
          acc[1][b] = acc[1][b] - lp;
          acc[2][b] = acc[2][b] - lz;
        }
      }
    }
    if (MODE == 0) {
      // time_step_memvars_cg4 with the masked predicted displacement; lane
      // k: cg point k % 4 of wavefield k / 4
      float *src = &sSrc[le * NSRC * B], *old = &sOld[le * NSRC * B];
      const float *V1 = U0, *V2 = (NCOMP == 2 ? U0 : U1), *V3 = (NCOMP == 2 ? U1 : U2);
      if (valid && k < 4 * B) {
        // lane k: cg point k % 4 of wavefield k / 4
        memvar_prepare<SRC, B, NSRC>(A, sG, axial, k >> 2, e, k & 3, V1, V2, V3, src, old);
      }
      __syncthreads();
      // time_step_memvars_cg4's update of the entries (k, b) this lane read
      // at the start (SOLVER/attenuation.f90:130-155); the monopole's
      // components 3 and 5 are not updated
      const int c = k >> 2, q = k & 3;
      if (valid && k < 24 && !(SRC == AXH_MONOPOLE && (c == 3 || c == 5))) {
        float *mv = A.memvar + (size_t)e * 24 * A.nsls * B;
        const int sk = 4 + 4 * src_slot<NSRC>(c) + q;
        double sd[B], od[B], st[B], ot[B];
#pragma unroll
        for (int b = 0; b < B; b++) {
          sd[b] = (double)src[sk * B + b];
          od[b] = (double)old[sk * B + b];
          st[b] = (double)src[q * B + b];
          ot[b] = (double)old[q * B + b];
        }
#pragma unroll
        for (int jj = 0; jj < (MV ? MS : 1); jj++) {
          if (jj >= A.nsls) break;
          const double E = A.expw[jj];
          const double *cf = sCf[MV ? le : 0][MV ? jj : 0];
          float nv[B];
#pragma unroll
          for (int b = 0; b < B; b++) {
            const float db = (float)(cf[0] * sd[b] + cf[1] * od[b]);
            if (c < 3) {
              const float trb = (float)(cf[2] * st[b] + cf[3] * ot[b]);
              nv[b] = (float)((E * (double)mvr[jj][b] + (double)db) + (double)trb);
            } else {
              nv[b] = (float)(E * (double)mvr[jj][b] + (double)db);
            }
          }
          stv<B>(mv + (jj * 24 + k) * B, nv);
        }
      }
    }
  }

  if (valid) {
    if (direct) {
      // the point's only copy: assembly (0 + f, as the gather sums), kick
      // and drift as in k_asm_solid
      const float mi = dmi;
#pragma unroll
      for (int cc = 0; cc < NCOMP; cc++) {
        const int c = (NCOMP == 2 && cc == 1) ? 2 : cc;
        const size_t p = ((size_t)c * A.nglob + gp) * B;
        float *v = dv[cc];
        if (MODE == 0) {
          float a[B];
          const float *a0 = da[MODE == 0 ? cc : 0];
#pragma unroll
          for (int b = 0; b < B; b++) {
            const float sm = 0.0f + acc[c][b];
            if (c == 2 && SRC == AXH_DIPOLE)
              a[b] = (float)(-2.0 * (double)mi * (double)sm);
            else
              a[b] = -mi * sm;
            v[b] = (float)((double)v[b] + A.hdt * (double)(a0[b] + a[b]));
          }
          stv<B>(A.acc0 + p, a);
          stv<B>(A.velo + p, dv[cc]);
        } else {
          float u[B];
#pragma unroll
          for (int b = 0; b < B; b++) {
            const float a = -mi * (0.0f + acc[c][b]);
            if (c == 2 && SRC == AXH_DIPOLE)
              v[b] = (float)((double)v[b] + 2.0 * (double)a * A.cv);
            else
              v[b] = (float)((double)v[b] + (double)a * A.cv);
            u[b] = (float)((double)sU[cc * 256 * B + t * B + b] + (double)v[b] * A.cd_next);
          }
          stv<B>(A.velo + p, dv[cc]);
          stv<B>(A.disp_w + p, u);
        }
      }
    } else {
      stv<B>(A.force + slot * B, acc[0]);
      if (SRC != AXH_MONOPOLE) stv<B>(A.force + ((size_t)A.nslot + slot) * B, acc[1]);
      stv<B>(A.force + (2 * (size_t)A.nslot + slot) * B, acc[2]);
    }
  }
}

// receiver samples of compute_recfile_seis_bare (SOLVER/seismograms.f90:742),
// blockIdx.y = wavefield of the group: at the steps with mod(iter, seis_it)
// == 0, sample q = iter / seis_it goes to slot mod(q - 1, cap) of the
// wavefield's staging ring
struct RecArgs {
  int src, nglob;
  const float *disp;     // [3][nglob][B]
  int nrec[kGroup], seis_it[kGroup], cap[kGroup];
  const int *rec_g[kGroup];
  float *ring[kGroup];   // [cap][nrec][3]
  StepRef step;
};

template <int B>
__device__ __forceinline__ void rec_sample(const RecArgs &A, int r, int b) {
  if (r >= A.nrec[b]) return;
  const int iter = step_of(A.step, b);
  if (iter % A.seis_it[b] != 0) return;
  const int slot = (iter / A.seis_it[b] - 1) % A.cap[b];
  const int g = A.rec_g[b][r];
  const float d1 = A.disp[(size_t)g * B + b], d2 = A.disp[((size_t)A.nglob + g) * B + b],
              d3 = A.disp[(2 * (size_t)A.nglob + g) * B + b];
  float *o = A.ring[b] + ((size_t)slot * A.nrec[b] + r) * 3;
  if (A.src == AXH_MONOPOLE) { o[0] = d1; o[1] = d3; o[2] = 0.0f; }
  else if (A.src == AXH_DIPOLE) { o[0] = d1 + d2; o[1] = d1 - d2; o[2] = d3; }
  else { o[0] = d1; o[1] = d2; o[2] = d3; }
}

// memory-variable update of the symplectic scheme, once per step after the
// final drift (SOLVER/time_evol_wave.F90:727-731); also refreshes the SLS sums
// read by the stage kernels.
// Each lane owns entries r = k + 25 w of its element's [24][B] block for every
// SLS: it converts their sources once, runs the SLS in order for each entry
// and sums the new values over the SLS (from zero, in SLS order) as it goes.
// The products ts_fac_t(jj) a_j and ts_fac_tm1(jj) a_j of the element are
// formed once per element, exactly as the per-entry expressions group them.
// Groups (B > 1): the lane's memory variables (NR entries of every SLS,
// MS >= nsls) are requested into registers before anything else, so their
// latency overlaps the displacement gather and the strain sources, and the
// workgroup needs no LDS for them (more workgroups per CU).  One wavefield:
// the tile's memory variables are staged in dynamic LDS (10 * 24 * nsls
// floats); measured, this is the faster form at B = 1 (36 vs 43 us per step
// on the 10 s mesh), the registers at B = 4 (127 vs 158 us).
// Blocks past the solid tiles take the step's receiver samples (R; nothing
// in the update writes the displacement they read), which saves the step a
// launch.
template <int SRC, int B, int MS>
__global__ __launch_bounds__(256) void k_memvar(SolidArgs A, RecArgs R) {
  constexpr int NCOMP = (SRC == AXH_MONOPOLE) ? 2 : 3;
  constexpr int NR = (24 * B + 24) / 25;  // entries per lane
  constexpr bool REG = B > 1;             // memory variables in registers
  extern __shared__ __align__(16) float sMv[];
  __shared__ __align__(16) float sU[NCOMP * 256 * B];
  __shared__ __align__(16) float sG[105];
  __shared__ __align__(16) float sSrc[kElemPerTile * 28 * B];
  __shared__ __align__(16) float sOld[kElemPerTile * 28 * B];
  __shared__ double sCf[kElemPerTile][kMaxSls][4];
  const int t = threadIdx.x, tile = blockIdx.x;
  if (tile >= A.ntile) {  // block-uniform: no barrier below is reached
    const int q = tile - A.ntile;
    rec_sample<B>(R, (q / B) * 256 + t, q % B);
    return;
  }
  const int le = t / 25, k = t - 25 * le;
  const int e = tile * kElemPerTile + le;
  const bool valid = (t < 250) && (e < A.nel);
  const size_t slot = (size_t)tile * 256 + t;
  const int nsls = A.nsls;
  const int n = 24 * nsls * B;
  const int eb = le * 25 * B;
  float *memvar = A.memvar + (size_t)(valid ? e : 0) * n;
  float mvr[REG ? NR : 1][REG ? MS : 1];
  if constexpr (REG) {
    if (valid) {
#pragma unroll
      for (int w = 0; w < NR; w++) {
        const int r = k + 25 * w;
        if (r >= 24 * B) continue;
#pragma unroll
        for (int jj = 0; jj < MS; jj++)
          if (jj < nsls) mvr[w][jj] = memvar[jj * 24 * B + r];
      }
    }
  }
  stage_g(A.G, sG);
  gather_disp<NCOMP, B>(A, valid, valid ? A.gid[slot] : 0, t, sU);
  const float *mv = &sMv[le * n];
  if constexpr (!REG) {
    // the tile's memory variables are one contiguous block (n is a multiple
    // of 24): copied by the whole workgroup in 16-byte pieces
    const int e0 = tile * kElemPerTile;
    const int n4 = std::min(kElemPerTile, A.nel - e0) * (n / 4);
    const float4 *src4 = reinterpret_cast<const float4 *>(A.memvar + (size_t)e0 * n);
    float4 *dst4 = reinterpret_cast<float4 *>(sMv);
    for (int x = t; x < n4; x += 256) dst4[x] = src4[x];
  }
  if (valid && k >= 4 * B && k < 4 * B + nsls) {
    const int jj = k - 4 * B;
    const double amu = A.aj_mu[(size_t)e * nsls + jj], ak = A.aj_k[(size_t)e * nsls + jj];
    sCf[le][jj][0] = A.tst[jj] * amu;
    sCf[le][jj][1] = A.tstm1[jj] * amu;
    sCf[le][jj][2] = A.tst[jj] * ak;
    sCf[le][jj][3] = A.tstm1[jj] * ak;
  }
  __syncthreads();
  const bool axial = valid && A.einfo[e] >= 0;
  float *src = &sSrc[le * 28 * B], *old = &sOld[le * 28 * B];
  const float *U0 = sU + eb, *U1 = sU + 256 * B + eb, *U2 = sU + 2 * 256 * B + eb;
  const float *V1 = U0, *V2 = (NCOMP == 2 ? U0 : U1), *V3 = (NCOMP == 2 ? U1 : U2);
  if (valid && k < 4 * B) {
    // lane k: cg point k % 4 of wavefield k / 4
    memvar_prepare<SRC, B>(A, sG, axial, k >> 2, e, k & 3, V1, V2, V3, src, old);
  }
  __syncthreads();
  if (!valid) return;
  // per entry: its sources (converted once) and its running SLS sum; the SLS
  // loop is outside so that the lane's entries proceed side by side
  double sd[NR], od[NR], st[NR], ot[NR];
  float rs[NR];
  bool upd[NR];
#pragma unroll
  for (int w = 0; w < NR; w++) {
    const int r = k + 25 * w;
    const int cq = r / B, b = r - cq * B;
    const int q = cq & 3, c = cq >> 2;
    upd[w] = r < 24 * B && !(SRC == AXH_MONOPOLE && (c == 3 || c == 5));
    rs[w] = 0.0f;
    if (upd[w]) {
      sd[w] = (double)src[(4 + cq) * B + b];
      od[w] = (double)old[(4 + cq) * B + b];
      st[w] = (double)src[q * B + b];
      ot[w] = (double)old[q * B + b];
    }
  }
#pragma unroll
  for (int jj = 0; jj < MS; jj++) {
    if (jj >= nsls) break;
    const double E = A.expw[jj];
    const double c0 = sCf[le][jj][0], c1 = sCf[le][jj][1], c2 = sCf[le][jj][2],
                 c3 = sCf[le][jj][3];
#pragma unroll
    for (int w = 0; w < NR; w++) {
      const int r = k + 25 * w;
      if (r >= 24 * B) continue;
      const int m = jj * 24 * B + r;
      if (!upd[w]) {
        // not updated for a monopole source: summed as they are
        rs[w] = rs[w] + (REG ? mvr[REG ? w : 0][REG ? jj : 0] : mv[m]);
        continue;
      }
      const int c = (r / B) >> 2;
      const double mvo = (double)(REG ? mvr[REG ? w : 0][REG ? jj : 0] : mv[m]);
      const float db = (float)(c0 * sd[w] + c1 * od[w]);
      float nv;
      if (c < 3) {
        const float trb = (float)(c2 * st[w] + c3 * ot[w]);
        nv = (float)((E * mvo + (double)db) + (double)trb);
      } else {
        nv = (float)(E * mvo + (double)db);
      }
      memvar[m] = nv;
      rs[w] = rs[w] + nv;
    }
  }
#pragma unroll
  for (int w = 0; w < NR; w++) {
    const int r = k + 25 * w;
    if (r < 24 * B) A.rsum[(size_t)e * 24 * B + r] = rs[w];
  }
}

// ---------------------------------------------------------------------------
// Memory variables at all 25 points of an element (att_coarse_grained
// false), one wavefield: the anelastic forces of glob_anel_stiffness_*_4
// (SOLVER/stiffness_mono.f90:409-507, stiffness_di.f90:604-723,
// stiffness_quad.f90:555-668), subtracted from the element forces the
// stiffness kernel wrote (before the assembly adds the coupling terms, the
// reference's order), and time_step_memvars_4 (SOLVER/attenuation.f90:210-334)
// with the strain of compute_strain_att_el_4 (attenuation.f90:542-606).  One
// block per solid tile, a lane per element point (i = k % 5, j = k / 5).
// Newmark runs both parts in one launch (the update reads the displacement the
// forces were computed from); a symplectic step runs the forces after every
// stage and the update once after the last drift.  The update also stores the
// sums over the SLS of the new memory variables (from zero in SLS order, the
// sums the forces form), so the stage launches read 6 floats per point instead
// of 6 n_sls.
// ---------------------------------------------------------------------------
constexpr int kAfTerms = 12;
enum { AF_Y = 0, AF_VSE, AF_VSX, AF_VZE, AF_VZX, AF_DSDE, AF_DZDE, AF_DSDX, AF_DZDX, AF_INVS,
       AF_DMU, AF_DK };

struct AnelFullArgs {
  int nel, nsls, nglob, nslot;
  const float *G;         // G0 | G1 | G1T | G2 | G2T
  const float *terms;     // [tile][kAfTerms][256]
  const float *axis;      // [nel][5][5]: Y0, V0_s_eta, V0_s_xi, V0_z_eta, V0_z_xi
  const int *einfo, *gid;
  const double *aj_mu, *aj_k;  // [nel][nsls]
  double expw[kMaxSls], tst[kMaxSls], tstm1[kMaxSls];
  const float *disp;      // [3][nglob][B]
  float *force;           // [3][nslot][B]
  float *memvar;          // [tile][nsls][6][256][B]
  float *trm1;            // [tile][256][B]
  float *devm1;           // [tile][6][256][B]
  float *rsum;            // [tile][6][256][B]: sums over the SLS
};

template <int SRC, bool STIFF, bool UPD, int B>
__global__ __launch_bounds__(256) void k_anel_full(AnelFullArgs A) {
  __shared__ float sG[105];
  __shared__ float sA[6][256 * B];  // S terms (forces), then the strain's fields
  __shared__ float sR[6][256 * B];  // sums of the memory variables
  const int t = threadIdx.x, tile = blockIdx.x;
  const int le = t / 25, k = t - 25 * le;
  const int i = k % 5, j = k / 5;
  const int e = tile * kElemPerTile + le;
  const bool valid = (t < 250) && (e < A.nel);
  const int eb = le * 25;  // the element's first lane
  const size_t slot = (size_t)tile * 256 + t;
  stage_g(A.G, sG);
  const float *G0 = sG, *G1 = sG + 5, *G1T = sG + 30, *G2 = sG + 55, *G2T = sG + 80;
  const bool ax = valid && A.einfo[e] >= 0;
  const float *Gs = ax ? G1 : G2, *GsT = ax ? G1T : G2T;
  auto term = [&](int a) {
    return valid ? A.terms[((size_t)tile * kAfTerms + a) * 256 + t] : 0.0f;
  };
  const size_t cs = (size_t)256 * B;  // one component of a tile
  const size_t mvs = 6 * cs;          // one SLS of a tile
  float *const mv = A.memvar + (size_t)tile * A.nsls * mvs + (size_t)t * B;
  float *const prs = A.rsum + (size_t)tile * mvs + (size_t)t * B;
  const int tb = t * B;
  if (STIFF) {
    float r[6][B];
#pragma unroll
    for (int c = 0; c < 6; c++) {
#pragma unroll
      for (int w = 0; w < B; w++) r[c][w] = 0.0f;
      if (valid && (SRC != AXH_MONOPOLE || (c != 3 && c != 5))) {
        if (UPD) {
          for (int q = 0; q < A.nsls; q++) {
            float x[B];
            ldv<B>(mv + q * mvs + c * cs, x);
#pragma unroll
            for (int w = 0; w < B; w++) r[c][w] = r[c][w] + x[w];
          }
        } else {
          ldv<B>(prs + c * cs, r[c]);
        }
      }
      stv<B>(&sR[c][tb], r[c]);
    }
    const float yl = term(AF_Y), vse = term(AF_VSE), vsx = term(AF_VSX), vze = term(AF_VZE),
                vzx = term(AF_VZX);
    {
      float S[6][B];
#pragma unroll
      for (int w = 0; w < B; w++) {
        const float r1 = r[0][w], r3 = r[2][w], r4 = r[3][w], r5 = r[4][w], r6 = r[5][w];
        if (SRC == AXH_MONOPOLE) {
          S[0][w] = vze * r1 + vse * r5;
          S[1][w] = vzx * r1 + vsx * r5;
          S[2][w] = 0.0f;
          S[3][w] = 0.0f;
        } else if (SRC == AXH_DIPOLE) {
          S[0][w] = vze * (r1 - r6) + vse * (r5 - r4);
          S[1][w] = vzx * (r1 - r6) + vsx * (r5 - r4);
          S[2][w] = vze * (r1 + r6) + vse * (r5 + r4);
          S[3][w] = vzx * (r1 + r6) + vsx * (r5 + r4);
        } else {
          S[0][w] = vze * r1 + vse * r5;
          S[1][w] = vzx * r1 + vsx * r5;
          S[2][w] = vze * r6 + vse * r4;
          S[3][w] = vzx * r6 + vsx * r4;
        }
        S[4][w] = vze * r5 + vse * r3;
        S[5][w] = vzx * r5 + vsx * r3;
      }
#pragma unroll
      for (int f = 0; f < 6; f++)
        if (SRC != AXH_MONOPOLE || (f != 2 && f != 3)) stv<B>(&sA[f][tb], S[f]);
    }
    __syncthreads();
    if (valid) {
      const float *av = A.axis + (size_t)e * 25;
      const int p0 = eb + 5 * j;
      float l1[B], l2[B], l3[B];
#pragma unroll
      for (int w = 0; w < B; w++) {
        const int o = eb * B + w;
        const float r2 = r[1][w], r4 = r[3][w], r6 = r[5][w];
        const float X1 = amul<B>(Gs, &sA[0][o], i, j), X2 = mulb<B>(&sA[1][o], G2T, i, j);
        const float X5 = amul<B>(Gs, &sA[4][o], i, j), X6 = mulb<B>(&sA[5][o], G2T, i, j);
        if (SRC == AXH_MONOPOLE) {
          l1[w] = X1 + X2 + yl * r2;
          l2[w] = 0.0f;
          l3[w] = X5 + X6;
        } else {
          const float X3 = amul<B>(Gs, &sA[2][o], i, j), X4 = mulb<B>(&sA[3][o], G2T, i, j);
          if (SRC == AXH_DIPOLE) {
            l1[w] = X1 + X2;
            l2[w] = X3 + X4 + 2.0f * yl * (r2 - r6);
            l3[w] = X5 + X6 - yl * r4;
          } else {
            l1[w] = X1 + X2 + yl * (r2 - 2.0f * r6);
            l2[w] = -X3 - X4 + yl * (r6 - 2.0f * r2);
            l3[w] = X5 + X6 - 2.0f * yl * r4;
          }
        }
        if (ax) {
          // the axial vectors at this lane's j, from the sums at (0, j)
          const float y0 = av[j], v0se = av[5 + j], v0sx = av[10 + j], v0ze = av[15 + j],
                      v0zx = av[20 + j];
          const int q0 = p0 * B + w;
          const float r10 = sR[0][q0], r20 = sR[1][q0], r30 = sR[2][q0], r40 = sR[3][q0],
                      r50 = sR[4][q0], r60 = sR[5][q0];
          const float g0 = G0[i];
          if (SRC == AXH_MONOPOLE) {
            const float V1 = v0ze * r10 + v0se * r50 + y0 * r20;
            const float V2 = v0ze * r50 + v0se * r30;
            l1[w] = l1[w] + g0 * V1;
            l3[w] = l3[w] + g0 * V2;
            if (i == 0) {
              // vxm_4(V3, G2T): V3 at every point (0, kk) of the element
              float V4 = 0.0f;
              for (int kk = 0; kk < 5; kk++) {
                const int pk = (eb + 5 * kk) * B + w;
                const float V3 = av[20 + kk] * sR[4][pk] + av[10 + kk] * sR[2][pk];
                V4 = V4 + V3 * G2T[kk + 5 * j];
              }
              l3[w] = l3[w] + V4;
            }
          } else if (SRC == AXH_DIPOLE) {
            const float V1 = v0ze * (r10 - r60) + v0se * (r50 - r40);
            const float V2 = v0zx * (r10 - r60) + v0sx * (r50 - r40);
            l1[w] = l1[w] + g0 * V1;
            if (i == 0) {
              float V3 = 0.0f;
              for (int kk = 0; kk < 5; kk++) {
                const int pk = (eb + 5 * kk) * B + w;
                const float V2k = av[20 + kk] * (sR[0][pk] - sR[5][pk]) +
                                  av[10 + kk] * (sR[4][pk] - sR[3][pk]);
                V3 = V3 + V2k * G2T[kk + 5 * j];
              }
              l1[w] = l1[w] + V3;
            }
            const float Vm = v0ze * (r10 + r60) + v0se * (r50 + r40) + y0 * 2.0f * (r20 - r60);
            l2[w] = l2[w] + g0 * Vm;
            // the z-component adds outerprod(G0, V2) (stiffness_di.f90:706-707)
            l3[w] = l3[w] + g0 * V2;
          } else {
            const float Vs = v0ze * r10 + y0 * (r20 - 2.0f * r60);
            const float Vp = -(v0ze * r60) + y0 * (r60 - 2.0f * r20);
            const float Vz = v0se * r30;
            l1[w] = l1[w] + g0 * Vs;
            l2[w] = l2[w] + g0 * Vp;
            l3[w] = l3[w] + g0 * Vz;
          }
        }
      }
      float x[B];
      float *f0 = A.force + slot * B, *f2 = A.force + (2 * (size_t)A.nslot + slot) * B;
      ldv<B>(f0, x);
#pragma unroll
      for (int w = 0; w < B; w++) x[w] = x[w] - l1[w];
      stv<B>(f0, x);
      if (SRC != AXH_MONOPOLE) {
        float *f1 = A.force + ((size_t)A.nslot + slot) * B;
        ldv<B>(f1, x);
#pragma unroll
        for (int w = 0; w < B; w++) x[w] = x[w] - l2[w];
        stv<B>(f1, x);
      }
      ldv<B>(f2, x);
#pragma unroll
      for (int w = 0; w < B; w++) x[w] = x[w] - l3[w];
      stv<B>(f2, x);
    }
    if (UPD) __syncthreads();  // sA is reused below
  }
  if (UPD) {
    const int g = valid ? A.gid[slot] : 0;
    float u1[B], u2[B], u3[B];
#pragma unroll
    for (int w = 0; w < B; w++) u1[w] = u2[w] = u3[w] = 0.0f;
    if (valid) {
      ldv<B>(A.disp + (size_t)g * B, u1);
      if (SRC != AXH_MONOPOLE) ldv<B>(A.disp + ((size_t)A.nglob + g) * B, u2);
      ldv<B>(A.disp + (2 * (size_t)A.nglob + g) * B, u3);
    }
    // fields whose gradients the strain needs (f1 = u_z throughout)
#pragma unroll
    for (int w = 0; w < B; w++) {
      if (SRC == AXH_MONOPOLE) {
        sA[0][tb + w] = u1[w];
        sA[1][tb + w] = u3[w];
      } else if (SRC == AXH_DIPOLE) {
        sA[0][tb + w] = u1[w] + u2[w];
        sA[1][tb + w] = u3[w];
        sA[2][tb + w] = u2[w];
        sA[3][tb + w] = u1[w] - u2[w];
      } else {
        sA[0][tb + w] = u1[w];
        sA[1][tb + w] = u3[w];
        sA[2][tb + w] = u2[w];
        sA[3][tb + w] = u1[w] - 2.0f * u2[w];
        sA[4][tb + w] = u2[w] - 2.0f * u1[w];
      }
    }
    __syncthreads();
    if (!valid) return;  // no barrier follows
    const float dse = term(AF_DSDE), dze = term(AF_DZDE), dsx = term(AF_DSDX),
                dzx = term(AF_DZDX), is = term(AF_INVS);
    const float dm = term(AF_DMU), dk = term(AF_DK);
    const double third = 1.0 / 3.0;
    float str[B], sdev[6][B];
#pragma unroll
    for (int w = 0; w < B; w++) {
      const int o = eb * B + w;
      // axisym_gradient_solid_el_4: d/ds and d/dz of field f at this point
      auto grad = [&](int f, float &ds, float &dz) {
        const float m1 = amul<B>(GsT, &sA[f][o], i, j);
        const float m2 = mulb<B>(&sA[f][o], G2, i, j);
        ds = dze * m1 + dzx * m2;
        dz = dse * m1 + dsx * m2;
      };
      // f_over_s_solid_el_4: f/s, and d/ds on the axis (dsdf_elem_solid)
      auto fos = [&](int f, float val) {
        if (ax && i == 0) {
          float ds, dz;
          grad(f, ds, dz);
          return ds;
        }
        return is * val;
      };
      float E[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      float b1s, b1z, b2s, b2z;
      grad(0, b1s, b1z);
      grad(1, b2s, b2z);
      E[0] = b1s;
      E[2] = b2z;
      E[4] = b1z + b2s;
      if (SRC == AXH_MONOPOLE) {
        E[1] = fos(0, u1[w]);
      } else if (SRC == AXH_DIPOLE) {
        E[1] = 2.0f * fos(2, u2[w]);
        float ps, pz;
        grad(3, ps, pz);
        const float f3 = fos(1, u3[w]);
        E[3] = -f3 - pz;
        E[5] = -E[1] - ps;
      } else {
        E[1] = fos(3, u1[w] - 2.0f * u2[w]);
        float ps, pz;
        grad(2, ps, pz);
        const float f3 = fos(1, u3[w]);
        const float f4 = fos(4, u2[w] - 2.0f * u1[w]);
        E[3] = -(2.0f * f3) - pz;
        E[5] = f4 - ps;
      }
      const float tr = E[0] + E[1] + E[2];
      str[w] = dk * tr;
#pragma unroll
      for (int c = 0; c < 6; c++) sdev[c][w] = 0.0f;
#pragma unroll
      for (int c = 0; c < 3; c++)
        sdev[c][w] = (float)((double)(dm * 2.0f) * ((double)E[c] - (double)tr * third));
      sdev[4][w] = dm * E[4];
      if (SRC != AXH_MONOPOLE) {
        sdev[3][w] = dm * E[3];
        sdev[5][w] = dm * E[5];
      }
    }
    float *ptr = A.trm1 + slot * B;
    float *pdev = A.devm1 + (size_t)tile * mvs + (size_t)t * B;
    float trm1[B], devm1[6][B];
    ldv<B>(ptr, trm1);
#pragma unroll
    for (int c = 0; c < 6; c++) ldv<B>(pdev + c * cs, devm1[c]);
    const double *amu = A.aj_mu + (size_t)A.nsls * e, *akap = A.aj_k + (size_t)A.nsls * e;
    float rs[6][B];
#pragma unroll
    for (int c = 0; c < 6; c++)
#pragma unroll
      for (int w = 0; w < B; w++) rs[c][w] = 0.0f;
    for (int q = 0; q < A.nsls; q++) {
      const double Tt = A.tst[q], Tm = A.tstm1[q], Ew = A.expw[q];
      const double ak = akap[q], am = amu[q];
      float trb[B];
#pragma unroll
      for (int w = 0; w < B; w++)
        trb[w] = (float)(Tt * ak * (double)str[w] + Tm * ak * (double)trm1[w]);
      float *m = mv + q * mvs;
#pragma unroll
      for (int c = 0; c < 6; c++) {
        if (SRC == AXH_MONOPOLE && (c == 3 || c == 5)) continue;
        float x[B];
        ldv_nt<B>(m + c * cs, x);
#pragma unroll
        for (int w = 0; w < B; w++) {
          const float db =
              (float)(Tt * am * (double)sdev[c][w] + Tm * am * (double)devm1[c][w]);
          const double old = (double)x[w];
          x[w] = c < 3 ? (float)((Ew * old + (double)db) + (double)trb[w])
                       : (float)(Ew * old + (double)db);
          rs[c][w] = rs[c][w] + x[w];
        }
        stv_nt<B>(m + c * cs, x);
      }
    }
    stv<B>(ptr, str);
#pragma unroll
    for (int c = 0; c < 6; c++) {
      stv<B>(pdev + c * cs, sdev[c]);
      if (!STIFF) stv<B>(prs + c * cs, rs[c]);  // read by the symplectic stages
    }
  }
}

// ---------------------------------------------------------------------------
// predictor / drift per global point (SOLVER/time_evol_wave.F90:357-367,
// 597-605, 720-725) with the axis masks (SOLVER/apply_masks.f90) and the
// blow-up test of runtime_info (SOLVER/time_evol_wave.F90:1042)
// ---------------------------------------------------------------------------
struct PredArgs {
  int n, mask0, mask1;   // masked component range [mask0, mask1]
  int comps;             // bit c set: component c is stepped
  int newmark, apply_mask, check_blow, stride, naxb;
  double dt, hdt2, cd;
  const uint8_t *flags;
  const int *axb_g;      // masked axis points on the solid-fluid boundary
  float *u, *v;          // [comp][stride][B]
  const float *a0;
  float *axb_val;        // unmasked prediction of the axb points [naxb][3][B]
  int *blow[kGroup];
  double thr[kGroup];
  StepRef step;
};

template <int B>
__device__ __forceinline__ void predict_point(const PredArgs &A, int g) {
  const uint8_t fl = A.flags[g];
  for (int c = 0; c < 3; c++) {
    if (!(A.comps & (1 << c))) continue;
    const size_t p = ((size_t)c * A.stride + g) * B;
    float u[B], v[B], un[B];
    ldv<B>(A.u + p, u);
    ldv<B>(A.v + p, v);
    if (A.check_blow && (fl & kFlagBlow)) {
#pragma unroll
      for (int b = 0; b < B; b++)
        if ((double)fabsf(u[b]) > A.thr[b]) atomicMin(A.blow[b], step_of(A.step, b));
    }
    if (A.newmark) {
      float a0[B];
      ldv<B>(A.a0 + p, a0);
#pragma unroll
      for (int b = 0; b < B; b++)
        un[b] = (float)((double)u[b] + A.dt * (double)v[b] + A.hdt2 * (double)a0[b]);
    } else {
#pragma unroll
      for (int b = 0; b < B; b++) un[b] = (float)((double)u[b] + (double)v[b] * A.cd);
    }
    if (fl & kFlagAxisBdry) {
      for (int q = 0; q < A.naxb; q++)
        if (A.axb_g[q] == g) stv<B>(A.axb_val + ((size_t)q * 3 + c) * B, un);
    }
    if (A.apply_mask && (fl & kFlagAxis) && c >= A.mask0 && c <= A.mask1) {
#pragma unroll
      for (int b = 0; b < B; b++) un[b] = 0.0f;
    }
    stv<B>(A.u + p, un);
  }
}

// The solid and the fluid prediction in one launch: the first nblk_s
// workgroups take the solid points, the others the fluid points.
template <int B>
__global__ __launch_bounds__(256) void k_predict(PredArgs S, PredArgs F, int nblk_s) {
  if ((int)blockIdx.x < nblk_s) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g < S.n) predict_point<B>(S, g);
  } else {
    const int g = (blockIdx.x - nblk_s) * 256 + threadIdx.x;
    if (g < F.n) predict_point<B>(F, g);
  }
}

// ---------------------------------------------------------------------------
// assembly + mass + update, per global point
// ---------------------------------------------------------------------------
struct AsmSolidArgs {
  int n, src, newmark, nslot, comps, mask0, mask1;
  int npts;              // points assembled here (the others are kFlagDirect)
  int ncontig;           // pts[ip] == ip for ip < ncontig (not read there)
  const int *pts;
  const int *gstart, *gslot;
  const uint8_t *flags;
  const float *minv;
  const float *gamma;    // absorbing boundaries: damping factor per point, or NULL
  double hdt, cv;
  const float *force;    // [3][nslot][B]
  float *velo, *acc0;    // [3][n][B]
  int nsrc[kGroup];
  const int *src_g[kGroup];
  const float *src_val[kGroup];
  const float *stf[kGroup];  // NULL when the source is not added here
  StepRef step;
  int nstages, stage;
  // symplectic: the drift of the next stage (or the final drift) right after
  // the kick, u = u + v*cd_next, masked when drift_mask is set
  int drift, drift_mask, naxb;
  double cd_next;
  float *disp;           // [3][n][B]
  const int *axb_g;
  float *axb_val;        // [naxb][3][B]
  // solid-fluid coupling (bdry_copy2solid): added to the force of a coupled
  // slot before the slot enters the sum
  int dipole;
  const int *cpl_slot;   // per solid slot: coupling entry or -1
  const int *cpl_gf;     // fluid point of a coupling entry
  const float *cpl_bm;   // [entry][2]
  const float *ddchi;    // [nglob_f][B]
};

// The copies of a point are gathered kAsmChunk at a time with the three
// components together, so that up to twelve loads are in flight per lane; each
// component's sum still runs over the copies in ascending slot order from
// zero.  The point's velocity (and acceleration, displacement) is loaded
// before the gather for the same reason.
template <int B>
constexpr int kAsmChunk = (B >= 3 ? 2 : 4);
template <int B>
__global__ __launch_bounds__(256) void k_asm_solid(AsmSolidArgs A) {
  const int ip = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (ip >= A.npts) return;
  const int g = ip < A.ncontig ? ip : A.pts[ip];
  const uint8_t fl = A.flags[g];
  const int q0 = A.gstart[g], q1 = A.gstart[g + 1];
  const float mi = A.minv[g];
  // a component is gathered unless it is off or axis-masked (masked copies
  // are zero before the gather)
  bool on[3];
#pragma unroll
  for (int c = 0; c < 3; c++)
    on[c] = (A.comps & (1 << c)) && !((fl & kFlagAxis) && c >= A.mask0 && c <= A.mask1);
  float v[3][B], a0[3][B], u[3][B];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (!(A.comps & (1 << c))) continue;
    const size_t p = ((size_t)c * A.n + g) * B;
    ldv<B>(A.velo + p, v[c]);
    if (A.newmark) ldv<B>(A.acc0 + p, a0[c]);
    if (A.drift || A.gamma) ldv<B>(A.disp + p, u[c]);
  }
  // absorbing boundaries: a -= 2 gamma v + gamma^2 u after the inverse mass
  // (SOLVER/time_evol_wave.F90:468-491,690-711), with v and u of this stage
  const float gm = A.gamma ? A.gamma[g] : 0.0f;
  float s[3][B];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int b = 0; b < B; b++) s[c][b] = 0.0f;
  const size_t cstride = (size_t)A.nslot * B;
  constexpr int W = kAsmChunk<B>;
  for (int q = q0; q < q1; q += W) {
    int sl[W];
#pragma unroll
    for (int w = 0; w < W; w++) sl[w] = (q + w < q1) ? A.gslot[q + w] : -1;
    float x[W][3][B];
#pragma unroll
    for (int w = 0; w < W; w++)
#pragma unroll
      for (int c = 0; c < 3; c++)
        if (sl[w] >= 0 && on[c]) ldv<B>(A.force + c * cstride + (size_t)sl[w] * B, x[w][c]);
    if (fl & kFlagCoupled) {
      // bdry_copy2solid (SOLVER/time_evol_wave.F90:1582-1611), added to the
      // coupled slot's force before it enters the sum
#pragma unroll
      for (int w = 0; w < W; w++) {
        const int p = sl[w] >= 0 ? A.cpl_slot[sl[w]] : -1;
        if (p < 0) continue;
        float f2[B];
        ldv<B>(A.ddchi + (size_t)A.cpl_gf[p] * B, f2);
        const float bm0 = A.cpl_bm[2 * p], bm1 = A.cpl_bm[2 * p + 1];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          if (!on[c] || (c == 1 && !A.dipole)) continue;
          const float bmc = (c == 2) ? bm1 : bm0;
#pragma unroll
          for (int b = 0; b < B; b++) x[w][c][b] = x[w][c][b] + bmc * f2[b];
        }
      }
    }
#pragma unroll
    for (int w = 0; w < W; w++)
#pragma unroll
      for (int c = 0; c < 3; c++)
        if (sl[w] >= 0 && on[c])
#pragma unroll
          for (int b = 0; b < B; b++) s[c][b] = s[c][b] + x[w][c][b];
  }
  int si[B];
  float stf[B];
#pragma unroll
  for (int b = 0; b < B; b++) {
    si[b] = -1;
    stf[b] = 0.0f;
    if (fl & kFlagSource) stf[b] = stf_at(A.stf[b], A.step, b, A.nstages, A.stage);
    if (stf[b] != 0.0f) {
      const int *sg = A.src_g[b];
      for (int k = 0; k < A.nsrc[b]; k++)
        if (sg[k] == g) { si[b] = k; break; }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (!(A.comps & (1 << c))) continue;
#pragma unroll
    for (int b = 0; b < B; b++)
      if (si[b] >= 0) s[c][b] = s[c][b] - A.src_val[b][3 * si[b] + c] * stf[b];
    const size_t p = ((size_t)c * A.n + g) * B;
    if (A.newmark) {
      float a[B];
#pragma unroll
      for (int b = 0; b < B; b++) {
        if (c == 2 && A.src == AXH_DIPOLE)
          a[b] = (float)(-2.0 * (double)mi * (double)s[c][b]);
        else
          a[b] = -mi * s[c][b];
        if (A.gamma) a[b] = (a[b] - (2.0f * gm) * v[c][b]) - (gm * gm) * u[c][b];
        v[c][b] = (float)((double)v[c][b] + A.hdt * (double)(a0[c][b] + a[b]));
      }
      stv<B>(A.acc0 + p, a);
    } else {
#pragma unroll
      for (int b = 0; b < B; b++) {
        float a = -mi * s[c][b];
        if (A.gamma) a = (a - (2.0f * gm) * v[c][b]) - (gm * gm) * u[c][b];
        if (c == 2 && A.src == AXH_DIPOLE)
          v[c][b] = (float)((double)v[c][b] + 2.0 * (double)a * A.cv);
        else
          v[c][b] = (float)((double)v[c][b] + (double)a * A.cv);
      }
    }
    stv<B>(A.velo + p, v[c]);
    if (A.drift) {
      // as k_predict's drift, with the velocity just computed
#pragma unroll
      for (int b = 0; b < B; b++)
        u[c][b] = (float)((double)u[c][b] + (double)v[c][b] * A.cd_next);
      if (fl & kFlagAxisBdry) {
        for (int k = 0; k < A.naxb; k++)
          if (A.axb_g[k] == g) stv<B>(A.axb_val + ((size_t)k * 3 + c) * B, u[c]);
      }
      if (A.drift_mask && (fl & kFlagAxis) && c >= A.mask0 && c <= A.mask1) {
#pragma unroll
        for (int b = 0; b < B; b++) u[c][b] = 0.0f;
      }
      stv<B>(A.disp + p, u[c]);
    }
  }
}

struct AsmFluidArgs {
  int n, newmark, mask_axis, mask_surface;
  const int *gstart, *gslot;
  const uint8_t *flags;
  const float *minv;
  const float *gamma;        // absorbing boundaries: damping factor per point, or NULL
  double hdt, cv;
  const float *force;        // [nslot_f][B]
  float *dchi, *ddchi;       // [n][B]
  // symplectic: the drift of the next stage (or the final drift) right after
  // the kick, chi = chi + dchi*cd_next, axis-masked when drift_mask is set
  int drift, drift_mask;
  double cd_next;
  float *chi;                // [n][B]
  // solid-fluid coupling (bdry_copy2fluid): added to the force of a coupled
  // slot before the slot enters the sum
  int dipole, nglob_s;
  const int *cpl_slot;       // per fluid slot: coupling entry or -1
  const int *cpl_gs;         // solid point of a coupling entry
  const int *cpl_axb;        // its side-buffer slot or -1
  const float *cpl_bm;       // [entry][2]
  const float *disp;         // [3][nglob_s][B], the stage's prediction
  const float *axb_val;      // [naxb][3][B]
};

template <int B>
__global__ __launch_bounds__(256) void k_asm_fluid(AsmFluidArgs A) {
  const int g = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (g >= A.n) return;
  const uint8_t fl = A.flags[g];
  const bool masked = (A.mask_axis && (fl & kFlagAxis)) ||
                      (A.mask_surface && (fl & kFlagSurface));
  float s[B];
#pragma unroll
  for (int b = 0; b < B; b++) s[b] = 0.0f;
  if (!masked) {
    // copies gathered four at a time (loads in flight together), summed in
    // ascending slot order from zero
    const bool coupled = fl & kFlagCoupled;
    const int q1 = A.gstart[g + 1];
    for (int q = A.gstart[g]; q < q1; q += 4) {
      int sl[4];
#pragma unroll
      for (int w = 0; w < 4; w++) sl[w] = (q + w < q1) ? A.gslot[q + w] : -1;
      float x[4][B];
#pragma unroll
      for (int w = 0; w < 4; w++)
        if (sl[w] >= 0) ldv<B>(A.force + (size_t)sl[w] * B, x[w]);
      if (coupled) {
        // bdry_copy2fluid (SOLVER/time_evol_wave.F90:1532-1580) with the
        // prediction before the axis mask
#pragma unroll
        for (int w = 0; w < 4; w++) {
          const int p = sl[w] >= 0 ? A.cpl_slot[sl[w]] : -1;
          if (p < 0) continue;
          const int gs = A.cpl_gs[p], kq = A.cpl_axb[p];
          float s1[B], s2[B], s3[B];
          if (kq >= 0) {
            ldv<B>(A.axb_val + ((size_t)kq * 3 + 0) * B, s1);
            ldv<B>(A.axb_val + ((size_t)kq * 3 + 1) * B, s2);
            ldv<B>(A.axb_val + ((size_t)kq * 3 + 2) * B, s3);
          } else {
            ldv<B>(A.disp + (size_t)gs * B, s1);
            ldv<B>(A.disp + ((size_t)A.nglob_s + gs) * B, s2);
            ldv<B>(A.disp + (2 * (size_t)A.nglob_s + gs) * B, s3);
          }
          const float bm1 = A.cpl_bm[2 * p], bm2 = A.cpl_bm[2 * p + 1];
#pragma unroll
          for (int b = 0; b < B; b++) {
            if (A.dipole)
              x[w][b] = x[w][b] - bm1 * (s1[b] + s2[b]) - bm2 * s3[b];
            else
              x[w][b] = x[w][b] - bm1 * s1[b] - bm2 * s3[b];
          }
        }
      }
#pragma unroll
      for (int w = 0; w < 4; w++)
        if (sl[w] >= 0)
#pragma unroll
          for (int b = 0; b < B; b++) s[b] = s[b] + x[w][b];
    }
  }
  const float mi = A.minv[g];
  float d[B], dd[B], c[B];
  ldv<B>(A.dchi + (size_t)g * B, d);
  if (A.drift || A.gamma) ldv<B>(A.chi + (size_t)g * B, c);
  // absorbing boundaries: ddchi -= 2 gamma dchi + gamma^2 chi after the
  // inverse mass (SOLVER/time_evol_wave.F90:432-434,663-665)
  const float gm = A.gamma ? A.gamma[g] : 0.0f;
  if (A.newmark) {
    float d0[B];
    ldv<B>(A.ddchi + (size_t)g * B, d0);
#pragma unroll
    for (int b = 0; b < B; b++) {
      dd[b] = -mi * s[b];
      if (A.gamma) dd[b] = (dd[b] - (2.0f * gm) * d[b]) - (gm * gm) * c[b];
      d[b] = (float)((double)d[b] + A.hdt * (double)(d0[b] + dd[b]));
    }
  } else {
#pragma unroll
    for (int b = 0; b < B; b++) {
      dd[b] = -s[b] * mi;
      if (A.gamma) dd[b] = (dd[b] - (2.0f * gm) * d[b]) - (gm * gm) * c[b];
      d[b] = (float)((double)d[b] + A.cv * (double)dd[b]);
    }
  }
  stv<B>(A.ddchi + (size_t)g * B, dd);
  stv<B>(A.dchi + (size_t)g * B, d);
  if (A.drift) {
#pragma unroll
    for (int b = 0; b < B; b++) c[b] = (float)((double)c[b] + (double)d[b] * A.cd_next);
    if (A.drift_mask && (fl & kFlagAxis)) {
#pragma unroll
      for (int b = 0; b < B; b++) c[b] = 0.0f;
    }
    stv<B>(A.chi + (size_t)g * B, c);
  }
}

template <int B>
__global__ __launch_bounds__(256) void k_receivers(RecArgs A) {
  rec_sample<B>(A, blockIdx.x * 256 + threadIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Kernel-wavefield snapshots, dump_type displ_only (dump_disp_global,
// SOLVER/wavefields_io.f90:1019-1097, with kwf_mapping_sol / kwf_mapping_flu,
// :743-781), blockIdx.y = wavefield b of the group: at the steps with
// mod(iter, strain_it) == 0, snapshot q = iter / strain_it goes to slot
// mod(q, cap) of wavefield b's ring, as [ncomp][npt] (monopole s, z; else s,
// phi, z).  A lane per point (grid stride): solid point k reads the
// displacement of its global point (sol_g, -1 for a point of one of the
// wavefield's source elements, zeroed), fluid point k the chi values of its
// row and column in its element and forms the fluid displacement
//   mxm1 = sum_k GT(i,k) chi(k,j)   (GT = G1T on axial elements, else G2T)
//   mxm2 = sum_k chi(i,k) G2(k,j)   (both summed from zero in k order)
//   s = inv_rho (dzdeta mxm1 + dzdxi mxm2), z = inv_rho (dsdeta mxm1 + dsdxi mxm2)
// as axisym_gradient_fluid (SOLVER/pointwise_derivatives.f90:509-544) and
// the product with inv_rho_fluid (wavefields_io.f90:1087-1089) evaluate it.
// The launch for the initial fields (init_only) writes only the wavefields
// at step 0.
// ---------------------------------------------------------------------------
struct KwfArgs {
  int nps, npt, ncomp, dipole, strain_it, cap, nglob_s, init_only;
  const int *sol_g[kGroup];  // per wavefield: [nps] global solid point, or -1
  const int *flu_g;      // [npf][10]: chi(0..4, j), then chi(i, 0..4)
  const int *flu_ij;     // [npf]: i | j << 3 | axial << 6
  const float *flu_c;    // [5][npf]: inv_rho, DsDeta, DzDeta, DsDxi, DzDxi (over J)
  const float *G;        // G0 | G1 | G1T | G2 | G2T
  const float *disp;     // [3][nglob_s][B]
  const float *chi;      // [nglob_f][B]
  float *ring[kGroup];   // per wavefield: [cap][ncomp][npt]
  StepRef step;
};

template <int B>
__global__ __launch_bounds__(256) void k_kwf(KwfArgs A) {
  const int b = blockIdx.y;
  const int iter = step_of(A.step, b);
  // uniform per block: not a snapshot step of this wavefield
  if (iter % A.strain_it != 0 || (A.init_only && iter != 0)) return;
  float *o = A.ring[b] + (size_t)((iter / A.strain_it) % A.cap) * A.ncomp * A.npt;
  const int npf = A.npt - A.nps;
  const int *sol_g = A.sol_g[b];
  for (int k = blockIdx.x * 256 + threadIdx.x; k < A.npt; k += gridDim.x * 256) {
    float s, p, z;
    if (k < A.nps) {
      const int g = sol_g[k];
      if (g < 0) {
        s = p = z = 0.0f;
      } else {
        const float u1 = A.disp[(size_t)g * B + b],
                    u2 = A.disp[((size_t)A.nglob_s + g) * B + b],
                    u3 = A.disp[(2 * (size_t)A.nglob_s + g) * B + b];
        if (A.dipole) { s = u1 + u2; p = u1 - u2; z = u3; }
        else { s = u1; p = u2; z = u3; }
      }
    } else {
      const int q = k - A.nps;
      const int ij = A.flu_ij[q];
      const int i = ij & 7, j = (ij >> 3) & 7;
      const float *GT = (ij >> 6) ? A.G + 30 : A.G + 80;
      const float *G2 = A.G + 55;
      const int *id = A.flu_g + (size_t)q * 10;
      float m1 = 0.0f, m2 = 0.0f;
#pragma unroll
      for (int kk = 0; kk < 5; kk++) m1 = m1 + GT[i + 5 * kk] * A.chi[(size_t)id[kk] * B + b];
#pragma unroll
      for (int kk = 0; kk < 5; kk++)
        m2 = m2 + A.chi[(size_t)id[5 + kk] * B + b] * G2[kk + 5 * j];
      const float rho = A.flu_c[q], dsdeta = A.flu_c[(size_t)npf + q],
                  dzdeta = A.flu_c[2 * (size_t)npf + q], dsdxi = A.flu_c[3 * (size_t)npf + q],
                  dzdxi = A.flu_c[4 * (size_t)npf + q];
      const float gs = dzdeta * m1 + dzdxi * m2;
      const float gz = dsdeta * m1 + dsdxi * m2;
      s = rho * gs;
      p = 0.0f;
      z = rho * gz;
    }
    o[k] = s;
    if (A.ncomp == 3) {
      o[(size_t)A.npt + k] = p;
      o[2 * (size_t)A.npt + k] = z;
    } else {
      o[(size_t)A.npt + k] = z;
    }
  }
}

// Start of a run: each wavefield's first step and its blow-up flag cleared
// (a kernel, not a memset: also safe where a caller captures the stream).
struct StepSet {
  int *base;
  int first[kGroup];
  int *blow[kGroup];
};
__global__ void k_step_set(StepSet A, int B) {
  const int b = threadIdx.x;
  if (b < B) {
    A.base[b] = A.first[b];
    *A.blow[b] = INT_MAX;
  }
}

// End of a replayed graph of n steps: every wavefield n steps on.
__global__ void k_step_advance(int *base, int B, int n) {
  const int b = threadIdx.x;
  if (b < B) base[b] += n;
}

// moving one wavefield's state between a group's interleaved arrays and its
// own (dst[x*B + b] = src[x], and back)
__global__ __launch_bounds__(256) void k_interleave(float *dst, const float *src, size_t n,
                                                    int B, int b) {
  for (size_t x = (size_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (size_t)gridDim.x * 256)
    dst[x * B + b] = src[x];
}
__global__ __launch_bounds__(256) void k_deinterleave(float *dst, const float *src, size_t n,
                                                      int B, int b) {
  for (size_t x = (size_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (size_t)gridDim.x * 256)
    dst[x] = src[x * B + b];
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

// Host-side layout of the mesh part, computed without touching the GPU:
// global numbering and gather lists, per-point values, flags, packed terms,
// coupling maps.  Every consistency requirement of the per-point storage is
// checked here.
struct MeshHost {
  int nel_s = 0, nel_f = 0, ntile_s = 0, ntile_f = 0, nglob_s = 0, nglob_f = 0;
  int nstages = 1, naxb = 0;
  double coefd[16] = {0}, coefv[16] = {0};
  std::vector<float> G, coef_s, m0_s, minv_s, coef_f, m0_f, minv_f, cpl_bm, cg_stage, cg_upd;
  std::vector<int> einfo_s, gid_s, gstart_s, gslot_s, einfo_f, gid_f, gstart_f, gslot_f,
      cpl_ss, cpl_gs, cpl_sf, cpl_gf, cpl_axb, axb_g, cpl_slot_s, cpl_slot_f;
  std::vector<uint8_t> flags_s, flags_f;
  std::vector<double> aj_mu, aj_k;
  bool zero_term = false;
  // memory variables at all points: packed [tile][kAfTerms][256] terms and
  // the axial vectors [nel][5][5]
  std::vector<float> af_terms, af_axis;
  std::vector<int> first_order_s;  // solid points in first-appearance order
  // absorbing boundaries: damping factor per point (empty when there are none)
  std::vector<float> gamma_s, gamma_f;
  double expw[kMaxSls] = {0}, tst[kMaxSls] = {0}, tstm1[kMaxSls] = {0};
};

// Host-side layout of the wavefield part: source points, fluid source,
// receivers, source time function.
struct WaveHost {
  std::vector<int> src_g, srcel_f, rec_g;
  std::vector<float> src_val, src_fl, stf;
};

// Renumber the solid points so that the points the assembly kernel handles
// (all but kFlagDirect) come first, each class in its previous order: that
// kernel then reads and writes one contiguous range of every per-point array,
// and the stiffness kernel's single-copy points another.
static void number_assembled_first(MeshHost &H) {
  const int n = H.nglob_s;
  std::vector<int> perm(n), inv(n);  // old -> new, new -> old
  int next = 0;
  for (int pass = 0; pass < 2; pass++)
    for (int g = 0; g < n; g++)
      if (((H.flags_s[g] & kFlagDirect) != 0) == (pass == 1)) {
        perm[g] = next;
        inv[next++] = g;
      }
  std::vector<int> gstart(n + 1, 0), gslot(H.gslot_s.size());
  std::vector<float> minv(n);
  std::vector<uint8_t> flags(n);
  for (int p = 0; p < n; p++) {
    const int g = inv[p];
    const int q0 = H.gstart_s[g], q1 = H.gstart_s[g + 1];
    std::copy(H.gslot_s.begin() + q0, H.gslot_s.begin() + q1, gslot.begin() + gstart[p]);
    gstart[p + 1] = gstart[p] + (q1 - q0);
    minv[p] = H.minv_s[g];
    flags[p] = H.flags_s[g];
  }
  if (!H.gamma_s.empty()) {
    std::vector<float> gm(n);
    for (int p = 0; p < n; p++) gm[p] = H.gamma_s[inv[p]];
    H.gamma_s.swap(gm);
  }
  H.gstart_s.swap(gstart);
  H.gslot_s.swap(gslot);
  H.minv_s.swap(minv);
  H.flags_s.swap(flags);
  for (auto &x : H.gid_s)
    if (x >= 0) x = perm[x];
  for (auto &x : H.cpl_gs) x = perm[x];
  for (auto &x : H.axb_g) x = perm[x];
  H.first_order_s.swap(perm);  // perm[old] lists the new ids in the old order
}

static void build_mesh_layout(const axh_ctx *ctx, MeshHost &H) {
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
  const bool absorbing = ctx->get("solid_absorbing_gamma") != nullptr;
  if (absorbing)
    H.gamma_s = per_point(NS, ctx->need("solid_absorbing_gamma").as<float>(), H.nel_s,
                          "solid_absorbing_gamma");
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
  if (zero_term_of(src) >= 0 && H.nel_s > 0) {
    const uint32_t *v = ctx->need(names[zero_term_of(src)]).as<uint32_t>();
    bool zero = true;
    for (size_t x = 0; x < (size_t)H.nel_s * kNpp && zero; x++) zero = v[x] == 0u;
    H.zero_term = zero;
  }

  // axial terms M0_w1..10 of the axial elements, the ones the source type
  // reads (monopole w1-3, dipole w1-4 and w6-10, quadrupole w1-6); the rest
  // stays zero
  auto m0_used = [&](int w) {
    if (src == AXH_MONOPOLE) return w <= 3;
    if (src == AXH_DIPOLE) return w != 5;
    return w <= 6;
  };
  H.einfo_s.assign(H.nel_s, -1);
  int nax = 0;
  for (int e = 0; e < H.nel_s; e++)
    if (axis_s[e]) {
      H.einfo_s[e] = nax++;
      for (int w = 1; w <= 10; w++) {
        const Array *a = m0_used(w) ? ctx->get("M0_w" + std::to_string(w)) : nullptr;
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
    if (absorbing)
      H.gamma_f = per_point(NF, ctx->need("fluid_absorbing_gamma").as<float>(), H.nel_f,
                            "fluid_absorbing_gamma");
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
        for (int j = 0; j < 5; j++)
          H.m0_f.push_back(m0a ? m0a->as<float>()[(size_t)e * 5 + j] : 0.0f);
      }
    if (multipole && naxf > 0 && !m0a) fail(AXH_ERR_STATE, "missing upload: M0_w_fl");
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
    H.cpl_slot_s.assign((size_t)H.ntile_s * kTile, -1);
    H.cpl_slot_f.assign((size_t)H.ntile_f * kTile, -1);
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
        // the assembly adds the coupling term to these slots before summing
        H.cpl_slot_s[s1] = H.cpl_slot_f[s2] = (int)H.cpl_ss.size();
        H.flags_s[g] |= kFlagCoupled;
        H.flags_f[H.gid_f[s2]] |= kFlagCoupled;
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

  // ---- points with a single element copy and nothing else to add (no axis
  // mask, no coupling term; source points are excluded when a wavefield with
  // a source there is attached) ----
  // (none with memory variables at all points: the stiffness kernel would
  // kick them before the anelastic forces are subtracted)
  const bool anel_full = d.anel && !d.coarse_grained;
  for (int g = 0; g < H.nglob_s; g++)
    if (!absorbing && !anel_full && H.gstart_s[g + 1] - H.gstart_s[g] == 1 &&
        !(H.flags_s[g] & (kFlagAxis | kFlagCoupled)))
      H.flags_s[g] |= kFlagDirect;
  number_assembled_first(H);

  // ---- attenuation ----
  if (d.anel && d.coarse_grained) {
    const char *stage_names[kCgStage] = {"Y_cg4", "V_s_eta_cg4", "V_s_xi_cg4", "V_z_eta_cg4",
                                         "V_z_xi_cg4"};
    const char *upd_names[kCgUpd - 1] = {"DsDeta_over_J_sol_cg4", "DzDeta_over_J_sol_cg4",
                                         "DsDxi_over_J_sol_cg4", "DzDxi_over_J_sol_cg4",
                                         "inv_s_solid_cg4", "delta_mu_cg4", "delta_kappa_cg4"};
    H.cg_stage.assign((size_t)H.nel_s * kCgStage * 4, 0.0f);
    H.cg_upd.assign((size_t)H.nel_s * kCgUpd * 4, 0.0f);
    for (int a = 0; a < kCgStage; a++) {
      const float *p = ctx->need(stage_names[a]).as<float>();
      for (int e = 0; e < H.nel_s; e++)
        for (int q = 0; q < 4; q++)
          H.cg_stage[((size_t)e * kCgStage + a) * 4 + q] = p[(size_t)e * 4 + q];
    }
    for (int a = 0; a < kCgUpd - 1; a++) {
      const float *p = ctx->need(upd_names[a]).as<float>();
      for (int e = 0; e < H.nel_s; e++)
        for (int q = 0; q < 4; q++)
          H.cg_upd[((size_t)e * kCgUpd + a) * 4 + q] = p[(size_t)e * 4 + q];
    }
  }
  if (anel_full) {
    pack_terms(ctx, {"Y", "V_s_eta", "V_s_xi", "V_z_eta", "V_z_xi", "DsDeta_over_J_sol",
                     "DzDeta_over_J_sol", "DsDxi_over_J_sol", "DzDxi_over_J_sol", "inv_s_solid",
                     "delta_mu", "delta_kappa"},
               H.nel_s, H.af_terms);
    const char *axis_names[5] = {"Y0", "V0_s_eta", "V0_s_xi", "V0_z_eta", "V0_z_xi"};
    H.af_axis.assign((size_t)H.nel_s * 25, 0.0f);
    for (int a = 0; a < 5; a++) {
      const float *p = ctx->need(axis_names[a]).as<float>();
      for (int e = 0; e < H.nel_s; e++)
        std::memcpy(&H.af_axis[(size_t)e * 25 + 5 * a], p + (size_t)e * 5, 5 * sizeof(float));
    }
  }
  if (d.anel) {
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
}

static void build_wave_layout(const axh_ctx *ctx, const std::vector<int> &gid_s, int nglob_s,
                              WaveHost &W) {
  const axh_dims &d = ctx->dims;
  const int nel_s = d.nel_solid, nel_f = d.nel_fluid;
  // source terms (solid), added after assembly to every copy in a source
  // element (SOLVER/time_evol_wave.F90:1082-1097): copies must agree
  if (!d.fluid_src && d.nelsrc > 0) {
    const int32_t *iels = ctx->need("ielsrc").as<int32_t>();
    const float *st = ctx->need("source_term").as<float>();
    std::vector<int> srcidx(nel_s, -1);
    for (int kk = 0; kk < d.nelsrc; kk++) {
      if (iels[kk] < 1 || iels[kk] > nel_s) fail(AXH_ERR_ARG, "ielsrc out of range");
      srcidx[iels[kk] - 1] = kk;
    }
    std::vector<int> at(nglob_s, -1);  // global point -> index in pts
    std::vector<int> pts;
    std::vector<float> val;
    auto nonzero = [&](int q) {
      return val[3 * q] != 0.0f || val[3 * q + 1] != 0.0f || val[3 * q + 2] != 0.0f;
    };
    for (int e = 0; e < nel_s; e++) {
      if (srcidx[e] < 0) continue;
      for (int k = 0; k < kNpp; k++) {
        const int g = gid_s[slot_of(e, k)];
        float x[3];
        for (int c = 0; c < 3; c++) x[c] = st[k + kNpp * ((size_t)srcidx[e] + 8 * c)];
        if (at[g] < 0) {
          at[g] = (int)pts.size();
          pts.push_back(g);
          for (int c = 0; c < 3; c++) val.push_back(x[c]);
        } else {
          for (int c = 0; c < 3; c++)
            if (val[3 * at[g] + c] != x[c])
              fail(AXH_ERR_UNSUPPORTED, "source term differs between copies of a point");
        }
      }
    }
    // a copy in an element without source holds zero
    for (int e = 0; e < nel_s; e++) {
      if (srcidx[e] >= 0) continue;
      for (int k = 0; k < kNpp; k++) {
        const int q = at[gid_s[slot_of(e, k)]];
        if (q >= 0 && nonzero(q))
          fail(AXH_ERR_UNSUPPORTED, "source term differs between copies of a point");
      }
    }
    // the points with a nonzero term, in ascending global index
    for (int g = 0; g < nglob_s; g++) {
      const int q = at[g];
      if (q >= 0 && nonzero(q)) {
        W.src_g.push_back(g);
        for (int c = 0; c < 3; c++) W.src_val.push_back(val[3 * q + c]);
      }
    }
  }
  if (nel_f > 0) {
    W.srcel_f.assign(nel_f, -1);
    W.src_fl.assign(25 * 8, 0.0f);
    if (d.fluid_src && d.nelsrc > 0) {
      const int32_t *iels = ctx->need("ielsrc").as<int32_t>();
      for (int kk = 0; kk < d.nelsrc; kk++) {
        if (iels[kk] < 1 || iels[kk] > nel_f) fail(AXH_ERR_ARG, "ielsrc out of range");
        W.srcel_f[iels[kk] - 1] = kk;
      }
      std::memcpy(W.src_fl.data(), ctx->need("source_term").as<float>(), 25 * 8 * 4);
    }
  }
  // ---- receivers ----
  if (d.num_rec > 0) {
    const int32_t *rf = ctx->need("recfile_el").as<int32_t>();
    W.rec_g.resize(d.num_rec);
    for (int r = 0; r < d.num_rec; r++) {
      int e = rf[r] - 1, i = rf[r + d.num_rec], j = rf[r + 2 * d.num_rec];
      if (e < 0 || e >= nel_s || i < 0 || i > 4 || j < 0 || j > 4)
        fail(AXH_ERR_ARG, "receiver index out of range");
      W.rec_g[r] = gid_s[slot_of(e, i + 5 * j)];
    }
  }
  const Array &stf = ctx->need("stf");
  W.stf.assign(stf.as<float>(), stf.as<float>() + stf.count());
}

void check_inputs(const axh_ctx *ctx) {
  MeshHost H;
  build_mesh_layout(ctx, H);
  WaveHost W;
  build_wave_layout(ctx, H.gid_s, H.nglob_s, W);
}

// 128-bit fingerprint of the mesh part the device would hold for ctx (every
// array of MeshHost and its sizes).  Two contexts with equal fingerprints
// step on the same device mesh, whatever their unused upload entries hold.
struct FpHash {
  uint64_t a = 0x9e3779b97f4a7c15ull, b = 0xc2b2ae3d27d4eb4full;
  void word(uint64_t w) {
    a = (a ^ w) * 0xff51afd7ed558ccdull;
    a ^= a >> 32;
    b = (b + w) * 0xc4ceb9fe1a85ec53ull;
    b ^= b >> 29;
  }
  void bytes(const void *p, size_t n) {
    const unsigned char *c = (const unsigned char *)p;
    word(n);
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
      uint64_t w;
      std::memcpy(&w, c + i, 8);
      word(w);
    }
    uint64_t t = 0;
    std::memcpy(&t, c + i, n - i);
    word(t);
  }
  template <class T> void vec(const std::vector<T> &v) { bytes(v.data(), v.size() * sizeof(T)); }
};

void mesh_fingerprint(axh_ctx *ctx, uint64_t fp[2]) {
  if (ctx->fp_valid) {
    fp[0] = ctx->fp[0];
    fp[1] = ctx->fp[1];
    return;
  }
  MeshHost H;
  build_mesh_layout(ctx, H);
  FpHash h;
  const int ints[] = {ctx->dims.src_type, ctx->dims.scheme, ctx->dims.anel, ctx->dims.n_sls,
                      ctx->dims.nel_bdry, H.nel_s, H.nel_f, H.nglob_s, H.nglob_f,
                      H.nstages, H.naxb};
  h.bytes(ints, sizeof ints);
  h.bytes(&ctx->dims.deltat, sizeof(double));
  h.bytes(H.coefd, sizeof H.coefd);
  h.bytes(H.coefv, sizeof H.coefv);
  h.bytes(H.expw, sizeof H.expw);
  h.bytes(H.tst, sizeof H.tst);
  h.bytes(H.tstm1, sizeof H.tstm1);
  for (const auto *v : {&H.G, &H.coef_s, &H.m0_s, &H.minv_s, &H.coef_f, &H.m0_f, &H.minv_f,
                        &H.cpl_bm, &H.cg_stage, &H.cg_upd, &H.gamma_s, &H.gamma_f,
                        &H.af_terms, &H.af_axis})
    h.vec(*v);
  for (const auto *v : {&H.einfo_s, &H.gid_s, &H.gstart_s, &H.gslot_s, &H.einfo_f, &H.gid_f,
                        &H.gstart_f, &H.gslot_f, &H.cpl_ss, &H.cpl_gs, &H.cpl_sf, &H.cpl_gf,
                        &H.cpl_axb, &H.axb_g, &H.cpl_slot_s, &H.cpl_slot_f})
    h.vec(*v);
  h.vec(H.flags_s);
  h.vec(H.flags_f);
  h.vec(H.aj_mu);
  h.vec(H.aj_k);
  fp[0] = ctx->fp[0] = h.a;
  fp[1] = ctx->fp[1] = h.b;
  ctx->fp_valid = true;
}

// Key of the batches a context may join: the dims every member shares and the
// mesh maps (numbering and axis flags), not the source type or its terms.
void batch_key(const axh_ctx *ctx, uint64_t key[2]) {
  const axh_dims &d = ctx->dims;
  FpHash h;
  const int ints[] = {d.abi_version, d.npol, d.nel_solid, d.nel_fluid, d.nglob_solid,
                      d.nglob_fluid, d.nel_bdry, d.scheme, d.anel, d.coarse_grained, d.n_sls};
  h.bytes(ints, sizeof ints);
  h.bytes(&d.deltat, sizeof(double));
  for (const char *nm : {"igloc_solid", "igloc_fluid", "axis_solid", "axis_fluid"}) {
    const Array *a = ctx->get(nm);
    if (!a) fail(AXH_ERR_STATE, std::string("missing upload: ") + nm);
    h.vec(a->bytes);
  }
  key[0] = h.a;
  key[1] = h.b;
}

void global_numbering(const axh_ctx *ctx, int domain, int32_t *gid, int32_t *nglob) {
  const bool solid = domain == 0;
  MeshHost H;
  build_mesh_layout(ctx, H);
  const int nel = solid ? H.nel_s : H.nel_f;
  const std::vector<int> &g = solid ? H.gid_s : H.gid_f;
  for (int e = 0; e < nel; e++)
    for (int k = 0; k < kNpp; k++) gid[(size_t)e * kNpp + k] = g[slot_of(e, k)];
  *nglob = solid ? H.nglob_s : H.nglob_f;
}

static std::shared_ptr<MeshDev> mesh_build(const axh_ctx *ctx) {
  const axh_dims &d = ctx->dims;
  MeshHost H;
  build_mesh_layout(ctx, H);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    fail(AXH_ERR_HIP, "no HIP device available");
  auto M = std::make_shared<MeshDev>();
  MeshDev &V = *M;
  V.src = d.src_type;
  V.scheme = d.scheme;
  V.anel = d.anel && d.coarse_grained;
  V.anel_full = d.anel && !d.coarse_grained;
  V.nsls = d.n_sls;
  V.nel_s = H.nel_s;
  V.nel_f = H.nel_f;
  V.ntile_s = H.ntile_s;
  V.ntile_f = H.ntile_f;
  V.nbdry = d.nel_bdry;
  V.nstages = H.nstages;
  V.nglob_s = H.nglob_s;
  V.nglob_f = H.nglob_f;
  V.naxb = H.naxb;
  std::memcpy(V.coefd, H.coefd, sizeof V.coefd);
  std::memcpy(V.coefv, H.coefv, sizeof V.coefv);
  std::memcpy(V.expw, H.expw, sizeof V.expw);
  std::memcpy(V.tst, H.tst, sizeof V.tst);
  std::memcpy(V.tstm1, H.tstm1, sizeof V.tstm1);
  DevAlloc &m = V.mem;
  V.G = m.upload(H.G);
  V.coef_s = m.upload(H.coef_s);
  V.zero_term = H.zero_term;
  V.einfo_s = m.upload(H.einfo_s);
  V.m0_s = m.upload(H.m0_s);
  V.gid_s = m.upload(H.gid_s);
  V.gstart_s = m.upload(H.gstart_s);
  V.gslot_s = m.upload(H.gslot_s);
  V.minv_s = m.upload(H.minv_s);
  if (!H.gamma_s.empty()) V.gamma_s = m.upload(H.gamma_s);
  if (V.nel_f > 0) {
    V.coef_f = m.upload(H.coef_f);
    V.einfo_f = m.upload(H.einfo_f);
    V.m0_f = m.upload(H.m0_f);
    V.gid_f = m.upload(H.gid_f);
    V.gstart_f = m.upload(H.gstart_f);
    V.gslot_f = m.upload(H.gslot_f);
    V.minv_f = m.upload(H.minv_f);
    if (!H.gamma_f.empty()) V.gamma_f = m.upload(H.gamma_f);
    V.flags_f = m.upload(H.flags_f);
  }
  if (V.nbdry > 0) {
    V.cpl_ss = m.upload(H.cpl_ss);
    V.cpl_gs = m.upload(H.cpl_gs);
    V.cpl_sf = m.upload(H.cpl_sf);
    V.cpl_gf = m.upload(H.cpl_gf);
    V.cpl_axb = m.upload(H.cpl_axb);
    V.cpl_bm = m.upload(H.cpl_bm);
    V.cpl_slot_s = m.upload(H.cpl_slot_s);
    V.cpl_slot_f = m.upload(H.cpl_slot_f);
    V.axb_g = m.upload(H.axb_g);
  }
  if (V.anel) {
    V.cg_stage = m.upload(H.cg_stage);
    V.cg_upd = m.upload(H.cg_upd);
  }
  if (V.anel_full) {
    V.af_terms = m.upload(H.af_terms);
    V.af_axis = m.upload(H.af_axis);
  }
  if (V.anel || V.anel_full) {
    V.aj_mu = m.upload(H.aj_mu);
    V.aj_k = m.upload(H.aj_k);
  }
  V.gid_s_h = std::move(H.gid_s);
  V.first_order_s_h = std::move(H.first_order_s);
  V.gid_f_h = std::move(H.gid_f);
  V.flags_s_h = std::move(H.flags_s);
  V.first_direct = V.nglob_s;
  for (int g = 0; g < V.nglob_s; g++)
    if (V.flags_s_h[g] & kFlagDirect) { V.first_direct = g; break; }
  return M;
}

// ---------------------------------------------------------------------------
// wavefield storage and sources
// ---------------------------------------------------------------------------
// Source, receivers and blow-up flag of one wavefield (not interleaved).
struct Member {
  DevAlloc mem;
  int nsrc = 0;               // solid source points
  std::vector<int> src_g_h;
  int *src_g = nullptr;
  float *src_val = nullptr;   // [nsrc][3]
  int fluid_src = 0;
  int *srcel_f = nullptr;     // per fluid element: source slot k or -1
  float *src_fl = nullptr;    // source_term_fl (25 x 8)
  int nrec = 0, seis_it = 1;
  int *rec_g = nullptr;
  // receiver samples of a run, staged on the device: room for seis_cap
  // samples, sized when the member is built (a longer run copies them out
  // whenever the buffer is full), so a run allocates nothing
  float *seis_d = nullptr;
  int seis_cap = 0;
  int *blow = nullptr;
  double blow_thr = 0.0;
  float *stf_d = nullptr;     // source time function [niter][nstages]
};

// State of B wavefields, interleaved (element x of wavefield b at x*B + b),
// and their element-force scratch.
struct Fields {
  int B = 1;
  DevAlloc mem;
  float *disp = nullptr, *velo = nullptr, *acc0 = nullptr, *force_s = nullptr;
  float *chi = nullptr, *dchi = nullptr, *ddchi = nullptr, *force_f = nullptr;
  float *memvar = nullptr, *rsum = nullptr, *trm1 = nullptr, *devm1 = nullptr;
  float *axb_val = nullptr;   // unmasked prediction of the axis boundary points
  // solid point flags with the source points of the group's wavefields
  // (kFlagSource set, kFlagDirect cleared) and the points the assembly
  // kernel handles (all but kFlagDirect, ascending)
  uint8_t *flags_s = nullptr;
  int nasm = 0;
  int *asm_pts = nullptr;
  int *step_base = nullptr;   // [kGroup] step of each wavefield (StepRef)
  // solid tiles holding a source point of the group (at most kMaxSrcTiles)
  std::vector<int> src_tiles;
};

// the state arrays of a wavefield (pointer into F, entries per wavefield)
static std::vector<std::pair<float *, size_t>> state_arrays(const MeshDev &M, Fields &F) {
  std::vector<std::pair<float *, size_t>> v = {
      {F.disp, 3 * (size_t)M.nglob_s}, {F.velo, 3 * (size_t)M.nglob_s},
      {F.acc0, 3 * (size_t)M.nglob_s}, {F.axb_val, 3 * (size_t)std::max(M.naxb, 1)}};
  if (M.nel_f > 0) {
    v.push_back({F.chi, (size_t)M.nglob_f});
    v.push_back({F.dchi, (size_t)M.nglob_f});
    v.push_back({F.ddchi, (size_t)M.nglob_f});
  }
  if (M.anel) {
    v.push_back({F.memvar, (size_t)M.nel_s * 24 * M.nsls});
    v.push_back({F.rsum, (size_t)M.nel_s * 24});
    v.push_back({F.trm1, (size_t)M.nel_s * 4});
    v.push_back({F.devm1, (size_t)M.nel_s * 24});
  }
  if (M.anel_full) {
    const size_t ns = (size_t)M.ntile_s * kTile;
    v.push_back({F.memvar, ns * 6 * M.nsls});
    v.push_back({F.rsum, ns * 6});
    v.push_back({F.trm1, ns});
    v.push_back({F.devm1, ns * 6});
  }
  return v;
}

// Point flags and assembly list of a group whose wavefields have their solid
// source terms at the points srcs[...] (the mesh part stays unchanged).  Only
// a single wavefield kicks its single-copy points in the stiffness kernel
// (for wider groups that costs the stiffness kernel more than it saves the
// assembly): groups of B > 1 assemble every point.
static void fields_set_sources(const MeshDev &M, Fields &F,
                               const std::vector<const std::vector<int> *> &srcs) {
  std::vector<uint8_t> fl = M.flags_s_h;
  if (F.B > 1)
    for (auto &f : fl) f &= (uint8_t)~kFlagDirect;
  {
    std::vector<char> is_src(M.nglob_s, 0);
    for (const auto *v : srcs)
      for (int g : *v) is_src[g] = 1;
    F.src_tiles.clear();
    for (size_t sl = 0; sl < M.gid_s_h.size(); sl++) {
      const int g = M.gid_s_h[sl];
      const int tl = (int)(sl / kTile);
      if (g >= 0 && is_src[g] && (F.src_tiles.empty() || F.src_tiles.back() != tl))
        F.src_tiles.push_back(tl);
    }
  }
  for (const auto *v : srcs)
    for (int g : *v) {
      fl[g] |= kFlagSource;
      fl[g] &= (uint8_t)~kFlagDirect;
    }
  std::vector<int> pts;
  pts.reserve(M.nglob_s);
  if (F.B > 1) {
    // every point, in first-appearance (element) order: neighbouring lanes
    // then gather neighbouring slots of the force array
    pts = M.first_order_s_h;
  } else {
    for (int g = 0; g < M.nglob_s; g++)
      if (!(fl[g] & kFlagDirect)) pts.push_back(g);
  }
  F.nasm = (int)pts.size();
  HIPCHECK(hipMemcpy(F.flags_s, fl.data(), fl.size(), hipMemcpyHostToDevice));
  if (!pts.empty())
    HIPCHECK(hipMemcpy(F.asm_pts, pts.data(), pts.size() * sizeof(int), hipMemcpyHostToDevice));
}

static void fields_alloc(const MeshDev &M, Fields &F, int B) {
  F.B = B;
  DevAlloc &m = F.mem;
  F.flags_s = m.alloc<uint8_t>(M.nglob_s);
  F.asm_pts = m.alloc<int>(M.nglob_s);
  F.step_base = m.zeros<int>(kGroup);
  const size_t b = B;
  F.disp = m.zeros<float>(3 * (size_t)M.nglob_s * b);
  F.velo = m.zeros<float>(3 * (size_t)M.nglob_s * b);
  F.acc0 = m.zeros<float>(3 * (size_t)M.nglob_s * b);
  F.force_s = m.zeros<float>(3 * (size_t)M.ntile_s * kTile * b);
  F.axb_val = m.zeros<float>(3 * (size_t)std::max(M.naxb, 1) * b);
  if (M.nel_f > 0) {
    F.chi = m.zeros<float>((size_t)M.nglob_f * b);
    F.dchi = m.zeros<float>((size_t)M.nglob_f * b);
    F.ddchi = m.zeros<float>((size_t)M.nglob_f * b);
    F.force_f = m.zeros<float>((size_t)M.ntile_f * kTile * b);
  }
  if (M.anel) {
    F.memvar = m.zeros<float>((size_t)M.nel_s * 24 * M.nsls * b);
    F.rsum = m.zeros<float>((size_t)M.nel_s * 24 * b);
    F.trm1 = m.zeros<float>((size_t)M.nel_s * 4 * b);
    F.devm1 = m.zeros<float>((size_t)M.nel_s * 24 * b);
  }
  if (M.anel_full) {
    const size_t ns = (size_t)M.ntile_s * kTile;
    F.memvar = m.zeros<float>(ns * 6 * M.nsls * b);
    F.rsum = m.zeros<float>(ns * 6 * b);
    F.trm1 = m.zeros<float>(ns * b);
    F.devm1 = m.zeros<float>(ns * 6 * b);
  }
}

// Copy wavefield src_b of S into wavefield dst_b of D (same mesh), through
// the (de)interleaving kernels on stream s.
static void copy_state(const MeshDev &M, Fields &D, int dst_b, Fields &S, int src_b,
                       hipStream_t s) {
  auto dv = state_arrays(M, D), sv = state_arrays(M, S);
  for (size_t a = 0; a < dv.size(); a++) {
    const size_t n = dv[a].second;
    const int nb = (int)std::min<size_t>((n + 255) / 256, 65536);
    if (S.B == 1) {
      hipLaunchKernelGGL(k_interleave, dim3(nb), dim3(256), 0, s, dv[a].first, sv[a].first, n,
                         D.B, dst_b);
    } else if (D.B == 1) {
      hipLaunchKernelGGL(k_deinterleave, dim3(nb), dim3(256), 0, s, dv[a].first, sv[a].first, n,
                         S.B, src_b);
    } else {
      fail(AXH_ERR_STATE, "state copy between two groups");
    }
  }
  HIPCHECK(hipGetLastError());
}

// Receiver samples a member stages on the device: a whole run's up to
// kSeisStage (4.3 MB at 181 receivers).
constexpr int kSeisStage = 2048;
static int seis_staging_samples(const axh_dims &d) {
  const int total = d.seis_it > 0 ? d.niter / d.seis_it : d.niter;
  return std::max(1, std::min(total, kSeisStage));
}

static void member_build(const axh_ctx *ctx, const MeshDev &M, Member &W) {
  const axh_dims &d = ctx->dims;
  WaveHost H;
  build_wave_layout(ctx, M.gid_s_h, M.nglob_s, H);
  DevAlloc &m = W.mem;
  W.nsrc = (int)H.src_g.size();
  W.src_g = m.upload(H.src_g);
  W.src_val = m.upload(H.src_val);
  W.fluid_src = d.fluid_src;
  if (M.nel_f > 0) {
    W.srcel_f = m.upload(H.srcel_f);
    W.src_fl = m.upload(H.src_fl);
  }
  W.nrec = d.num_rec;
  W.seis_it = d.seis_it;
  if (W.nrec > 0) {
    W.rec_g = m.upload(H.rec_g);
    W.seis_cap = seis_staging_samples(d);
    W.seis_d = m.alloc<float>((size_t)W.seis_cap * 3 * W.nrec);
  }
  W.blow = m.alloc<int>(1);
  // blow-up threshold of runtime_info: maxval(abs(disp(1,1,:,:))) >
  // 10*abs(magnitude), the real(4) maximum promoted to double
  W.blow_thr = 10.0 * std::fabs(d.magnitude);
  W.stf_d = m.upload(H.stf);
  W.src_g_h = std::move(H.src_g);
}

// Stream, events and per-kernel timing of the runs of a context or a batch.
struct Runner {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool profile = false;
  std::vector<hipEvent_t> kev;  // pairs
  std::vector<int> kev_class;
  double kms[KC_N] = {0};
  long kcount[KC_N] = {0};
  // replayed graph of graph_steps steps of one group (0: launch directly)
  int graph_steps = 0;
  hipGraphExec_t gexec = nullptr;
  const void *gfields = nullptr;
  const int *gbase = nullptr;
  int gsteps = 0;
  uint64_t gkwf = 0;   // kernel-wavefield gather captured in the graph (0: none)
  void init() {
    if (stream) return;
    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreate(&ev0));
    HIPCHECK(hipEventCreate(&ev1));
  }
  ~Runner() {
    // teardown: errors are not actionable here
    if (stream) (void)hipStreamSynchronize(stream);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    for (auto e : kev) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// Device buffers a run writes besides the fields: the step base of the
// group, each wavefield's blow-up flag and its receiver-sample ring of cap
// samples.  A context's own (run_buffers) or a fixed scratch area.
struct RunBuffers {
  int *step_base = nullptr;
  int *blow[kGroup] = {};
  float *ring[kGroup] = {};
  int cap[kGroup] = {};
};

static RunBuffers run_buffers(Fields &F, Member *const *mb, int B) {
  RunBuffers r;
  r.step_base = F.step_base;
  for (int b = 0; b < B; b++) {
    r.blow[b] = mb[b]->blow;
    r.ring[b] = mb[b]->seis_d;
    r.cap[b] = std::max(mb[b]->seis_cap, 1);
  }
  return r;
}

// The receiver samples of a group's run, per wavefield: sample q (the step
// q * seis_it) is staged in slot mod(q - 1, cap) of the wavefield's ring;
// q0 is the last sample before the run, qdone the last one whose step is
// queued, qout the last one queued for copying out.
struct GroupTail {
  int B = 0;
  Member *mb[kGroup] = {};
  float *seis[kGroup] = {};
  RunBuffers rb;
  int q0[kGroup] = {}, qdone[kGroup] = {}, qout[kGroup] = {};
  // queue the copy of samples qout+1 .. qdone into seis (at most two pieces)
  void copy_out(int b, hipStream_t s) {
    const Member &w = *mb[b];
    const int n = qdone[b] - qout[b];
    if (n <= 0) return;
    if (seis[b]) {
      const size_t row = 3 * (size_t)w.nrec;
      const int cap = rb.cap[b];
      const int slot = qout[b] % cap;
      const int first = std::min(n, cap - slot);
      float *dst = seis[b] + (size_t)(qout[b] - q0[b]) * row;
      HIPCHECK(hipMemcpyAsync(dst, rb.ring[b] + slot * row, first * row * sizeof(float),
                              hipMemcpyDeviceToHost, s));
      if (n > first)
        HIPCHECK(hipMemcpyAsync(dst + first * row, rb.ring[b], (n - first) * row * sizeof(float),
                                hipMemcpyDeviceToHost, s));
    }
    qout[b] = qdone[b];
  }
  // steps it+1 .. it+n of wavefield b are about to be queued: their samples
  // must fit the ring beside those not yet copied out (stream order keeps a
  // copy ahead of the kernels that refill its slots)
  void reserve(int b, int it, int n, hipStream_t s) {
    const Member &w = *mb[b];
    if (w.nrec <= 0) return;
    const int q1 = (it + n) / std::max(w.seis_it, 1);
    if (q1 - qout[b] > rb.cap[b]) copy_out(b, s);
    qdone[b] = q1;
  }
};

// Copy-out of the kernel-wavefield snapshots of a run.  Snapshot q sits in
// ring slot q mod cap until its copy is queued; copies are queued in the run's
// stream order (after the steps that wrote them, before those that reuse the
// slots) into one of two pinned staging halves of H snapshots, from where the
// host moves them to the caller's buffer while the GPU steps on.  A half is
// reused once its previous copy has completed and been moved.
struct KwfTail {
  int s = 1, cap = 2, H = 1;
  size_t snap = 0;              // floats per snapshot
  const float *ring = nullptr;
  float *stage[2] = {};
  hipEvent_t ev[2] = {};
  int pend_q[2] = {}, pend_n[2] = {};
  int half = 0;
  float *out = nullptr;         // the caller's buffer, from snapshot first_q on
  int first_q = 0, qout = -1, qdone = -1;  // copy queued up to qout; steps up to qdone

  void drain(int h) {
    if (pend_n[h] == 0) return;
    HIPCHECK(hipEventSynchronize(ev[h]));
    std::memcpy(out + (size_t)(pend_q[h] - first_q) * snap, stage[h],
                (size_t)pend_n[h] * snap * sizeof(float));
    pend_n[h] = 0;
  }
  // queue the copy of snapshots qout+1 .. qout+n (n <= H) into the next half
  void flush(int n, hipStream_t st) {
    const int h = half;
    half ^= 1;
    drain(h);
    const int q0 = qout + 1, slot = q0 % cap, n1 = std::min(n, cap - slot);
    HIPCHECK(hipMemcpyAsync(stage[h], ring + (size_t)slot * snap, (size_t)n1 * snap * sizeof(float),
                            hipMemcpyDeviceToHost, st));
    if (n > n1)
      HIPCHECK(hipMemcpyAsync(stage[h] + (size_t)n1 * snap, ring,
                              (size_t)(n - n1) * snap * sizeof(float), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipEventRecord(ev[h], st));
    pend_q[h] = q0;
    pend_n[h] = n;
    qout += n;
  }
  // steps it+1 .. it+n are about to be queued: copy out every H snapshots,
  // and whatever the launch would otherwise overwrite
  void reserve(int it, int n, hipStream_t st) {
    const int q1 = (it + n) / s;
    while (qdone - qout >= H) flush(H, st);
    while (q1 - qout > cap && qout < qdone) flush(std::min(H, qdone - qout), st);
    if (q1 - qout > cap) fail(AXH_ERR_STATE, "kernel-wavefield ring too small for one launch");
    qdone = q1;
  }
  // the copies of the rest, queued; complete() moves every pending half
  void finish(hipStream_t st) {
    while (qout < qdone) flush(std::min(H, qdone - qout), st);
  }
  void complete() {
    drain(half);
    drain(half ^ 1);
  }
};

// The kernel-wavefield gather of a context (axh_run_kwf): the point tables of
// k_kwf and its snapshot ring, for one grid, strain_it and source mask.
struct KwfDev {
  DevAlloc mem;
  int nps = 0, npf = 0, ncomp = 2, strain_it = 0, src_mask = 0, ring_req = 0;
  int cap = 2, half = 1;   // ring slots; snapshots per pinned staging half
  uint64_t gen = 0;  // distinct for every build (keys the captured graph)
  int *sol_g = nullptr, *flu_g = nullptr, *flu_ij = nullptr;
  float *flu_c = nullptr, *ring = nullptr;
  float *stage = nullptr;  // pinned host memory, 2 * half snapshots
  hipEvent_t ev[2] = {};
  KwfDev() = default;
  KwfDev(const KwfDev &) = delete;
  KwfDev &operator=(const KwfDev &) = delete;
  ~KwfDev() {
    // teardown: errors are not actionable here
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stage) (void)hipHostFree(stage);
  }
};

// One context on the device: the (shared) mesh part, its source data, its own
// fields (B = 1) and, while a batch steps it, where its state lives instead.
struct Device {
  std::shared_ptr<MeshDev> mesh;
  Member member;
  Fields own;
  Fields *live = nullptr;  // a batch group holding the state, or NULL
  int live_slot = 0;
  GroupTail pending;       // a run queued by axh_run_begin, until axh_run_end
  int32_t pending_it1 = 0;
  KwfTail pending_kwf;     // its kernel-wavefield copies (axh_run_kwf_begin)
  bool pending_has_kwf = false;
  std::unique_ptr<KwfDev> kwf;  // built by the first axh_run_kwf
  Runner run;  // declared last: its stream is drained before buffers are freed
};

// Groups of a batch: interleaved fields of up to kGroup members each.
struct BatchState {
  std::vector<std::unique_ptr<Fields>> groups;
  std::vector<std::vector<int>> index;          // members of each group
  std::vector<std::shared_ptr<MeshDev>> mesh;   // each group's mesh part
  Runner run;                                   // attach / detach copies
  std::vector<std::unique_ptr<Runner>> runs;    // each group's stream
};

void device_build(axh_ctx *ctx) {
  std::unique_ptr<Device> D(new Device);
  D->mesh = ctx->bound_mesh ? ctx->bound_mesh : mesh_build(ctx);
  member_build(ctx, *D->mesh, D->member);
  fields_alloc(*D->mesh, D->own, 1);
  fields_set_sources(*D->mesh, D->own, {&D->member.src_g_h});
  HIPCHECK(hipDeviceSynchronize());
  ctx->dev.reset(D.release());
}

void DeviceDeleter::operator()(Device *d) const { device_destroy(d); }
void BatchStateDeleter::operator()(BatchState *b) const { delete b; }

void device_destroy(Device *d) { delete d; }

size_t device_bytes(const Device *d) {
  return d ? d->mesh->mem.bytes + d->member.mem.bytes + d->own.mem.bytes +
                 (d->kwf ? d->kwf->mem.bytes : 0)
           : 0;
}

const char *kernel_class_name(int k) { return (k >= 0 && k < KC_N) ? kClassName[k] : nullptr; }

static bool runner_stats(const Runner *r, int k, double *ms, int64_t *n) {
  if (k < 0 || k >= KC_N) return false;
  *ms = r ? r->kms[k] : 0.0;
  *n = r ? r->kcount[k] : 0;
  return true;
}

bool kernel_stats(const Device *d, int k, double *ms, int64_t *n) {
  return runner_stats(d ? &d->run : nullptr, k, ms, n);
}

bool batch_kernel_stats(const axh_batch *b, int k, double *ms, int64_t *n) {
  if (!runner_stats(nullptr, k, ms, n)) return false;
  if (!b || !b->state) return true;
  for (const auto &r : b->state->runs) {
    double m = 0.0;
    int64_t c = 0;
    runner_stats(r.get(), k, &m, &c);
    *ms += m;
    *n += c;
  }
  return true;
}

bool in_running_batch(const axh_ctx *ctx) { return ctx->dev && ctx->dev->live; }

// Put every member on the device mesh of the first member of its source type
// (building what is missing) and move the members' state into the batch's
// interleaved groups: the members of each source type, in batch order, split
// into groups of at most the group width, sizes as equal as possible.
void batch_attach(axh_batch *bt) {
  if (bt->state) return;
  const std::vector<axh_ctx *> &members = bt->members;
  std::unique_ptr<BatchState, BatchStateDeleter> S(new BatchState);
  S->run.init();
  const int width = std::max(1, std::min(bt->group_width, kGroup));
  for (int type = AXH_MONOPOLE; type <= AXH_QUADPOLE; type++) {
    std::vector<int> of_type;
    for (size_t b = 0; b < members.size(); b++)
      if (members[b]->dims.src_type == type) of_type.push_back((int)b);
    if (of_type.empty()) continue;
    axh_ctx *first = members[of_type[0]];
    if (!first->dev) device_build(first);
    std::shared_ptr<MeshDev> mesh = first->dev->mesh;
    for (size_t q = 1; q < of_type.size(); q++) {
      axh_ctx *c = members[of_type[q]];
      if (!c->dev) {
        std::unique_ptr<Device> D(new Device);
        D->mesh = mesh;
        member_build(c, *mesh, D->member);
        fields_alloc(*mesh, D->own, 1);
        fields_set_sources(*mesh, D->own, {&D->member.src_g_h});
        c->dev.reset(D.release());
      } else if (c->dev->mesh != mesh) {
        // same mesh part (checked when the batch was created): the global
        // numbering is the same, so the wavefield part carries over unchanged
        if (c->dev->run.stream) HIPCHECK(hipStreamSynchronize(c->dev->run.stream));
        c->dev->mesh = mesh;
      }
    }
    const int n = (int)of_type.size();
    const int ng = (n + width - 1) / width;
    int next = 0;
    for (int g = 0; g < ng; g++) {
      const int size = n / ng + (g < n % ng ? 1 : 0);
      std::unique_ptr<Fields> F(new Fields);
      fields_alloc(*mesh, *F, size);
      std::vector<const std::vector<int> *> srcs;
      for (int b = 0; b < size; b++)
        srcs.push_back(&members[of_type[next + b]]->dev->member.src_g_h);
      fields_set_sources(*mesh, *F, srcs);
      std::vector<int> idx;
      for (int b = 0; b < size; b++, next++) {
        Device &D = *members[of_type[next]]->dev;
        if (D.run.stream) HIPCHECK(hipStreamSynchronize(D.run.stream));
        copy_state(*mesh, *F, b, D.own, 0, S->run.stream);
        idx.push_back(of_type[next]);
      }
      S->index.push_back(idx);
      S->groups.push_back(std::move(F));
      S->mesh.push_back(mesh);
      S->runs.emplace_back(new Runner);
      S->runs.back()->init();
    }
  }
  HIPCHECK(hipStreamSynchronize(S->run.stream));
  for (size_t g = 0; g < S->groups.size(); g++)
    for (size_t b = 0; b < S->index[g].size(); b++) {
      Device &D = *members[S->index[g][b]]->dev;
      D.live = S->groups[g].get();
      D.live_slot = (int)b;
    }
  bt->state = std::move(S);
}

// Move the members' state back to their own fields; the groups are freed.
void batch_detach(axh_batch *bt) {
  if (!bt->state) return;
  BatchState &S = *bt->state;
  std::string err;
  for (size_t g = 0; g < S.groups.size(); g++)
    for (size_t b = 0; b < S.index[g].size(); b++) {
      Device &D = *bt->members[S.index[g][b]]->dev;
      try {
        copy_state(*D.mesh, D.own, 0, *S.groups[g], (int)b, S.run.stream);
      } catch (const Error &e) {
        err = e.msg;
      }
      D.live = nullptr;
      D.live_slot = 0;
    }
  const hipError_t rc = hipStreamSynchronize(S.run.stream);
  bt->state.reset();
  if (!err.empty()) fail(AXH_ERR_HIP, err);
  if (rc != hipSuccess) fail(AXH_ERR_HIP, std::string("batch detach: ") + hipGetErrorString(rc));
}

// ---------------------------------------------------------------------------
// the time loop of one group (B wavefields on one mesh)
// ---------------------------------------------------------------------------
static inline int nblk(long n) { return (int)((n + 255) / 256); }

struct Launcher {
  Runner &R;
  template <class F> void operator()(int kc, F &&launch) {
    if (R.profile) {
      hipEvent_t a, b;
      HIPCHECK(hipEventCreate(&a));
      HIPCHECK(hipEventCreate(&b));
      HIPCHECK(hipEventRecord(a, R.stream));
      launch();
      HIPCHECK(hipEventRecord(b, R.stream));
      R.kev.push_back(a);
      R.kev.push_back(b);
      R.kev_class.push_back(kc);
    } else {
      launch();
    }
  }
};

// one launch: the solid tiles, then the fluid tiles (the Newmark
// memory-variable update sized for 5 linear solids where that suffices)
template <int SRC, int MODE, bool ANEL, int B>
static void launch_solid(const MeshDev &M, hipStream_t s, const SolidArgs &A,
                         const FluidArgs &F) {
  const int nblocks = M.ntile_s + (M.nel_f > 0 ? M.ntile_f : 0);
  if constexpr (B == 1 && zero_term_of(SRC) >= 0) {
    // one wavefield of an isotropic model: the +0 term is not read (a group's
    // launch is bound by instruction issue, not by the term bytes)
    if (M.zero_term) {
      if (MODE == 0 && ANEL && M.nsls <= 5)
        hipLaunchKernelGGL((k_solid<SRC, MODE, ANEL, 1, 5, true>), dim3(nblocks), dim3(256), 0,
                           s, A, F);
      else
        hipLaunchKernelGGL((k_solid<SRC, MODE, ANEL, 1, kMaxSls, true>), dim3(nblocks),
                           dim3(256), 0, s, A, F);
      return;
    }
  }
  if (MODE == 0 && ANEL && M.nsls <= 5)
    hipLaunchKernelGGL((k_solid<SRC, MODE, ANEL, B, 5>), dim3(nblocks), dim3(256), 0, s, A, F);
  else
    hipLaunchKernelGGL((k_solid<SRC, MODE, ANEL, B>), dim3(nblocks), dim3(256), 0, s, A, F);
}

template <int B>
static void dispatch_solid(const MeshDev &M, hipStream_t s, const SolidArgs &A,
                           const FluidArgs &F, int mode) {
  const bool an = M.anel != 0;
#define AXH_SOLID_CASE(S)                                                              \
  if (mode == 0) {                                                                     \
    if (an) launch_solid<S, 0, true, B>(M, s, A, F); else launch_solid<S, 0, false, B>(M, s, A, F); \
  } else {                                                                             \
    if (an) launch_solid<S, 1, true, B>(M, s, A, F); else launch_solid<S, 1, false, B>(M, s, A, F); \
  }
  if (M.src == AXH_MONOPOLE) { AXH_SOLID_CASE(AXH_MONOPOLE) }
  else if (M.src == AXH_DIPOLE) { AXH_SOLID_CASE(AXH_DIPOLE) }
  else { AXH_SOLID_CASE(AXH_QUADPOLE) }
#undef AXH_SOLID_CASE
}

template <bool STIFF, bool UPD, int B>
static void dispatch_anel_full(const MeshDev &M, hipStream_t s, const AnelFullArgs &A) {
  const dim3 grid(M.ntile_s), blk(256);
  if (M.src == AXH_MONOPOLE)
    hipLaunchKernelGGL((k_anel_full<AXH_MONOPOLE, STIFF, UPD, B>), grid, blk, 0, s, A);
  else if (M.src == AXH_DIPOLE)
    hipLaunchKernelGGL((k_anel_full<AXH_DIPOLE, STIFF, UPD, B>), grid, blk, 0, s, A);
  else
    hipLaunchKernelGGL((k_anel_full<AXH_QUADPOLE, STIFF, UPD, B>), grid, blk, 0, s, A);
}

// the SLS capacity of the kernel: 5 (the reference's default number of
// linear solids) or kMaxSls
template <int B, int MS>
static void dispatch_memvar_ms(const MeshDev &M, hipStream_t s, const SolidArgs &A,
                               const RecArgs &R, int nrec_blk) {
  const dim3 grid(M.ntile_s + nrec_blk);
  // one wavefield stages the tile's memory variables in dynamic LDS
  const size_t lds = B > 1 ? 0 : sizeof(float) * kElemPerTile * 24 * M.nsls;
  if (M.src == AXH_MONOPOLE)
    hipLaunchKernelGGL((k_memvar<AXH_MONOPOLE, B, MS>), grid, dim3(256), lds, s, A, R);
  else if (M.src == AXH_DIPOLE)
    hipLaunchKernelGGL((k_memvar<AXH_DIPOLE, B, MS>), grid, dim3(256), lds, s, A, R);
  else
    hipLaunchKernelGGL((k_memvar<AXH_QUADPOLE, B, MS>), grid, dim3(256), lds, s, A, R);
}
template <int B>
static void dispatch_memvar(const MeshDev &M, hipStream_t s, const SolidArgs &A,
                            const RecArgs &R, int nrec_blk) {
  if (M.nsls <= 5) dispatch_memvar_ms<B, 5>(M, s, A, R, nrec_blk);
  else dispatch_memvar_ms<B, kMaxSls>(M, s, A, R, nrec_blk);
}

// Queues on the stream of R the advance of each of the B wavefields of F
// (member data mb[b], dims of member 0) by nsteps, wavefield b through its
// steps start[b] .. start[b]+nsteps-1; seis[b] (may be NULL) is where
// wavefield b's samples go.  The host waits only when a wavefield's sample
// ring fills up inside the run; group_finish completes the run.  Groups on
// different streams (a batch's source types) run concurrently.
//
// Every launch reads its step from the group's device step base (StepRef),
// so the launches of one step are the same for all steps.  With
// R.graph_steps = S > 0 (and no per-kernel profiling) the first S steps are
// captured once into a graph of the group (ending with the advance of the
// step base by S) that is replayed for every S steps of the run; the steps
// left over are launched directly.  Both give the same results bit for bit.
// A kernel-wavefield gather to run with the steps of a group's wavefields;
// gen identifies their tables and rings (a replayed graph is captured for one
// gen); tail[b] copies wavefield b's snapshots out.
struct KwfLaunch {
  KwfArgs args;
  uint64_t gen = 0;
  KwfTail *tail[kGroup] = {};
};
constexpr int kKwfBlocks = 1024;   // grid of k_kwf (grid-stride over the points)

template <int B>
static void run_group(Runner &R, const MeshDev &M, Fields &F, Member *const *mb,
                      const axh_ctx *const *ctx, const int32_t *start, int32_t nsteps,
                      float *const *seis, const RunBuffers &RB, GroupTail &T,
                      const KwfLaunch *kw = nullptr) {
  if (F.B != B) fail(AXH_ERR_STATE, "group width mismatch");
  const axh_dims &d = ctx[0]->dims;
  const bool multipole = M.src != AXH_MONOPOLE;
  const bool newmark = M.scheme == AXH_NEWMARK2;
  const double dt = d.deltat, hdt = 0.5 * dt, hdt2 = 0.5 * (dt * dt);
  const size_t nslot_s = (size_t)M.ntile_s * kTile;
  Launcher L{R};
  hipStream_t st_ = R.stream;

  // the run's first steps and cleared blow-up flags
  StepSet SS{};
  SS.base = RB.step_base;
  for (int b = 0; b < B; b++) {
    SS.first[b] = start[b];
    SS.blow[b] = RB.blow[b];
  }
  hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, st_, SS, B);

  SolidArgs SA{};
  SA.nel = M.nel_s; SA.nsls = M.nsls; SA.nglob = M.nglob_s; SA.nslot = (int)nslot_s;
  SA.ntile = M.ntile_s;
  SA.G = M.G; SA.coef = M.coef_s; SA.m0 = M.m0_s; SA.einfo = M.einfo_s; SA.gid = M.gid_s;
  SA.cg_stage = M.cg_stage; SA.cg_upd = M.cg_upd; SA.aj_mu = M.aj_mu; SA.aj_k = M.aj_k;
  for (int s = 0; s < kMaxSls; s++) {
    SA.expw[s] = M.expw[s]; SA.tst[s] = M.tst[s]; SA.tstm1[s] = M.tstm1[s];
  }
  SA.disp = F.disp; SA.force = F.force_s; SA.memvar = F.memvar; SA.rsum = F.rsum;
  SA.trm1 = F.trm1; SA.devm1 = F.devm1;
  SA.flags = F.flags_s; SA.minv = M.minv_s; SA.velo = F.velo; SA.acc0 = F.acc0;
  SA.first_direct = M.first_direct;
  if (F.src_tiles.size() > (size_t)kMaxSrcTiles)
    fail(AXH_ERR_STATE, "source points in more tiles than the group's source elements");
  SA.nsrc_tile = (int)F.src_tiles.size();
  for (size_t q = 0; q < F.src_tiles.size(); q++) SA.src_tile[q] = F.src_tiles[q];
  SA.disp_w = F.disp; SA.hdt = hdt;
  AnelFullArgs AN{};
  if (M.anel_full) {
    AN.nel = M.nel_s; AN.nsls = M.nsls; AN.nglob = M.nglob_s; AN.nslot = (int)nslot_s;
    AN.G = M.G; AN.terms = M.af_terms; AN.axis = M.af_axis; AN.einfo = M.einfo_s;
    AN.gid = M.gid_s; AN.aj_mu = M.aj_mu; AN.aj_k = M.aj_k;
    for (int s = 0; s < kMaxSls; s++) {
      AN.expw[s] = M.expw[s]; AN.tst[s] = M.tst[s]; AN.tstm1[s] = M.tstm1[s];
    }
    AN.disp = F.disp; AN.force = F.force_s; AN.memvar = F.memvar; AN.trm1 = F.trm1;
    AN.devm1 = F.devm1; AN.rsum = F.rsum;
  }
  FluidArgs FA{};
  FA.nel = M.nel_f; FA.multipole = multipole; FA.G = M.G; FA.coef = M.coef_f;
  FA.m0 = M.m0_f; FA.einfo = M.einfo_f; FA.gid = M.gid_f; FA.chi = F.chi;
  FA.force = F.force_f;
  // solid axis masks: monopole s; dipole +,z (2,3); quadrupole all
  const int mask0 = (M.src == AXH_DIPOLE) ? 1 : 0;
  const int mask1 = (M.src == AXH_MONOPOLE) ? 0 : 2;
  const int comps = multipole ? 7 : 5;
  PredArgs PS{};
  PS.n = M.nglob_s; PS.mask0 = mask0; PS.mask1 = mask1; PS.comps = comps;
  PS.dt = dt; PS.hdt2 = hdt2; PS.flags = F.flags_s; PS.stride = M.nglob_s;
  PS.naxb = M.naxb; PS.axb_g = M.axb_g; PS.u = F.disp; PS.v = F.velo; PS.a0 = F.acc0;
  PS.axb_val = F.axb_val;
  PredArgs PF{};
  PF.n = M.nglob_f; PF.mask0 = 0; PF.mask1 = 0; PF.comps = 1; PF.dt = dt; PF.hdt2 = hdt2;
  PF.flags = M.flags_f; PF.stride = M.nglob_f; PF.naxb = 0; PF.axb_g = nullptr;
  PF.u = F.chi; PF.v = F.dchi; PF.a0 = F.ddchi; PF.axb_val = F.axb_val;
  AsmSolidArgs AS{};
  AS.n = M.nglob_s; AS.src = M.src; AS.newmark = newmark; AS.nslot = (int)nslot_s;
  AS.comps = comps; AS.mask0 = mask0; AS.mask1 = mask1; AS.gstart = M.gstart_s;
  AS.npts = F.nasm; AS.pts = F.asm_pts;
  // one wavefield: the assembled points are first the mesh's range
  // [0, first_direct) in order, then the group's source points beyond it
  AS.ncontig = F.B == 1 ? std::min(M.first_direct, F.nasm) : 0;
  AS.gslot = M.gslot_s; AS.flags = F.flags_s; AS.minv = M.minv_s; AS.hdt = hdt;
  AS.gamma = M.gamma_s;
  AS.force = F.force_s; AS.velo = F.velo; AS.acc0 = F.acc0;
  AS.disp = F.disp; AS.naxb = M.naxb; AS.axb_g = M.axb_g; AS.axb_val = F.axb_val;
  AS.dipole = M.src == AXH_DIPOLE; AS.cpl_slot = M.cpl_slot_s; AS.cpl_gf = M.cpl_gf;
  AS.cpl_bm = M.cpl_bm; AS.ddchi = F.ddchi;
  AsmFluidArgs AF{};
  AF.n = M.nglob_f; AF.newmark = newmark; AF.mask_axis = multipole;
  AF.mask_surface = newmark; AF.gstart = M.gstart_f; AF.gslot = M.gslot_f;
  AF.flags = M.flags_f; AF.minv = M.minv_f; AF.hdt = hdt; AF.force = F.force_f;
  AF.gamma = M.gamma_f;
  AF.dchi = F.dchi; AF.ddchi = F.ddchi; AF.chi = F.chi;
  AF.dipole = M.src == AXH_DIPOLE; AF.nglob_s = M.nglob_s; AF.cpl_slot = M.cpl_slot_f;
  AF.cpl_gs = M.cpl_gs; AF.cpl_axb = M.cpl_axb; AF.cpl_bm = M.cpl_bm; AF.disp = F.disp;
  AF.axb_val = F.axb_val;
  RecArgs RA{};
  RA.src = M.src; RA.nglob = M.nglob_s; RA.disp = F.disp;
  int maxrec = 0;
  for (int b = 0; b < B; b++) {
    Member &w = *mb[b];
    FA.srcel[b] = w.srcel_f; FA.src_fl[b] = w.src_fl;
    PS.blow[b] = RB.blow[b]; PS.thr[b] = w.blow_thr;
    PF.blow[b] = RB.blow[b]; PF.thr[b] = w.blow_thr;
    AS.nsrc[b] = w.nsrc; AS.src_g[b] = w.src_g; AS.src_val[b] = w.src_val;
    RA.nrec[b] = w.nrec; RA.rec_g[b] = w.rec_g; RA.ring[b] = RB.ring[b];
    RA.seis_it[b] = std::max(w.seis_it, 1); RA.cap[b] = RB.cap[b];
    maxrec = std::max(maxrec, w.nrec);
  }

  HIPCHECK(hipEventRecord(R.ev0, st_));
  T.B = B;
  T.rb = RB;
  for (int b = 0; b < B; b++) {
    T.mb[b] = mb[b];
    T.seis[b] = seis ? seis[b] : nullptr;
    const int si = std::max(mb[b]->seis_it, 1);
    T.q0[b] = T.qdone[b] = T.qout[b] = (start[b] - 1) / si;
  }
  const StepRef sref{RB.step_base, 0};
  PS.step = PF.step = FA.step = AS.step = RA.step = sref;
  KwfArgs KA{};
  const int kwf_blocks = kw ? std::max(1, std::min(kKwfBlocks, nblk(kw->args.npt))) : 0;
  if (kw) {
    KA = kw->args;
    KA.step = sref;
    bool any0 = false;
    for (int b = 0; b < B; b++) any0 = any0 || start[b] == 1;
    if (any0 && KA.npt > 0) {
      // the initial fields (iter 0) of the wavefields at their first step, as
      // dump_stuff(0, ...) before the loop
      KA.step.koff = -1;
      KA.init_only = 1;
      L(KC_REC, [&] {
        hipLaunchKernelGGL(k_kwf<B>, dim3(kwf_blocks, B), dim3(256), 0, st_, KA);
      });
      KA.init_only = 0;
    }
  }
  FA.nstages = AS.nstages = M.nstages;
  for (int b = 0; b < B; b++) {
    FA.stf[b] = (newmark && mb[b]->fluid_src) ? mb[b]->stf_d : nullptr;
    AS.stf[b] = (newmark && mb[b]->fluid_src) ? nullptr : mb[b]->stf_d;
  }
  // the launches of the step koff steps past the step base; with sample_all
  // the receiver kernels run every step (a graph's steps are all alike), else
  // only where a wavefield takes a sample (it: the step on the host)
  auto one_step = [&](int koff, bool sample_all, int step) {
    PS.step.koff = PF.step.koff = FA.step.koff = AS.step.koff = RA.step.koff = koff;
    for (int st = 0; st < M.nstages; st++) {
      // predictor / first drift of the step; the symplectic drifts of the
      // later stages and the final drift are fused into the kicks below
      if (newmark || st == 0) {
        PS.newmark = PF.newmark = newmark;
        PS.cd = PF.cd = newmark ? 0.0 : M.coefd[st];
        PS.apply_mask = 1;
        PF.apply_mask = multipole;
        PS.check_blow = 1;
        PF.check_blow = 0;
        L(KC_PRED, [&] {
          const int nbs = nblk(PS.n), nbf = M.nel_f > 0 ? nblk(PF.n) : 0;
          hipLaunchKernelGGL(k_predict<B>, dim3(nbs + nbf), dim3(256), 0, st_, PS, PF, nbs);
        });
      }
      // drift after the kick: coefd(st+1), masked except the final one
      // (SOLVER/time_evol_wave.F90:597-605,720-725)
      const bool last = st + 1 == M.nstages;
      AS.drift = AF.drift = !newmark;
      AS.cd_next = AF.cd_next = newmark ? 0.0 : M.coefd[st + 1];
      AS.drift_mask = !last;
      AF.drift_mask = !last && multipole;
      FA.stage = AS.stage = st;
      SA.direct = 1;
      SA.cv = newmark ? 0.0 : M.coefv[st];
      SA.cd_next = newmark ? 0.0 : M.coefd[st + 1];
      L(KC_SOLID, [&] { dispatch_solid<B>(M, st_, SA, FA, newmark ? 0 : 1); });
      SA.direct = 0;
      if (M.anel_full) {
        // anelastic forces; Newmark also updates the memory variables here
        L(KC_MEMVAR, [&] {
          if (newmark) dispatch_anel_full<true, true, B>(M, st_, AN);
          else dispatch_anel_full<true, false, B>(M, st_, AN);
        });
      }
      if (M.nel_f > 0) {
        AF.cv = newmark ? 0.0 : M.coefv[st];
        L(KC_ASM_F, [&] {
          hipLaunchKernelGGL(k_asm_fluid<B>, dim3(nblk(AF.n)), dim3(256), 0, st_, AF);
        });
      }
      AS.cv = newmark ? 0.0 : M.coefv[st];
      L(KC_ASM_S, [&] {
        hipLaunchKernelGGL(k_asm_solid<B>, dim3(nblk(AS.npts)), dim3(256), 0, st_, AS);
      });
    }
    bool any = sample_all && maxrec > 0;
    for (int b = 0; b < B && !any; b++)
      any = mb[b]->nrec > 0 && (start[b] + step) % mb[b]->seis_it == 0;
    if (!newmark && M.anel_full)
      L(KC_MEMVAR, [&] { dispatch_anel_full<false, true, B>(M, st_, AN); });
    if (!newmark && M.anel) {
      // the receiver samples ride along as extra blocks of the update
      const int nrb = any ? nblk(maxrec) * B : 0;
      L(KC_MEMVAR, [&] { dispatch_memvar<B>(M, st_, SA, RA, nrb); });
      any = false;
    }
    if (any)
      L(KC_REC, [&] {
        hipLaunchKernelGGL(k_receivers<B>, dim3(nblk(maxrec), B), dim3(256), 0, st_, RA);
      });
    // kernel-wavefield snapshot of the step (a graph's steps all launch it;
    // the kernel returns at once where the step is not a snapshot step)
    bool snap = kw && KA.npt > 0 && sample_all;
    for (int b = 0; b < B && kw && KA.npt > 0 && !snap; b++)
      snap = (start[b] + step) % KA.strain_it == 0;
    if (snap) {
      KA.step.koff = koff;
      L(KC_REC, [&] {
        hipLaunchKernelGGL(k_kwf<B>, dim3(kwf_blocks, B), dim3(256), 0, st_, KA);
      });
    }
  };
  auto reserve = [&](int step, int n) {
    for (int b = 0; b < B; b++) T.reserve(b, start[b] + step - 1, n, st_);
    for (int b = 0; b < B && kw; b++)
      if (kw->tail[b]) kw->tail[b]->reserve(start[b] + step - 1, n, st_);
  };

  int step = 0;
  const int S = R.graph_steps;
  bool fits = S > 0 && !R.profile && nsteps >= S;
  for (int b = 0; b < B && fits; b++)
    fits = mb[b]->nrec <= 0 || S / std::max(mb[b]->seis_it, 1) + 1 <= RB.cap[b];
  if (kw) fits = fits && S / kw->args.strain_it + 1 <= kw->args.cap;
  if (fits) {
    const uint64_t kgen = kw ? kw->gen : 0;
    if (!R.gexec || R.gfields != &F || R.gsteps != S || R.gbase != RB.step_base ||
        R.gkwf != kgen) {
      if (R.gexec) HIPCHECK(hipGraphExecDestroy(R.gexec));
      R.gexec = nullptr;
      hipGraph_t g = nullptr;
      HIPCHECK(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < S; k++) one_step(k, true, k);
      hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, st_, RB.step_base, B, S);
      HIPCHECK(hipStreamEndCapture(st_, &g));
      const hipError_t rc = hipGraphInstantiate(&R.gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (rc != hipSuccess) {
        R.gexec = nullptr;
        fail(AXH_ERR_HIP, std::string("graph of the step: ") + hipGetErrorString(rc));
      }
      R.gfields = &F;
      R.gsteps = S;
      R.gbase = RB.step_base;
      R.gkwf = kgen;
    }
    for (; step + S <= nsteps; step += S) {
      reserve(step, S);
      HIPCHECK(hipGraphLaunch(R.gexec, st_));
    }
  }
  // the rest, launched directly (koff counts from the current step base)
  for (int k = 0; step < nsteps; step++, k++) {
    reserve(step, 1);
    one_step(k, false, step);
  }
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipEventRecord(R.ev1, st_));
}

// Completes a group's run: copies out the staged samples, waits for the
// stream and collects the per-kernel times; blow[b] receives wavefield b's
// blow-up step (INT_MAX: none).  Returns the device time of the run [ms].
static double group_finish(Runner &R, GroupTail &T, int *blow) {
  for (int b = 0; b < T.B; b++) {
    HIPCHECK(hipMemcpyAsync(&blow[b], T.rb.blow[b], sizeof(int), hipMemcpyDeviceToHost,
                            R.stream));
    T.copy_out(b, R.stream);
  }
  HIPCHECK(hipStreamSynchronize(R.stream));
  float ms = 0.0f;
  HIPCHECK(hipEventElapsedTime(&ms, R.ev0, R.ev1));
  if (R.profile) {
    for (size_t q = 0; q < R.kev_class.size(); q++) {
      float kms = 0.0f;
      HIPCHECK(hipEventElapsedTime(&kms, R.kev[2 * q], R.kev[2 * q + 1]));
      R.kms[R.kev_class[q]] += kms;
      R.kcount[R.kev_class[q]] += 1;
    }
    for (auto e : R.kev) (void)hipEventDestroy(e);
    R.kev.clear();
    R.kev_class.clear();
  }
  return ms;
}

static void run_any(Runner &R, const MeshDev &M, Fields &F, Member *const *mb,
                    const axh_ctx *const *ctx, const int32_t *start, int32_t nsteps,
                    float *const *seis, const RunBuffers &RB, GroupTail &T,
                    const KwfLaunch *kw = nullptr) {
  switch (F.B) {
    case 1: return run_group<1>(R, M, F, mb, ctx, start, nsteps, seis, RB, T, kw);
    case 2: return run_group<2>(R, M, F, mb, ctx, start, nsteps, seis, RB, T, kw);
    case 3: return run_group<3>(R, M, F, mb, ctx, start, nsteps, seis, RB, T, kw);
    case 4: return run_group<4>(R, M, F, mb, ctx, start, nsteps, seis, RB, T, kw);
    default: fail(AXH_ERR_STATE, "group width out of range");
  }
}

// Queues steps it0..it1 of a context on its stream; device_run_end
// completes them.  Between the two the host is free (e.g. to write the
// previous run's samples) and seis must stay untouched.
void device_run_begin(axh_ctx *ctx, int32_t it0, int32_t it1, float *seis) {
  Device &V = *ctx->dev;
  if (V.live) fail(AXH_ERR_STATE, "the context is stepped by a batch");
  V.run.init();
  V.run.profile = ctx->profile;
  V.run.graph_steps = ctx->graph_steps;
  Member *mb[1] = {&V.member};
  const axh_ctx *cx[1] = {ctx};
  float *out[1] = {seis};
  const int32_t start[1] = {it0};
  V.pending = GroupTail();
  run_any(V.run, *V.mesh, V.own, mb, cx, start, it1 - it0 + 1, out,
          run_buffers(V.own, mb, 1), V.pending);
  V.pending_it1 = it1;
}

void device_run_end(axh_ctx *ctx) {
  Device &V = *ctx->dev;
  int blow = INT_MAX;
  ctx->last_run_ms = group_finish(V.run, V.pending, &blow);
  if (V.pending_has_kwf) {
    V.pending_has_kwf = false;
    V.pending_kwf.complete();
  }
  ctx->iter_done = V.pending_it1;
  if (blow != INT_MAX)
    fail(AXH_ERR_BLOWUP, "displacements blew up at time step " + std::to_string(blow));
}

void device_run(axh_ctx *ctx, int32_t it0, int32_t it1, float *seis) {
  device_run_begin(ctx, it0, it1, seis);
  device_run_end(ctx);
}

// ---------------------------------------------------------------------------
// kernel-wavefield runs (axh_run_kwf)
// ---------------------------------------------------------------------------
constexpr size_t kKwfRingBytes = AXH_KWF_RING_BYTES;  // snapshot ring on the device
constexpr size_t kKwfStageBytes = (size_t)16 << 20;   // one pinned staging half

// The point tables of k_kwf from the context's grid, in point order: the
// solid points' global ids (-1 in a source element when src_mask), the fluid
// points' row and column ids with their element terms; and a ring of
// min(snapshots of the whole run, kKwfRingBytes / snapshot) snapshots, or
// the caller's ring size.
static void kwf_build(axh_ctx *ctx, Device &V) {
  const axh_dims &d = ctx->dims;
  const MeshDev &M = *V.mesh;
  const int32_t *np = ctx->need("kwf_npoints").as<int32_t>();
  const int32_t *mapping = ctx->need("kwf_mapping").as<int32_t>();
  const int32_t *mask = ctx->need("kwf_mask").as<int32_t>();
  std::unique_ptr<KwfDev> K(new KwfDev);
  K->nps = np[0];
  K->npf = np[1];
  K->ncomp = d.src_type == AXH_MONOPOLE ? 2 : 3;
  K->strain_it = ctx->kwf_strain_it;
  K->src_mask = ctx->kwf_src_mask;
  const int npt = K->nps + K->npf;
  // source elements zeroed (src_dump_type 'mask'): ielsrc holds solid
  // elements unless the source sits in the fluid, where nothing is zeroed
  std::vector<char> zero(d.nel_solid, 0);
  if (K->src_mask && !d.fluid_src && d.nelsrc > 0) {
    const int32_t *ielsrc = ctx->need("ielsrc").as<int32_t>();
    for (int q = 0; q < d.nelsrc; q++)
      if (ielsrc[q] >= 1 && ielsrc[q] <= d.nel_solid) zero[ielsrc[q] - 1] = 1;
  }
  std::vector<int> sol_g(std::max(K->nps, 1), -1), flu_g(10 * (size_t)std::max(K->npf, 1), 0),
      flu_ij(std::max(K->npf, 1), 0);
  std::vector<float> flu_c(5 * (size_t)std::max(K->npf, 1), 0.0f);
  const float *fc[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  if (K->npf > 0)
    for (int q = 0; q < 5; q++) fc[q] = ctx->need(kFluidGradNames[q]).as<float>();
  const int32_t *axf = d.nel_fluid > 0 ? ctx->need("axis_fluid").as<int32_t>() : nullptr;
  for (int e = 0; e < d.nel_solid + d.nel_fluid; e++)
    for (int p = 0; p < kNpp; p++) {
      const size_t x = (size_t)e * kNpp + p;
      if (!mask[x]) continue;
      const int k = mapping[x] - 1;
      if (e < d.nel_solid) {
        if (k < 0 || k >= K->nps) fail(AXH_ERR_STATE, "bad kernel-wavefield grid");
        sol_g[k] = zero[e] ? -1 : M.gid_s_h[slot_of(e, p)];
      } else {
        const int ef = e - d.nel_solid, q = k - K->nps;
        if (q < 0 || q >= K->npf) fail(AXH_ERR_STATE, "bad kernel-wavefield grid");
        const int i = p % 5, j = p / 5;
        for (int kk = 0; kk < 5; kk++) {
          flu_g[(size_t)q * 10 + kk] = M.gid_f_h[slot_of(ef, kk + 5 * j)];
          flu_g[(size_t)q * 10 + 5 + kk] = M.gid_f_h[slot_of(ef, i + 5 * kk)];
        }
        flu_ij[q] = i | (j << 3) | ((axf[ef] ? 1 : 0) << 6);
        for (int c = 0; c < 5; c++) flu_c[(size_t)c * K->npf + q] = fc[c][(size_t)ef * kNpp + p];
      }
    }
  const size_t snap = (size_t)K->ncomp * std::max(npt, 1);
  const size_t snap_b = snap * sizeof(float);
  const int total = d.niter / K->strain_it + 1;
  // staging halves of up to 16 MiB (at least one snapshot); the ring holds two halves and the
  // snapshots of one replayed graph (at least 2 slots: a run from step 1
  // takes the initial snapshot and the first), at most 1 GiB
  if (ctx->kwf_ring > 0) {
    K->cap = ctx->kwf_ring;
    K->half = (int)std::max<size_t>(1, std::min<size_t>(K->cap / 2, kKwfStageBytes / snap_b));
  } else {
    K->half = (int)std::max<size_t>(1, std::min<size_t>(total, kKwfStageBytes / snap_b));
    const size_t want = 2 * (size_t)K->half + std::min(total, AXH_MAX_GRAPH_STEPS / K->strain_it + 2);
    K->cap = (int)std::max<size_t>(2, std::min(want, std::max<size_t>(2, kKwfRingBytes / snap_b)));
  }
  K->ring_req = ctx->kwf_ring;
  K->sol_g = K->mem.upload(sol_g);
  K->flu_g = K->mem.upload(flu_g);
  K->flu_ij = K->mem.upload(flu_ij);
  K->flu_c = K->mem.upload(flu_c);
  K->ring = K->mem.alloc<float>(snap * K->cap);
  HIPCHECK(hipHostMalloc((void **)&K->stage, 2 * (size_t)K->half * snap_b, hipHostMallocDefault));
  for (auto &e : K->ev) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  static std::atomic<uint64_t> builds{0};
  K->gen = ++builds;
  V.kwf = std::move(K);
}

// The context's gather tables and ring for its current settings (built on
// first use and when axh_set_kwf changed them).
static KwfDev &kwf_of(axh_ctx *ctx) {
  Device &V = *ctx->dev;
  if (!V.kwf || V.kwf->strain_it != ctx->kwf_strain_it || V.kwf->src_mask != ctx->kwf_src_mask ||
      V.kwf->ring_req != ctx->kwf_ring)
    kwf_build(ctx, V);
  return *V.kwf;
}

// The gather of a group of B wavefields with tables K[b] on mesh part M and
// fields F, their snapshots copied to out[b] from step start[b] on.  The
// wavefields must share the grid, strain_it and ring size.
static void kwf_group(KwfDev *const *K, int B, const MeshDev &M, const Fields &F,
                      const int32_t *start, float *const *out, KwfLaunch &kl, KwfTail *kt) {
  const KwfDev &K0 = *K[0];
  for (int b = 1; b < B; b++)
    if (K[b]->nps != K0.nps || K[b]->npf != K0.npf || K[b]->ncomp != K0.ncomp ||
        K[b]->strain_it != K0.strain_it || K[b]->cap != K0.cap || K[b]->half != K0.half)
      fail(AXH_ERR_STATE, "the kernel-wavefield grids or settings of a group's members differ");
  KwfArgs &A = kl.args;
  A = KwfArgs{};
  A.nps = K0.nps; A.npt = K0.nps + K0.npf; A.ncomp = K0.ncomp;
  A.dipole = M.src == AXH_DIPOLE; A.strain_it = K0.strain_it; A.cap = K0.cap;
  A.nglob_s = M.nglob_s;
  A.flu_g = K0.flu_g; A.flu_ij = K0.flu_ij; A.flu_c = K0.flu_c;
  A.G = M.G; A.disp = F.disp; A.chi = F.chi;
  kl.gen = 0;
  const size_t snap = (size_t)K0.ncomp * A.npt;
  for (int b = 0; b < B; b++) {
    const KwfDev &Kb = *K[b];
    A.sol_g[b] = Kb.sol_g;
    A.ring[b] = Kb.ring;
    kl.gen = kl.gen * 0x9E3779B97F4A7C15ull + Kb.gen;
    KwfTail &t = kt[b];
    t = KwfTail();
    t.s = Kb.strain_it;
    t.cap = Kb.cap;
    t.H = Kb.half;
    t.snap = snap;
    t.ring = Kb.ring;
    t.stage[0] = Kb.stage;
    t.stage[1] = Kb.stage + (size_t)Kb.half * snap;
    t.ev[0] = Kb.ev[0];
    t.ev[1] = Kb.ev[1];
    t.out = out[b];
    t.first_q = start[b] == 1 ? 0 : (start[b] + Kb.strain_it - 1) / Kb.strain_it;
    t.qout = t.qdone = t.first_q - 1;
    kl.tail[b] = A.npt > 0 ? &t : nullptr;
  }
}

// Queues steps it0..it1 with the kernel-wavefield snapshots of their steps
// (and of the initial fields when it0 = 1), written to kwf in step order;
// device_run_end completes them.  The snapshots are copied out through the
// pinned staging halves while the following steps run (KwfTail); between the
// two calls the host is free, and seis and kwf must stay untouched.
void device_run_kwf_begin(axh_ctx *ctx, int32_t it0, int32_t it1, float *seis, float *kwf) {
  Device &V = *ctx->dev;
  if (V.live) fail(AXH_ERR_STATE, "the context is stepped by a batch");
  KwfDev *K[1] = {&kwf_of(ctx)};
  KwfLaunch kl;
  float *kout[1] = {kwf};
  const int32_t start[1] = {it0};
  kwf_group(K, 1, *V.mesh, V.own, start, kout, kl, &V.pending_kwf);
  V.pending_has_kwf = kl.tail[0] != nullptr;
  V.run.init();
  V.run.profile = ctx->profile;
  V.run.graph_steps = ctx->graph_steps;
  Member *mb[1] = {&V.member};
  const axh_ctx *cx[1] = {ctx};
  float *out[1] = {seis};
  V.pending = GroupTail();
  run_any(V.run, *V.mesh, V.own, mb, cx, start, it1 - it0 + 1, out, run_buffers(V.own, mb, 1),
          V.pending, &kl);
  if (V.pending_has_kwf) V.pending_kwf.finish(V.run.stream);
  V.pending_it1 = it1;
}

void device_run_kwf(axh_ctx *ctx, int32_t it0, int32_t it1, float *seis, float *kwf) {
  device_run_kwf_begin(ctx, it0, it1, seis, kwf);
  device_run_end(ctx);
}

// The groups of a batch are queued one after the other, each on its own
// stream (groups of different source types, or more groups than one width
// holds, run concurrently), then completed in order.  With kwf, every
// member's kernel-wavefield snapshots of the steps go to kwf[member].
void batch_run(axh_batch *bt, int32_t nsteps, float *const *seis, float *const *kwf) {
  BatchState &S = *bt->state;
  std::vector<int> blow(bt->members.size(), INT_MAX);
  std::vector<GroupTail> tails(S.groups.size());
  std::vector<KwfLaunch> klaunch(kwf ? S.groups.size() : 0);
  std::vector<std::array<KwfTail, kGroup>> ktails(kwf ? S.groups.size() : 0);
  for (size_t g = 0; g < S.groups.size(); g++) {
    Runner &R = *S.runs[g];
    R.profile = bt->profile;
    R.graph_steps = bt->graph_steps;
    const std::vector<int> &idx = S.index[g];
    const MeshDev &M = *S.mesh[g];
    Member *mb[kGroup];
    const axh_ctx *cx[kGroup];
    float *out[kGroup];
    int32_t start[kGroup];
    for (size_t b = 0; b < idx.size(); b++) {
      axh_ctx *c = bt->members[idx[b]];
      if (c->dev->mesh.get() != &M) fail(AXH_ERR_STATE, "batch members on different meshes");
      mb[b] = &c->dev->member;
      cx[b] = c;
      out[b] = seis ? seis[idx[b]] : nullptr;
      start[b] = c->iter_done + 1;
    }
    KwfLaunch *kl = nullptr;
    if (kwf) {
      KwfDev *K[kGroup];
      float *kout[kGroup];
      for (size_t b = 0; b < idx.size(); b++) {
        K[b] = &kwf_of(bt->members[idx[b]]);
        kout[b] = kwf[idx[b]];
      }
      kwf_group(K, (int)idx.size(), M, *S.groups[g], start, kout, klaunch[g], ktails[g].data());
      kl = &klaunch[g];
    }
    run_any(R, M, *S.groups[g], mb, cx, start, nsteps, out,
            run_buffers(*S.groups[g], mb, (int)idx.size()), tails[g], kl);
    for (size_t b = 0; kl && b < idx.size(); b++)
      if (kl->tail[b]) kl->tail[b]->finish(R.stream);
  }
  // device time of the batch: first start to last end over the streams
  // (times relative to group 0's start)
  double first = 0.0, last = 0.0;
  for (size_t g = 0; g < S.groups.size(); g++) {
    int bl[kGroup];
    group_finish(*S.runs[g], tails[g], bl);
    const std::vector<int> &idx = S.index[g];
    for (size_t b = 0; kwf && b < idx.size(); b++)
      if (klaunch[g].tail[b]) klaunch[g].tail[b]->complete();
    for (size_t b = 0; b < idx.size(); b++) blow[idx[b]] = bl[b];
    float t0 = 0.0f, t1 = 0.0f;
    HIPCHECK(hipEventElapsedTime(&t0, S.runs[0]->ev0, S.runs[g]->ev0));
    HIPCHECK(hipEventElapsedTime(&t1, S.runs[0]->ev0, S.runs[g]->ev1));
    first = std::min(first, (double)t0);
    last = std::max(last, (double)t1);
  }
  bt->last_run_ms = last - first;
  for (axh_ctx *c : bt->members) {
    c->iter_done += nsteps;
    c->last_run_ms = bt->last_run_ms;
  }
  for (size_t b = 0; b < blow.size(); b++)
    if (blow[b] != INT_MAX)
      fail(AXH_ERR_BLOWUP, "batch member " + std::to_string(b) +
                               ": displacements blew up at time step " + std::to_string(blow[b]));
}

// ---------------------------------------------------------------------------
// runs that allocate nothing (axh_run_noalloc)
// ---------------------------------------------------------------------------
// The device buffers a run writes besides the fields -- the step base, the
// blow-up flag and the receiver-sample ring -- in one fixed array per GPU
// instead of the context's own.  Every call on a GPU shares that GPU's array
// and finds it as the last call left it (each run sets the step base and
// clears the flag itself); calls that overlap on one GPU race.
constexpr size_t kScratchBytes = (size_t)AXH_NOALLOC_SCRATCH_BYTES;
constexpr size_t kScratchHead = 256;   // step base [kGroup], blow-up flag
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;

// the array's device address on the current GPU (in host code `scratch` only
// names a host-side stand-in), looked up once per GPU
static unsigned char *scratch_ptr() {
  static std::mutex mu;
  static unsigned char *ptr[kMaxDevices] = {};
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail(AXH_ERR_HIP, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void *p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !p)
      fail(AXH_ERR_HIP, std::string("device address of the scratch array: ") +
                            hipGetErrorString(rc));
    ptr[dev] = (unsigned char *)p;
  }
  return ptr[dev];
}

int noalloc_max_receivers() {
  return (int)((kScratchBytes - kScratchHead) / (3 * sizeof(float)));
}

void device_run_noalloc(axh_ctx *ctx, int32_t it0, int32_t it1, float *seis) {
  Device &V = *ctx->dev;
  if (V.live) fail(AXH_ERR_STATE, "the context is stepped by a batch");
  if (!V.run.stream)
    fail(AXH_ERR_STATE, "axh_run_noalloc needs a context that has run (axh_run) before");
  unsigned char *base = scratch_ptr();
  Member *mb[1] = {&V.member};
  RunBuffers RB;
  RB.step_base = (int *)base;
  RB.blow[0] = (int *)base + kGroup;
  RB.ring[0] = (float *)(base + kScratchHead);
  RB.cap[0] = (int)((kScratchBytes - kScratchHead) /
                    (3 * sizeof(float) * (size_t)std::max(V.member.nrec, 1)));
  // per-kernel profiling creates events: off for this entry point
  V.run.profile = false;
  V.run.graph_steps = ctx->graph_steps;
  const axh_ctx *cx[1] = {ctx};
  float *out[1] = {seis};
  int blow = INT_MAX;
  const int32_t start[1] = {it0};
  GroupTail T;
  run_any(V.run, *V.mesh, V.own, mb, cx, start, it1 - it0 + 1, out, RB, T);
  ctx->last_run_ms = group_finish(V.run, T, &blow);
  ctx->iter_done = it1;
  if (blow != INT_MAX)
    fail(AXH_ERR_BLOWUP, "displacements blew up at time step " + std::to_string(blow));
}

// The live copy of ctx's state: its own fields, or its slot in a batch group.
static Fields &live_fields(Device &V, int &slot) {
  if (V.live) {
    slot = V.live_slot;
    return *V.live;
  }
  slot = 0;
  return V.own;
}

static void sync_member(Device &V) {
  if (V.run.stream) HIPCHECK(hipStreamSynchronize(V.run.stream));
  if (V.live) HIPCHECK(hipDeviceSynchronize());
}

// one wavefield's values of an interleaved device array
static std::vector<float> fetch(const float *dsrc, size_t n, int B, int slot) {
  std::vector<float> h(n * B), out(n);
  HIPCHECK(hipMemcpy(h.data(), dsrc, h.size() * 4, hipMemcpyDeviceToHost));
  for (size_t x = 0; x < n; x++) out[x] = h[x * B + slot];
  return out;
}
static void store(float *ddst, const std::vector<float> &v, int B, int slot) {
  std::vector<float> h(v.size() * B);
  HIPCHECK(hipMemcpy(h.data(), ddst, h.size() * 4, hipMemcpyDeviceToHost));
  for (size_t x = 0; x < v.size(); x++) h[x * B + slot] = v[x];
  HIPCHECK(hipMemcpy(ddst, h.data(), h.size() * 4, hipMemcpyHostToDevice));
}

// element-local fields from the per-point device arrays
void device_download(axh_ctx *ctx, float *disp, float *velo, float *chi,
                     float *dchi, float *ddchi) {
  Device &V = *ctx->dev;
  const MeshDev &M = *V.mesh;
  sync_member(V);
  int slot = 0;
  Fields &F = live_fields(V, slot);
  auto scatter_s = [&](const float *dsrc, float *out) {
    const std::vector<float> h = fetch(dsrc, 3 * (size_t)M.nglob_s, F.B, slot);
    for (int c = 0; c < 3; c++)
      for (int e = 0; e < M.nel_s; e++)
        for (int k = 0; k < kNpp; k++)
          out[((size_t)c * M.nel_s + e) * kNpp + k] =
              h[(size_t)c * M.nglob_s + M.gid_s_h[slot_of(e, k)]];
  };
  if (disp) scatter_s(F.disp, disp);
  if (velo) scatter_s(F.velo, velo);
  if (M.nel_f > 0 && (chi || dchi || ddchi)) {
    auto scatter_f = [&](const float *dsrc, float *out) {
      const std::vector<float> h = fetch(dsrc, M.nglob_f, F.B, slot);
      for (int e = 0; e < M.nel_f; e++)
        for (int k = 0; k < kNpp; k++) out[(size_t)e * kNpp + k] = h[M.gid_f_h[slot_of(e, k)]];
    };
    if (chi) scatter_f(F.chi, chi);
    if (dchi) scatter_f(F.dchi, dchi);
    if (ddchi) scatter_f(F.ddchi, ddchi);
  }
}

void device_upload_fields(axh_ctx *ctx, const float *disp, const float *velo,
                          const float *chi, const float *dchi, const float *ddchi) {
  Device &V = *ctx->dev;
  const MeshDev &M = *V.mesh;
  sync_member(V);
  int slot = 0;
  Fields &F = live_fields(V, slot);
  auto gather_s = [&](const float *in, float *ddst, const char *what) {
    std::vector<float> h(3 * (size_t)M.nglob_s, 0.0f);
    std::vector<char> seen(h.size(), 0);
    for (int c = 0; c < 3; c++)
      for (int e = 0; e < M.nel_s; e++)
        for (int k = 0; k < kNpp; k++) {
          size_t g = (size_t)c * M.nglob_s + M.gid_s_h[slot_of(e, k)];
          float x = in[((size_t)c * M.nel_s + e) * kNpp + k];
          if (!seen[g]) { h[g] = x; seen[g] = 1; }
          else if (h[g] != x) fail(AXH_ERR_ARG, std::string(what) + ": copies of a point differ");
        }
    store(ddst, h, F.B, slot);
  };
  if (disp) gather_s(disp, F.disp, "disp");
  if (velo) gather_s(velo, F.velo, "velo");
  if (M.nel_f > 0 && (chi || dchi || ddchi)) {
    auto gather_f = [&](const float *in, float *ddst, const char *what) {
      std::vector<float> h(M.nglob_f, 0.0f);
      std::vector<char> seen(h.size(), 0);
      for (int e = 0; e < M.nel_f; e++)
        for (int k = 0; k < kNpp; k++) {
          size_t g = M.gid_f_h[slot_of(e, k)];
          float x = in[(size_t)e * kNpp + k];
          if (!seen[g]) { h[g] = x; seen[g] = 1; }
          else if (h[g] != x) fail(AXH_ERR_ARG, std::string(what) + ": copies of a point differ");
        }
      store(ddst, h, F.B, slot);
    };
    if (chi) gather_f(chi, F.chi, "chi");
    if (dchi) gather_f(dchi, F.dchi, "dchi");
    if (ddchi) gather_f(ddchi, F.ddchi, "ddchi");
  }
}

// ---------------------------------------------------------------------------
// resident meshes and footprints
// ---------------------------------------------------------------------------
void footprint(axh_ctx *ctx, size_t *mesh_bytes, size_t *wave_bytes) {
  const axh_dims &d = ctx->dims;
  const size_t nel_s = d.nel_solid, nel_f = d.nel_fluid;
  const size_t ns = (size_t)((nel_s + kElemPerTile - 1) / kElemPerTile) * kTile;
  const size_t nf = (size_t)((nel_f + kElemPerTile - 1) / kElemPerTile) * kTile;
  const size_t gs = d.nglob_solid, gf = d.nglob_fluid;
  const size_t nc = d.src_type == AXH_MONOPOLE ? kNcMono : d.src_type == AXH_DIPOLE ? kNcDi
                                                                                 : kNcQuad;
  size_t nax = 0;
  if (const Array *a = ctx->get("axis_solid"))
    for (size_t e = 0; e < a->count(); e++) nax += a->as<int32_t>()[e] != 0;
  size_t m = 105 * 4;
  m += ns * nc * 4 + nel_s * 4 + nax * 50 * 4 + ns * 4 + (gs + 1) * 4 + nel_s * 25 * 4 + gs * 4;
  if (nel_f) m += nf * 4 * 4 + nel_f * 4 + nel_f * 5 * 4 + nf * 4 + (gf + 1) * 4 + nel_f * 25 * 4 +
                  gf * 4 + gf;
  if (d.nel_bdry) m += (size_t)d.nel_bdry * 5 * (5 * 4 + 8) + kMaxAxisBdry * 4 + (ns + nf) * 4;
  if (d.anel) m += nel_s * (kCgStage + kCgUpd) * 16 + 2 * nel_s * d.n_sls * 8;
  size_t w = 3 * gs * 4 * 3 + 3 * ns * 4 + 3 * kMaxAxisBdry * 4 + gs + gs * 4;
  if (nel_f) w += 3 * gf * 4 + nf * 4 + nel_f * 4 + 200 * 4;
  if (d.anel) w += nel_s * (24 * d.n_sls + 24 + 4 + 24) * 4;
  w += (size_t)std::max(d.nelsrc, 0) * 25 * 16 + (size_t)d.num_rec * 4 + 4;
  if (const Array *a = ctx->get("stf")) w += a->count() * 4;   // device stf table
  if (d.num_rec > 0) w += (size_t)seis_staging_samples(d) * 3 * d.num_rec * 4;
  *mesh_bytes = m;
  *wave_bytes = w;
}

}  // namespace axh

// a handle on a device mesh part
struct axh_mesh {
  std::shared_ptr<axh::MeshDev> mesh;
  uint64_t fp[2];
};

namespace axh {

axh_mesh *mesh_get(axh_ctx *ctx) {
  uint64_t fp[2];
  mesh_fingerprint(ctx, fp);
  if (!ctx->dev) device_build(ctx);
  std::unique_ptr<axh_mesh> m(new axh_mesh);
  m->mesh = ctx->dev->mesh;
  m->fp[0] = fp[0];
  m->fp[1] = fp[1];
  return m.release();
}

void mesh_bind(axh_ctx *ctx, const axh_mesh *m) {
  if (ctx->dev) fail(AXH_ERR_STATE, "the context is already on the device");
  uint64_t fp[2];
  mesh_fingerprint(ctx, fp);
  if (fp[0] != m->fp[0] || fp[1] != m->fp[1])
    fail(AXH_ERR_ARG, "the context's mesh terms differ from the mesh's");
  ctx->bound_mesh = m->mesh;
}

size_t mesh_bytes(const axh_mesh *m) { return m->mesh->mem.bytes; }

void mesh_release(axh_mesh *m) { delete m; }

}  // namespace axh
