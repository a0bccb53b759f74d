// axh_meshdb.cpp -- reader of the reference MESHER's mesh database.
//
// The database (Mesh/meshdb.dat0000) is a Fortran unformatted sequential file
// written by MESHER/pdb.f90:2213-2386 and read by the SOLVER in
// SOLVER/get_mesh.f90:47-409 (read_mesh_basics / read_mesh_advanced /
// read_mesh_axel of SOLVER/data_mesh.f90:195-328).  Each record is a 4-byte
// little-endian length, the payload and the length again.  This reader walks
// the records in that order, so a request loop or a test can obtain the
// mesh-only inputs of the time loop (numbering, solid/fluid element lists,
// axis and solid-fluid boundary elements, the spectral operators) without the
// Fortran host.  Only single-slice databases (nproc_mesh = 1) are read: the
// GPU loop keeps the whole mesh on one device.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "axh_internal.h"

using namespace axh;

struct axh_meshdb {
  axh_meshdb_info info{};
  std::vector<double> xi_k, eta, dxi, wt, wt_axial_k;
  std::vector<float> G0, G1, G1T, G2, G2T;  // realkind 4 databases
  std::vector<double> crd_s, crd_z;         // control points, scaled by router
  std::vector<int32_t> lnods;               // [nelem][8]
  std::vector<char> eltype;                 // [nelem][6]
  std::vector<int32_t> coarsing;            // LOGICAL (0/1)
  std::vector<int32_t> ielsolid, ielfluid, igloc_solid, igloc_fluid;
  std::vector<int32_t> bdry_solid_el, bdry_fluid_el, bdry_jpol_solid, bdry_jpol_fluid;
  std::vector<int32_t> ax_el, ax_el_solid, ax_el_fluid;
  std::vector<int32_t> axis_solid, axis_fluid;  // per element 0/1
};

namespace {

struct Records {
  FILE *f = nullptr;
  std::string path;
  explicit Records(const std::string &p) : path(p) {
    f = std::fopen(p.c_str(), "rb");
    if (!f) fail(AXH_ERR_IO, "cannot open mesh database " + p);
  }
  ~Records() {
    if (f) std::fclose(f);
  }
  // the next record's payload
  std::vector<char> next(const char *what) {
    uint32_t n = 0, m = 0;
    if (std::fread(&n, 4, 1, f) != 1) fail(AXH_ERR_IO, path + ": truncated before " + what);
    std::vector<char> b(n);
    if (n && std::fread(b.data(), 1, n, f) != n) fail(AXH_ERR_IO, path + ": truncated " + what);
    if (std::fread(&m, 4, 1, f) != 1 || m != n)
      fail(AXH_ERR_IO, path + ": bad record markers at " + what);
    return b;
  }
  int32_t i32(const char *what) {
    auto b = next(what);
    if (b.size() != 4) fail(AXH_ERR_IO, path + ": " + what + " is not one integer");
    int32_t v;
    std::memcpy(&v, b.data(), 4);
    return v;
  }
  template <class T> std::vector<T> vec(const char *what, size_t n) {
    auto b = next(what);
    if (b.size() < n * sizeof(T))
      fail(AXH_ERR_IO, path + ": record " + what + " shorter than expected");
    std::vector<T> v(n);
    if (n) std::memcpy(v.data(), b.data(), n * sizeof(T));
    return v;
  }
  template <class T> void fields(const char *what, std::initializer_list<T *> out) {
    auto b = next(what);
    if (b.size() != out.size() * sizeof(T)) fail(AXH_ERR_IO, path + ": record " + what);
    size_t k = 0;
    for (T *p : out) std::memcpy(p, b.data() + sizeof(T) * k++, sizeof(T));
  }
};

}  // namespace

extern "C" {

int axh_meshdb_open(const char *path, axh_meshdb **out) {
  return guarded([&] {
    if (!path || !out) fail(AXH_ERR_ARG, "null argument");
    *out = nullptr;
    std::unique_ptr<axh_meshdb> db(new axh_meshdb);
    axh_meshdb_info &I = db->info;
    Records R(path);
    // read_mesh_basics (SOLVER/data_mesh.f90:195-213)
    I.nproc = R.i32("nproc_mesh");
    I.npol = R.i32("npol");
    I.nelem = R.i32("nelem");
    I.npoint = R.i32("npoint");
    I.nel_solid = R.i32("nel_solid");
    I.nel_fluid = R.i32("nel_fluid");
    const int npoint_solid = R.i32("npoint_solid");
    const int npoint_fluid = R.i32("npoint_fluid");
    I.nglob_solid = R.i32("nglob_solid");
    I.nglob_fluid = R.i32("nglob_fluid");
    I.nel_bdry = R.i32("nel_bdry");
    I.ndisc = R.i32("ndisc");
    const int lfbkgrd = R.i32("lfbkgrdmodel");
    if (I.nproc != 1) fail(AXH_ERR_UNSUPPORTED, "mesh database of more than one slice");
    if (I.npol != 4) fail(AXH_ERR_UNSUPPORTED, "npol /= 4");
    const int np = I.npol + 1;
    if (npoint_solid != I.nel_solid * np * np || npoint_fluid != I.nel_fluid * np * np ||
        I.npoint != I.nelem * np * np || I.nelem != I.nel_solid + I.nel_fluid)
      fail(AXH_ERR_IO, "inconsistent element counts in the mesh database");
    // read_mesh_advanced (SOLVER/data_mesh.f90:217-306)
    db->xi_k = R.vec<double>("xi_k", np);
    db->eta = R.vec<double>("eta", np);
    db->dxi = R.vec<double>("dxi", np);
    db->wt = R.vec<double>("wt", np);
    db->wt_axial_k = R.vec<double>("wt_axial_k", np);
    {
      // G0 is realkind: its record length tells the precision
      auto g0 = R.next("G0");
      I.realkind_bytes = (int)(g0.size() / np);
      if (I.realkind_bytes != 4)
        fail(AXH_ERR_UNSUPPORTED, "mesh database written for realkind /= 4");
      db->G0.resize(np);
      std::memcpy(db->G0.data(), g0.data(), 4 * np);
    }
    db->G1 = R.vec<float>("G1", np * np);
    db->G1T = R.vec<float>("G1T", np * np);
    db->G2 = R.vec<float>("G2", np * np);
    db->G2T = R.vec<float>("G2T", np * np);
    I.npoin = R.i32("npoin");
    db->crd_s = R.vec<double>("crd_nodes(:,1)", I.npoin);
    db->crd_z = R.vec<double>("crd_nodes(:,2)", I.npoin);
    db->lnods.resize((size_t)I.nelem * 8);
    for (int e = 0; e < I.nelem; e++) {
      auto v = R.vec<int32_t>("lnods", 8);
      std::memcpy(&db->lnods[(size_t)e * 8], v.data(), 32);
    }
    I.nglob = R.i32("nglob");
    db->eltype = R.vec<char>("eltype", (size_t)I.nelem * 6);
    db->coarsing = R.vec<int32_t>("coarsing", I.nelem);
    db->ielsolid = R.vec<int32_t>("ielsolid", I.nel_solid);
    db->ielfluid = R.vec<int32_t>("ielfluid", I.nel_fluid);
    // the records hold all npoint entries; the solid/fluid ones come first
    db->igloc_solid = R.vec<int32_t>("igloc_solid", npoint_solid);
    db->igloc_fluid = R.vec<int32_t>("igloc_fluid", npoint_fluid);
    I.have_bdry = R.i32("have_bdry_elem") != 0;
    if (I.have_bdry) {
      db->bdry_solid_el = R.vec<int32_t>("bdry_solid_el", I.nel_bdry);
      db->bdry_fluid_el = R.vec<int32_t>("bdry_fluid_el", I.nel_bdry);
      db->bdry_jpol_solid = R.vec<int32_t>("bdry_jpol_solid", I.nel_bdry);
      db->bdry_jpol_fluid = R.vec<int32_t>("bdry_jpol_fluid", I.nel_bdry);
    }
    // read_db (SOLVER/get_mesh.f90:101-219)
    {
      auto b = R.next("pts_wavelngth,period,courant,deltat");
      if (b.size() != 32) fail(AXH_ERR_IO, "numerical parameters record");
      std::memcpy(&I.pts_wavelngth, b.data(), 8);
      std::memcpy(&I.period, b.data() + 8, 8);
      std::memcpy(&I.courant, b.data() + 16, 8);
      std::memcpy(&I.deltat, b.data() + 24, 8);
    }
    {
      auto b = R.next("bkgrdmodel");
      const size_t n = std::min<size_t>({b.size(), (size_t)lfbkgrd, sizeof(I.bkgrdmodel) - 1});
      std::memcpy(I.bkgrdmodel, b.data(), n);
      I.bkgrdmodel[n] = 0;
    }
    (void)R.next("override_ext_q");
    {
      auto b = R.next("router,have_fluid");
      if (b.size() != 12) fail(AXH_ERR_IO, "router record");
      std::memcpy(&I.router, b.data(), 8);
      int32_t hf;
      std::memcpy(&hf, b.data() + 8, 4);
      I.have_fluid = hf != 0;
    }
    for (int d = 0; d < I.ndisc; d++) (void)R.next("discont");
    (void)R.next("rmin,minh_ic,maxh_ic,maxh_icb");
    (void)R.next("hmin_glob,hmax_glob");
    (void)R.next("min_distance");
    for (int k = 0; k < 4; k++) (void)R.next("char_time");
    R.fields<int32_t>("naxel", {&I.naxel, &I.naxel_solid, &I.naxel_fluid});
    // read_mesh_axel (SOLVER/data_mesh.f90:312-326)
    db->ax_el = R.vec<int32_t>("ax_el", I.naxel);
    db->ax_el_solid = R.vec<int32_t>("ax_el_solid", I.naxel_solid);
    db->ax_el_fluid = R.vec<int32_t>("ax_el_fluid", I.naxel_fluid);
    // (the message-passing records of a single slice are empty)
    db->axis_solid.assign(I.nel_solid, 0);
    db->axis_fluid.assign(I.nel_fluid, 0);
    for (int32_t e : db->ax_el_solid) {
      if (e < 1 || e > I.nel_solid) fail(AXH_ERR_IO, "ax_el_solid out of range");
      db->axis_solid[e - 1] = 1;
    }
    for (int32_t e : db->ax_el_fluid) {
      if (e < 1 || e > I.nel_fluid) fail(AXH_ERR_IO, "ax_el_fluid out of range");
      db->axis_fluid[e - 1] = 1;
    }
    for (int32_t g : db->igloc_solid)
      if (g < 1 || g > I.nglob_solid) fail(AXH_ERR_IO, "igloc_solid out of range");
    for (int32_t g : db->igloc_fluid)
      if (g < 1 || g > I.nglob_fluid) fail(AXH_ERR_IO, "igloc_fluid out of range");
    *out = db.release();
  });
}

int axh_meshdb_close(axh_meshdb *db) {
  delete db;
  return AXH_OK;
}

int axh_meshdb_get_info(const axh_meshdb *db, axh_meshdb_info *info) {
  return guarded([&] {
    if (!db || !info) fail(AXH_ERR_ARG, "null argument");
    *info = db->info;
  });
}

int axh_meshdb_array(const axh_meshdb *db, const char *name, void *out, size_t nbytes) {
  return guarded([&] {
    if (!db || !name || !out) fail(AXH_ERR_ARG, "null argument");
    const std::string n(name);
    const void *src = nullptr;
    size_t len = 0;
    auto pick = [&](const auto &v) {
      src = v.data();
      len = v.size() * sizeof(v[0]);
    };
    if (n == "igloc_solid") pick(db->igloc_solid);
    else if (n == "igloc_fluid") pick(db->igloc_fluid);
    else if (n == "ielsolid") pick(db->ielsolid);
    else if (n == "ielfluid") pick(db->ielfluid);
    else if (n == "axis_solid") pick(db->axis_solid);
    else if (n == "axis_fluid") pick(db->axis_fluid);
    else if (n == "ax_el_solid") pick(db->ax_el_solid);
    else if (n == "ax_el_fluid") pick(db->ax_el_fluid);
    else if (n == "bdry_solid_el") pick(db->bdry_solid_el);
    else if (n == "bdry_fluid_el") pick(db->bdry_fluid_el);
    else if (n == "bdry_jpol_solid") pick(db->bdry_jpol_solid);
    else if (n == "bdry_jpol_fluid") pick(db->bdry_jpol_fluid);
    else if (n == "G0") pick(db->G0);
    else if (n == "G1") pick(db->G1);
    else if (n == "G1T") pick(db->G1T);
    else if (n == "G2") pick(db->G2);
    else if (n == "G2T") pick(db->G2T);
    else if (n == "crd_s") pick(db->crd_s);
    else if (n == "crd_z") pick(db->crd_z);
    else if (n == "lnods") pick(db->lnods);
    else if (n == "coarsing") pick(db->coarsing);
    else if (n == "eltype") pick(db->eltype);
    else fail(AXH_ERR_ARG, "unknown mesh database array " + n);
    if (nbytes != len)
      fail(AXH_ERR_ARG, n + ": " + std::to_string(len) + " bytes, not " + std::to_string(nbytes));
    if (len) std::memcpy(out, src, len);
  });
}

// The mesh-only uploads of a context from the database: the spectral
// operators and the numbering and axis maps.  The precomputed terms, the
// masses and the solid-fluid coupling matrix come from the reference's setup.
int axh_meshdb_upload(const axh_meshdb *db, axh_ctx *ctx) {
  const axh_meshdb_info &I = db ? db->info : axh_meshdb_info{};
  int rc = guarded([&] {
    if (!db || !ctx) fail(AXH_ERR_ARG, "null argument");
    const axh_dims &d = ctx->dims;
    if (d.nel_solid != I.nel_solid || d.nel_fluid != I.nel_fluid ||
        d.nglob_solid != I.nglob_solid || d.nglob_fluid != I.nglob_fluid ||
        d.nel_bdry != I.nel_bdry)
      fail(AXH_ERR_ARG, "the context's dims are not the database's");
  });
  if (rc != AXH_OK) return rc;
  rc = axh_upload_spectral(ctx, db->G0.data(), db->G1.data(), db->G1T.data(), db->G2.data(),
                           db->G2T.data());
  if (rc != AXH_OK) return rc;
  return axh_upload_mesh_maps(ctx, db->igloc_solid.data(), db->igloc_fluid.data(),
                              db->axis_solid.data(), db->axis_fluid.data());
}

}  // extern "C"
