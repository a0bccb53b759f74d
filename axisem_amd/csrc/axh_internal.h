// axh_internal.h -- shared declarations of the GPU time-loop library.
#pragma once

#include <cstdint>
#include <exception>
#include <map>
#include <new>
#include <memory>
#include <string>
#include <vector>

#include "axihip.h"

namespace axh {

constexpr int kNpol = 4;
constexpr int kNpp = 25;           // GLL points per element
constexpr int kElemPerTile = 10;   // elements per 256-lane workgroup
constexpr int kTile = 256;         // slots per tile (250 used)
constexpr int kMaxSls = 8;
constexpr int kMaxBatch = AXH_MAX_BATCH;  // wavefields per kernel launch

enum DType : uint32_t { kF32 = 1, kF64 = 2, kI32 = 3 };

// One uploaded array, kept in the reference layout.
struct Array {
  DType dtype = kF32;
  std::vector<uint64_t> dims;
  std::vector<char> bytes;
  size_t count() const {
    size_t n = 1;
    for (auto d : dims) n *= d;
    return n;
  }
  template <class T> const T *as() const {
    return reinterpret_cast<const T *>(bytes.data());
  }
};

struct Device;  // device-side layout and buffers (axh_device.hip)
struct DeviceDeleter {
  void operator()(Device *d) const;
};
struct MeshDev;     // device mesh part, shared (axh_device.hip)
struct BatchState;  // device groups of a batch (axh_device.hip)
struct BatchStateDeleter {
  void operator()(BatchState *b) const;
};

// Error raised inside the library; converted to a status code at the ABI.
struct Error {
  int code;
  std::string msg;
};
[[noreturn]] void fail(int code, const std::string &msg);

// Runs f, converting a failure into its status code (the message goes to
// axh_last_error): no exception crosses the C-ABI.
int set_error(int code, const std::string &msg);
template <class F> int guarded(F &&f) {
  try {
    f();
    return AXH_OK;
  } catch (const Error &e) {
    return set_error(e.code, e.msg);
  } catch (const std::bad_alloc &) {
    return set_error(AXH_ERR_ARG, "host allocation failed");
  } catch (const std::exception &e) {
    return set_error(AXH_ERR_ARG, e.what());
  } catch (...) {
    return set_error(AXH_ERR_ARG, "unknown error");
  }
}

}  // namespace axh

struct axh_ctx {
  axh_dims dims{};
  std::map<std::string, axh::Array> arrays;
  std::unique_ptr<axh::Device, axh::DeviceDeleter> dev;
  int32_t iter_done = 0;  // last completed step
  double last_run_ms = 0.0;
  bool profile = false;
  int graph_steps = 0;    // steps per replayed graph (0: direct launches)
  bool running = false;   // between axh_run_begin and axh_run_end
  // kernel-wavefield snapshots of axh_run_kwf (axh_set_kwf)
  int kwf_strain_it = 0;
  int kwf_src_mask = 1;
  int kwf_ring = 0;       // device ring capacity in snapshots (0: automatic)
  int batch_refs = 0;     // batches holding this context
  // device mesh part to use when the device part is built (axh_mesh_bind)
  std::shared_ptr<axh::MeshDev> bound_mesh;
  // cached fingerprint of the mesh part (cleared by every upload)
  bool fp_valid = false;
  uint64_t fp[2] = {0, 0};

  const axh::Array *get(const std::string &name) const {
    auto it = arrays.find(name);
    return it == arrays.end() ? nullptr : &it->second;
  }
  const axh::Array &need(const std::string &name) const;
  void put(const std::string &name, axh::DType t, std::vector<uint64_t> dims,
           const void *src);
};

// Wavefields on one mesh stepped together (axh_batch_*).  The members are not
// owned; while a batch holds a context, the context takes no uploads and
// cannot be destroyed; from the first batch run until the batch is destroyed
// the context's state lives in the batch's device groups.
struct axh_batch {
  std::vector<axh_ctx *> members;
  std::unique_ptr<axh::BatchState, axh::BatchStateDeleter> state;  // NULL before the first run
  double last_run_ms = 0.0;
  bool profile = false;
  int graph_steps = 0;              // steps per replayed graph (0: direct)
  int group_width = AXH_MAX_GROUP;  // wavefields per kernel instance
};

namespace axh {
// names of the solid 3-D terms in enum axh_solid_term order
extern const char *const kSolid3dNames[AXH_N_SOLID_3D];
extern const char *const kFluid3dNames[4];
// inv_rho_fluid and the fluid's derivative terms over J (axh_upload_fluid_gradient)
extern const char *const kFluidGradNames[5];

// device lifecycle (axh_device.hip)
void device_build(axh_ctx *ctx);
void device_destroy(Device *d);
void device_run(axh_ctx *ctx, int32_t it0, int32_t it1, float *seis);
void device_run_begin(axh_ctx *ctx, int32_t it0, int32_t it1, float *seis);
void device_run_end(axh_ctx *ctx);
void device_run_noalloc(axh_ctx *ctx, int32_t it0, int32_t it1, float *seis);
void device_run_kwf(axh_ctx *ctx, int32_t it0, int32_t it1, float *seis, float *kwf);
void device_run_kwf_begin(axh_ctx *ctx, int32_t it0, int32_t it1, float *seis, float *kwf);
void convolve_traces(int nt, int ntraces, const float *seis, int nj, const float *stf,
                     double dt, double scale, float *out);
int noalloc_max_receivers();
void device_download(axh_ctx *ctx, float *disp, float *velo, float *chi,
                     float *dchi, float *ddchi);
void device_upload_fields(axh_ctx *ctx, const float *disp, const float *velo,
                          const float *chi, const float *dchi,
                          const float *ddchi);
size_t device_bytes(const Device *d);
const char *kernel_class_name(int k);
void check_inputs(const axh_ctx *ctx);
void global_numbering(const axh_ctx *ctx, int domain, int32_t *gid, int32_t *nglob);
bool kernel_stats(const Device *d, int k, double *ms, int64_t *n);
// batches: put every member on the first member's device mesh (building what
// is missing) and move their state into the batch's groups; step them
// together; move the state back
void batch_attach(axh_batch *b);
void batch_detach(axh_batch *b);
bool in_running_batch(const axh_ctx *ctx);
void mesh_fingerprint(axh_ctx *ctx, uint64_t fp[2]);
void batch_key(const axh_ctx *ctx, uint64_t key[2]);
// device bytes of the mesh part and of one wavefield (B = 1) without a GPU
void footprint(axh_ctx *ctx, size_t *mesh_bytes, size_t *wave_bytes);
axh_mesh *mesh_get(axh_ctx *ctx);
void mesh_bind(axh_ctx *ctx, const axh_mesh *m);
size_t mesh_bytes(const axh_mesh *m);
void mesh_release(axh_mesh *m);
// advance every member nsteps from its own last completed step
void batch_run(axh_batch *b, int32_t nsteps, float *const *seis, float *const *kwf = nullptr);
bool batch_kernel_stats(const axh_batch *b, int k, double *ms, int64_t *n);
int symplectic_coefficients(int scheme, double dt, double *cd, double *cv,
                            double *cf);
}  // namespace axh
