// axh_post.hip -- seismogram post-processing on the GPU.
//
// The reference's post-processing program (SOLVER/UTILS/post_processing.F90)
// sums the moment-tensor runs with their radiation factors, convolves each
// trace with a source time function and rotates the components.  The
// convolution is the costly part: for every trace, every one of nt samples
// sums nj = 2 shift_fact1 t0 / dt terms (convolve_with_stf,
// SOLVER/UTILS/post_processing.F90:1014-1084), tens of millions of terms per
// trace at 5 s and 1800 s.  Here one GPU thread owns one output sample of one
// trace and runs its terms in the reference's order and precision, so the
// result is the reference's bit for bit; a workgroup's 256 consecutive samples
// read a sliding window of the trace that stays in the caches.
#include <hip/hip_runtime.h>

#include <string>

#include "axh_internal.h"

#pragma clang fp contract(off)

namespace axh {

#define HIPCHECK_POST(x)                                                     \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess)                                                    \
      fail(AXH_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));     \
  } while (0)

// out(i) for sample i = 1 + blockIdx.x*256 + threadIdx.x of trace blockIdx.y:
// seis_fil(i) = seis_fil(i) + seis(i-j) * stf(j) * dt over j = 1..nj with
// j < i, the product of two REAL(4) values scaled by a REAL(8) dt and added
// in REAL(8) to the REAL(4) partial sum, then seis_fil * scale in REAL(8).
// The terms of one sample run in ascending j, as in the reference loop.
__global__ __launch_bounds__(256) void k_convolve(int nt, int nj, const float *__restrict__ seis,
                                                  const float *__restrict__ stf, double dt,
                                                  double scale, float *__restrict__ out) {
  extern __shared__ float s_stf[];
  for (int j = threadIdx.x; j < nj; j += 256) s_stf[j] = stf[j];
  __syncthreads();
  const int i = 1 + blockIdx.x * 256 + threadIdx.x;   // 1-based sample
  if (i > nt) return;
  const float *tr = seis + (size_t)blockIdx.y * nt;
  float acc = 0.0f;
  const int jmax = min(nj, i - 1);
  for (int j = 1; j <= jmax; j++) {
    const float p = tr[i - j - 1] * s_stf[j - 1];
    acc = (float)((double)acc + (double)p * dt);
  }
  out[(size_t)blockIdx.y * nt + i - 1] = (float)((double)acc * scale);
}

void convolve_traces(int nt, int ntraces, const float *seis, int nj, const float *stf,
                     double dt, double scale, float *out) {
  if (nt <= 0 || ntraces <= 0) return;
  const size_t n = (size_t)nt * ntraces;
  float *d_seis = nullptr, *d_out = nullptr, *d_stf = nullptr;
  hipStream_t s = nullptr;
  auto release = [&] {
    if (s) (void)hipStreamSynchronize(s);
    if (d_seis) (void)hipFree(d_seis);
    if (d_out) (void)hipFree(d_out);
    if (d_stf) (void)hipFree(d_stf);
    if (s) (void)hipStreamDestroy(s);
  };
  try {
    HIPCHECK_POST(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIPCHECK_POST(hipMalloc(&d_seis, n * sizeof(float)));
    HIPCHECK_POST(hipMalloc(&d_out, n * sizeof(float)));
    HIPCHECK_POST(hipMalloc(&d_stf, (size_t)std::max(nj, 1) * sizeof(float)));
    HIPCHECK_POST(hipMemcpyAsync(d_seis, seis, n * sizeof(float), hipMemcpyHostToDevice, s));
    if (nj > 0)
      HIPCHECK_POST(hipMemcpyAsync(d_stf, stf, (size_t)nj * sizeof(float),
                                   hipMemcpyHostToDevice, s));
    const dim3 grid((nt + 255) / 256, ntraces);
    hipLaunchKernelGGL(k_convolve, grid, dim3(256), (size_t)std::max(nj, 1) * sizeof(float), s,
                       nt, nj, d_seis, d_stf, dt, scale, d_out);
    HIPCHECK_POST(hipGetLastError());
    HIPCHECK_POST(hipMemcpyAsync(out, d_out, n * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHECK_POST(hipStreamSynchronize(s));
  } catch (...) {
    release();
    throw;
  }
  release();
}

}  // namespace axh
