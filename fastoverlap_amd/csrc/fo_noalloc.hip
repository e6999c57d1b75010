// Allocation-free spherical alignment: fo_sph_align_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) pass device-resident
// positions and outputs; every intermediate of the alignment (descriptors,
// harmonics, radial kernels, coefficients, plane candidates, peaks, the
// assignment solver's arrays) is laid out inside one file-scope device array
// of kScratchBytes.  The layout is the one fo_sph_batch_create allocates, the
// kernels are the same, so the results are identical to that path's.
//
// Each device holds its own copy of the array from the moment the library's
// code object is loaded until the process ends.  Its device address is
// resolved once per device (fo_context_create -> noalloc_attach, with the
// device current) and only that address is used on the host.  Every call on a
// device shares the array: calls on one device are serialised by a lock, and
// each call clears what must start at zero (the batch binding zeroes the FP32
// coefficient padding and uploads the descriptors).  Overlapping use of the
// same device's array from outside this entry point would race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "fastoverlap_hip.h"
#include "fo_host.h"
#include "fo_internal.h"

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

// host code must never pass this name to a kernel or a copy: it is a
// host-side stand-in; the device address is resolved per device below
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_attach_mu;
void* g_ptr[kMaxDevices] = {};
hipError_t g_attach_err[kMaxDevices] = {};
std::mutex g_call_mu[kMaxDevices];

// dist / rotation matrix / orientation of each pair into the caller's arrays
__global__ void k_unpack_results(int P, const fo::SphResult* __restrict__ res,
                                 double* __restrict__ dist, double* __restrict__ rmat,
                                 int* __restrict__ inverted) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  if (dist) dist[p] = res[p].dist;
  if (rmat)
    for (int k = 0; k < 9; ++k) rmat[9 * p + k] = res[p].rmat[k];
  if (inverted) inverted[p] = res[p].inverted;
}

}  // namespace

namespace fo {

void noalloc_attach(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> g(g_attach_mu);
  if (g_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess && p) {
    g_ptr[device] = p;
    g_attach_err[device] = hipSuccess;
  } else {
    (void)hipGetLastError();
    g_attach_err[device] = e == hipSuccess ? hipErrorInvalidSymbol : e;
  }
}

}  // namespace fo

extern "C" {

int fo_sph_align_noalloc(fo_context* ctx, const fo_sph_batch_desc* desc, const double* d_pos1,
                         const double* d_pos2, double* d_dist, double* d_X2, int* d_perm,
                         double* d_rmat, int* d_inverted) {
  if (!ctx || !desc || !d_pos1 || !d_pos2)
    return fo::set_error(FO_ERR_INVALID, "NULL argument");
  if (desc->flags & (FO_SPH_KEEP_GRID | FO_SPH_GENERAL))
    return fo::set_error(FO_ERR_INVALID,
                         "fo_sph_align_noalloc aligns positions only (no KEEP_GRID / GENERAL)");
  const int dev = fo::ctx_device(ctx);
  if (dev < 0 || dev >= kMaxDevices) return fo::set_error(FO_ERR_UNSUPPORTED, "device id too large");
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> g(g_attach_mu);
    base = g_ptr[dev];
    if (!base) {
      const std::string m = std::string("allocation-free scratch not resolved on this device: ") +
                            hipGetErrorString(g_attach_err[dev]);
      return fo::set_error(FO_ERR_HIP, m.c_str());
    }
  }
  std::lock_guard<std::mutex> serial(g_call_mu[dev]);
  if (hipSetDevice(dev) != hipSuccess) return fo::set_error(FO_ERR_HIP, "hipSetDevice failed");
  fo_sph_batch* b = nullptr;
  int rc = fo::sph_batch_plan(ctx, desc, &b, /*build_tables=*/false);
  if (rc) return rc;
  const size_t need = fo::sph_batch_bytes(b);
  if (need > kScratchBytes) {
    fo_sph_batch_destroy(b);
    const std::string m = "alignment needs " + std::to_string(need) +
                          " bytes of scratch; the allocation-free entry point holds " +
                          std::to_string(kScratchBytes) + " (use fo_sph_batch_create)";
    return fo::set_error(FO_ERR_NOMEM, m.c_str());
  }
  hipStream_t s = fo::ctx_stream(ctx);
  rc = fo::sph_batch_bind(b, base, s);
  if (!rc) rc = fo_sph_batch_set_positions_device(b, d_pos1, d_pos2);
  if (!rc) rc = fo_sph_batch_run(b, FO_STAGE_ALL, 0);
  if (!rc) {
    fo::SphOutputs o;
    fo::sph_batch_outputs(b, &o);
    hipLaunchKernelGGL(k_unpack_results, dim3((o.npairs + 63) / 64), dim3(64), 0, s, o.npairs,
                       o.res, d_dist, d_rmat, d_inverted);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && d_X2)
      e = hipMemcpyAsync(d_X2, o.X2, sizeof(double) * 3 * (size_t)o.natoms,
                         hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess && d_perm)
      e = hipMemcpyAsync(d_perm, o.perm, sizeof(int) * (size_t)o.natoms, hipMemcpyDeviceToDevice,
                         s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      const std::string m = std::string("fo_sph_align_noalloc: ") + hipGetErrorString(e);
      rc = fo::set_error(FO_ERR_HIP, m.c_str());
    }
  }
  fo_sph_batch_destroy(b);  // host-side plan only: the scratch stays with the device
  return rc;
}

}  // extern "C"
