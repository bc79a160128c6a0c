// grains_gpu_force_sums_noalloc: the per-particle sums of contact-force
// magnitudes of grains_gpu_force_sums, for callers that must not allocate
// device memory (code under HIP graph capture, tight serving loops).  The
// per-storage-index sums live in a file-scope device array of fixed size
// instead of a buffer the library allocates on first use; the result goes
// to the caller's device buffer in id order.  No counterpart in the
// reference.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/grains_gpu.h"
#include "dem_launch.hpp"
#include "dem_system.hpp"

namespace {

// Largest system the static scratch serves (doubles, one per particle:
// 32 MiB held on every GPU the library runs on until the process ends).
constexpr int kNoAllocMaxParticles = 1 << 22;
static __device__ __align__(256) double scratch[kNoAllocMaxParticles];

// Device address of `scratch` on each GPU, looked up once per device with
// that device current (the array's host-side name is not a device address).
constexpr int kMaxDevices = 64;
double* scratch_ptr[kMaxDevices] = {};
std::mutex scratch_mutex;

int fail(grains_gpu_system* s, int code, const std::string& msg) {
  if (s) s->last_error = msg;
  return code;
}

// out[id - 1] = sums[s] over storage indices s
__global__ __launch_bounds__(256) void k_sums_by_id(const double* __restrict__ sums, const int* __restrict__ id, int n,
                                                    double* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = id[s] - 1;
  if ((unsigned)i < (unsigned)n) out[i] = sums[s];
}

}  // namespace

int grains_gpu_force_sums_noalloc(grains_gpu_system* s, double* d_out) {
  if (!s || !s->initialized || !d_out) return GRAINS_ARGUMENT_ERROR;
  const int n = s->nparticles;
  if (n > kNoAllocMaxParticles)
    return fail(s, GRAINS_ARGUMENT_ERROR,
                "grains_gpu_force_sums_noalloc: " + std::to_string(n) + " particles exceed the static scratch of " +
                    std::to_string(kNoAllocMaxParticles));
  if (s->device < 0 || s->device >= kMaxDevices)
    return fail(s, GRAINS_ARGUMENT_ERROR, "grains_gpu_force_sums_noalloc: device index out of range");
  hipError_t e = hipSetDevice(s->device);
  if (e != hipSuccess) return fail(s, GRAINS_DEVICE_ERROR, std::string("hipSetDevice: ") + hipGetErrorString(e));
  double* ptr;
  {
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[s->device]) {
      double* p = nullptr;
      e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return fail(s, GRAINS_DEVICE_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      scratch_ptr[s->device] = p;
    }
    ptr = scratch_ptr[s->device];
  }
  if (n == 0) return GRAINS_OK;
  // every particle's sum is written by the evaluation: nothing to clear
  s->fsum_on = true;
  s->fsum_buf = ptr;
  int code = s->forces(dem::kModeForcesOnly, 0, false);
  s->fsum_on = false;
  s->fsum_buf = nullptr;
  if (code) return code;
  hipLaunchKernelGGL(k_sums_by_id, dim3((n + 255) / 256), dim3(256), 0, s->stream, ptr, s->P.id, n, d_out);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(s, GRAINS_DEVICE_ERROR, std::string("k_sums_by_id: ") + hipGetErrorString(e));
  return s->check_status();
}
