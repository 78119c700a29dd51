// mcpm_noalloc.hip — the energy and momentum check for callers that cannot
// allocate device memory (code under stream capture, tight loops, harnesses
// that forbid hipMalloc).
//
// mcpm_energy_momentum_noalloc runs the kernels of mcpm_energy_momentum on
// caller-owned FP64 velocity arrays.  Its only scratch, the per-workgroup
// partial sums, lives in the file-scope device array `scratch` below: 4
// doubles for each of the ENERGY_GRID workgroups (32 KiB), whatever the
// particle count, held by each GPU the library runs on until the process
// ends.  Host code never passes `scratch` itself to a kernel (the name is a
// host-side stand-in, not a device address): on a device's first call the
// array's address on that device is looked up once with
// hipGetSymbolAddress and kept per device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mcpm.h"
#include "mcpm_kernels.h"

using namespace mcpm;

namespace {

constexpr int NOALLOC_SCRATCH_DOUBLES = 4 * ENERGY_GRID;
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

std::mutex g_mu;
double *g_ptr[MAX_DEVICES] = {nullptr};

// Device address of `scratch` on the current device (looked up once).
int scratch_ptr(double **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return MCPM_ERR_HIP;
  if (dev < 0 || dev >= MAX_DEVICES) return MCPM_ERR_HIP;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MCPM_ERR_HIP;
    g_ptr[dev] = (double *)p;
  }
  *out = g_ptr[dev];
  return MCPM_OK;
}

}  // namespace

extern "C" {

int mcpm_energy_momentum_noalloc(const double *u1, const double *u2, const double *u3,
                                 int64_t n, double mass_fn, double *out, void *stream) {
  if (n < 0 || !out) return MCPM_ERR_ARG;
  if (n > 0 && (!u1 || !u2 || !u3)) return MCPM_ERR_ARG;
  double *part = nullptr;
  int st = scratch_ptr(&part);
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  // every workgroup writes its four partials (zeros when its slice is
  // empty), so nothing of the previous call's contents is read
  hipLaunchKernelGGL(k_energy_soa, dim3(ENERGY_GRID), dim3(256), 0, s, u1, u2, u3, n, part);
  if (hipGetLastError() != hipSuccess) return MCPM_ERR_HIP;
  hipLaunchKernelGGL(k_energy_finish, dim3(1), dim3(64), 0, s, part, ENERGY_GRID, mass_fn, out);
  if (hipGetLastError() != hipSuccess) return MCPM_ERR_HIP;
  return MCPM_OK;
}

}  // extern "C"
