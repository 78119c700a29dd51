// Random-access rates of the device, measured on arrays shaped like the
// collision schedule's: particle records, one 4-byte priority word per
// particle and one mark byte per 2^MARK_SHIFT particles.  The reservation
// rounds are bound by
// these rates (every access of an operation goes to an unpredictable
// particle), not by streaming bandwidth, so the rates are the ceiling
// their timings are set against.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mcpm.h"
#include "mcpm_kernels.h"

namespace mcpm {
namespace {

// Index of access t: a 32-bit mix of t, reduced to [0, n).
__device__ __forceinline__ uint32_t probe_index(uint32_t t, uint32_t n) {
  uint32_t x = t * 2654435761u + 0x9e3779b9u;
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return (uint32_t)(((uint64_t)x * n) >> 32);
}

// Gathers of whole records (w uint4 each).
__global__ void k_probe_gather_rec(const uint4 *__restrict__ a, uint32_t n, int w, uint32_t nops,
                                   uint4 *__restrict__ sink) {
  uint4 acc = make_uint4(0u, 0u, 0u, 0u);
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nops;
       t += gridDim.x * blockDim.x) {
    const uint4 *r = a + (size_t)probe_index(t, n) * w;
    for (int q = 0; q < w; ++q) {
      const uint4 v = r[q];
      acc.x ^= v.x;
      acc.y ^= v.y;
      acc.z ^= v.z;
      acc.w ^= v.w;
    }
  }
  // never true for the fill pattern; keeps the loads live
  if (acc.x == 0x5a5a5a5au && acc.y == 0xa5a5a5a5u) sink[0] = acc;
}

// Gathers of one 4-byte word of n.
__global__ void k_probe_gather_word(const uint32_t *__restrict__ a, uint32_t n, uint32_t nops,
                                    uint32_t *__restrict__ sink) {
  uint32_t acc = 0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nops;
       t += gridDim.x * blockDim.x)
    acc ^= a[probe_index(t, n)];
  if (acc == 0x5a5a5a5au) sink[0] = acc;
}

// 32-bit atomic minimum into one word of n, result unused.
__global__ void k_probe_atomic_word(uint32_t *__restrict__ a, uint32_t n, uint32_t nops) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nops;
       t += gridDim.x * blockDim.x)
    atomicMin(&a[probe_index(t, n)], t);
}

// Stores of one 4-byte word of n.
__global__ void k_probe_scatter_word(uint32_t *__restrict__ a, uint32_t n, uint32_t nops) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nops;
       t += gridDim.x * blockDim.x)
    a[probe_index(t, n)] = t;
}

// Gathers of one byte of n.
__global__ void k_probe_gather_byte(const uint8_t *__restrict__ a, uint32_t n, uint32_t nops,
                                    uint32_t *__restrict__ sink) {
  uint32_t acc = 0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nops;
       t += gridDim.x * blockDim.x)
    acc += a[probe_index(t, n)];
  if (acc == 0x5a5a5a5au) sink[0] = acc;
}

// Stores of whole records.
__global__ void k_probe_scatter_rec(uint4 *__restrict__ a, uint32_t n, int w, uint32_t nops) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nops;
       t += gridDim.x * blockDim.x) {
    uint4 *r = a + (size_t)probe_index(t, n) * w;
    for (int q = 0; q < w; ++q) r[q] = make_uint4(t, t, t, t);
  }
}

}  // namespace
}  // namespace mcpm

using namespace mcpm;

extern "C" int mcpm_probe_random_access(int64_t n_records, int32_t record_bytes,
                                        mcpm_access_rates *out) {
  if (!out || n_records < 1024 || n_records > (int64_t)1 << 30 ||
      (record_bytes != 16 && record_bytes != 32))
    return MCPM_ERR_ARG;
  const uint32_t n = (uint32_t)n_records, w = (uint32_t)record_bytes / 16u;
  const uint32_t nops = n < (1u << 24) ? n : (1u << 24);
  const uint32_t nb = n >> MARK_SHIFT;  // mark bytes
  uint4 *rec = nullptr, *sink = nullptr;
  uint32_t *words = nullptr;
  uint8_t *bytes = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = MCPM_OK;
  constexpr int NKIND = 6;
  double best[NKIND] = {1e30, 1e30, 1e30, 1e30, 1e30, 1e30};
  if (hipMalloc((void **)&rec, (size_t)n * w * sizeof(uint4)) != hipSuccess ||
      hipMalloc((void **)&words, (size_t)n * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void **)&bytes, (size_t)nb) != hipSuccess ||
      hipMalloc((void **)&sink, 2 * sizeof(uint4)) != hipSuccess ||
      hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    st = MCPM_ERR_HIP;
    goto done;
  }
  if (hipMemsetAsync(rec, 0x11, (size_t)n * w * sizeof(uint4), s) != hipSuccess ||
      hipMemsetAsync(words, 0xFF, (size_t)n * sizeof(uint32_t), s) != hipSuccess ||
      hipMemsetAsync(bytes, 0, (size_t)nb, s) != hipSuccess) {
    st = MCPM_ERR_HIP;
    goto done;
  }
  // best of five launches of each kind, 8192 workgroups of 256
  for (int rep = 0; rep < 5 && st == MCPM_OK; ++rep) {
    for (int kind = 0; kind < NKIND; ++kind) {
      hipEventRecord(e0, s);
      if (kind == 0)
        hipLaunchKernelGGL(k_probe_gather_rec, dim3(8192), dim3(256), 0, s, rec, n, (int)w, nops,
                           sink);
      else if (kind == 1)
        hipLaunchKernelGGL(k_probe_gather_word, dim3(8192), dim3(256), 0, s, words, n, nops,
                           (uint32_t *)sink);
      else if (kind == 2)
        hipLaunchKernelGGL(k_probe_atomic_word, dim3(8192), dim3(256), 0, s, words, n, nops);
      else if (kind == 3)
        hipLaunchKernelGGL(k_probe_scatter_rec, dim3(8192), dim3(256), 0, s, rec, n, (int)w, nops);
      else if (kind == 4)
        hipLaunchKernelGGL(k_probe_scatter_word, dim3(8192), dim3(256), 0, s, words, n, nops);
      else
        hipLaunchKernelGGL(k_probe_gather_byte, dim3(8192), dim3(256), 0, s, bytes, nb, nops,
                           (uint32_t *)sink);
      hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) {
        st = MCPM_ERR_HIP;
        break;
      }
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      if (ms > 0.f && ms < best[kind]) best[kind] = ms;
    }
  }
  if (st == MCPM_OK) {
    out->gather_record = nops / (best[0] * 1e-3);
    out->gather_word = nops / (best[1] * 1e-3);
    out->atomic_min_word = nops / (best[2] * 1e-3);
    out->scatter_record = nops / (best[3] * 1e-3);
    out->scatter_word = nops / (best[4] * 1e-3);
    out->gather_byte = nops / (best[5] * 1e-3);
  }
done:
  if (s) hipStreamSynchronize(s);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (s) hipStreamDestroy(s);
  if (rec) hipFree(rec);
  if (words) hipFree(words);
  if (bytes) hipFree(bytes);
  if (sink) hipFree(sink);
  return st;
}
