// Streaming-copy probe for the roofline harness (no reference counterpart):
// the best 1-read + 1-write HBM rate a plain kernel reaches on this device,
// the practical ceiling the bench reports beside the 8 TB/s HBM peak.
#include "fs2_internal.h"

namespace fs2 {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

// Grid-stride float4 copy: at any moment the grid reads one contiguous
// window of grid x 256 x 16 bytes and writes another.
__global__ void __launch_bounds__(256) k_copy4(const f4* __restrict__ a, f4* __restrict__ b,
                                               int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

}  // namespace

double copy_probe_ms(uint64_t bytes, int grid, int reps) {
  const int64_t n = (int64_t)(bytes / 16);
  if (n <= 0 || grid <= 0 || reps <= 0) fail(FS2_ERR_VALUE, "copy probe: bad arguments");
  void* a = dmalloc((size_t)n * 16);
  void* b = nullptr;
  try {
    b = dmalloc((size_t)n * 16);
  } catch (...) {
    dfree(a);
    throw;
  }
  hipStream_t s = library_stream();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double best = 1e30;
  hipError_t err = hipMemsetAsync(a, 0, (size_t)n * 16, s);
  if (err == hipSuccess) err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  for (int r = 0; r <= reps && err == hipSuccess; ++r) {   // r = 0 warms up
    hipLaunchKernelGGL(k_copy4, dim3((unsigned)grid), dim3(256), 0, s, (const f4*)a, (f4*)b, n);
    err = hipGetLastError();
    if (err == hipSuccess) err = hipEventRecord(e0, s);
    if (err == hipSuccess) hipLaunchKernelGGL(k_copy4, dim3((unsigned)grid), dim3(256), 0, s,
                                              (const f4*)a, (f4*)b, n);
    if (err == hipSuccess) err = hipGetLastError();
    if (err == hipSuccess) err = hipEventRecord(e1, s);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    if (err == hipSuccess && r > 0) best = std::min(best, (double)ms);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  dfree(a);
  dfree(b);
  if (err != hipSuccess) fail(FS2_ERR_DEVICE, "copy probe: %s", hipGetErrorString(err));
  return best;
}

}  // namespace fs2
