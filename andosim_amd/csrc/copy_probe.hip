// Calibration program for HBM traffic counters: a plain float copy of known
// size (read N*4 bytes, write N*4 bytes per launch).  Profiling FETCH_SIZE and
// WRITE_SIZE on it gives the scale that turns those counters into bytes for
// the cloth-step kernels.
//
//   copy_probe [MiB per buffer, default 64] [launches, default 5]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(size_t n, const float4* __restrict__ a, float4* __restrict__ b) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

int main(int argc, char** argv) {
  size_t mib = argc > 1 ? std::strtoul(argv[1], nullptr, 10) : 64;
  int launches = argc > 2 ? std::atoi(argv[2]) : 5;
  size_t bytes = mib << 20;
  size_t n4 = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  if (hipMemset(a, 0, bytes) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float total = 0.0f;
  for (int k = 0; k < launches; ++k) {
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, n4, a, b);
    (void)hipEventRecord(e1, 0);
    (void)hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (k) total += ms;
  }
  if (hipGetLastError() != hipSuccess) return 1;
  double gbs = launches > 1 ? 2.0 * bytes / (total / (launches - 1) * 1e-3) / 1e9 : 0.0;
  std::printf("copy_probe: %zu bytes read + %zu bytes written per launch, %d launches, %.1f GB/s\n", bytes, bytes,
              launches, gbs);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
