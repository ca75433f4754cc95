// Calibration probe for HBM traffic counters: copies a buffer of known size
// with a plain vectorised kernel so that rocprofv3's FETCH_SIZE / WRITE_SIZE
// per dispatch can be scaled to bytes. Usage: hbm_copy [MiB] [repeats]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copyKernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

int main(int argc, char **argv) {
  size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
  int reps = argc > 2 ? std::atoi(argv[2]) : 5;
  size_t bytes = mib << 20, n = bytes / sizeof(uint4);
  uint4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (int r = 0; r < reps; ++r) {
    (void)hipEventRecord(e0);
    copyKernel<<<4096, 256>>>(a, b, n);
    (void)hipEventRecord(e1);
    if (hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    std::printf("copy %zu MiB: %.3f ms, %.1f GB/s (read+write)\n", mib, ms, 2.0 * bytes / (ms * 1e6));
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
