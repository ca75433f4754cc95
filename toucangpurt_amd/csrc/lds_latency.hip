// Latency probe for the dependent LDS chains that make up one exec level of
// the streamed kernel: one workgroup of 704 threads (the streamed kernel's
// 8 compute + 3 loader waves), wave 0 runs a dependent chain while the others
// wait, or all waves run workgroup barriers. Times are s_memtime ticks (the
// shader clock), converted to ns with s_memrealtime (100 MHz) over the same
// interval. Usage: lds_latency [iterations]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

constexpr uint32_t kWords = 8192;  // 32 KB table
enum Mode : uint32_t { M_U8 = 0, M_BPERM, M_TOP, M_MID, M_BARRIER, M_HANDOFF, M_COUNT };
const char *kNames[M_COUNT] = {
    "ds_read_u8 chain (value-pool operand read)",
    "ds_bpermute_b32 chain (middle-level operand)",
    "top op: 3 byte reads + LUT byte read",
    "middle op: 3 bpermutes + LUT byte read",
    "workgroup barrier, 11 waves",
    "byte write + barrier + byte read (level hand-off)",
};

__global__ void chain(uint32_t mode, uint32_t iters, const uint32_t *init, uint64_t *out) {
  __shared__ uint32_t tab[kWords];
  for (uint32_t k = threadIdx.x; k < kWords; k += blockDim.x) tab[k] = init[k];
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint8_t *tb = reinterpret_cast<const uint8_t *>(tab);
  uint8_t *tw = reinterpret_cast<uint8_t *>(tab);
  uint32_t v = init[threadIdx.x] & 4095u;
  const uint64_t t0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime();
  if (mode == M_BARRIER) {
    for (uint32_t i = 0; i < iters; ++i) __syncthreads();
  } else if (mode == M_HANDOFF) {
    // wave w writes byte (i, w), then every wave reads the byte its
    // neighbour wrote; the read feeds the next write
    const uint32_t nw = blockDim.x >> 6;
    for (uint32_t i = 0; i < iters; ++i) {
      const uint32_t row = (i & 7u) * 1024u;
      tw[row + wave * 64u + lane] = uint8_t(v);
      __syncthreads();
      v = tb[row + ((wave + 1) % nw) * 64u + (lane ^ (v & 1u))] + 1u;
    }
  } else if (wave == 0) {
    for (uint32_t i = 0; i < iters; ++i) {
      if (mode == M_U8) {
        v = tb[(v * 131u + lane) & 4095u];
      } else if (mode == M_BPERM) {
        v = uint32_t(__builtin_amdgcn_ds_bpermute(int((v & 63u) << 2), int(v + i)));
      } else if (mode == M_TOP) {
        const uint32_t a = tb[(v + lane) & 4095u], b = tb[(v * 3u + lane) & 4095u], c = tb[(v * 5u + lane) & 4095u];
        v = tb[(((a << 8) | (b << 4) | c) + 8192u) & 16383u];
      } else {
        const uint32_t a = uint32_t(__builtin_amdgcn_ds_bpermute(int((v & 31u) << 2), int(v)));
        const uint32_t b = uint32_t(__builtin_amdgcn_ds_bpermute(int(((v >> 5) & 31u) << 2), int(v)));
        const uint32_t c = uint32_t(__builtin_amdgcn_ds_bpermute(int(((v >> 10) & 31u) << 2), int(v)));
        v = tb[(((a << 8) | (b << 4) | c) + 8192u) & 16383u] * 1025u + lane;
      }
    }
  }
  const uint64_t t1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = t1 - t0;
    out[1] = r1 - r0;
  }
  if (v == 0xFFFFFFFFu) out[2] = v;  // keeps the chain live
}

}  // namespace

int main(int argc, char **argv) {
  const uint32_t iters = argc > 1 ? uint32_t(std::atoi(argv[1])) : 4000;
  std::vector<uint32_t> init(kWords);
  uint32_t x = 12345u;
  for (auto &w : init) {
    x = x * 1664525u + 1013904223u;
    w = x;
  }
  uint32_t *dInit = nullptr;
  uint64_t *dOut = nullptr;
  if (hipMalloc(&dInit, kWords * 4) != hipSuccess || hipMalloc(&dOut, 64) != hipSuccess) return 1;
  if (hipMemcpy(dInit, init.data(), kWords * 4, hipMemcpyHostToDevice) != hipSuccess) return 1;
  for (uint32_t m = 0; m < M_COUNT; ++m) {
    for (int rep = 0; rep < 2; ++rep) {  // the first launch warms up
      hipLaunchKernelGGL(chain, dim3(1), dim3(704), 0, 0, m, iters, dInit, dOut);
      if (hipDeviceSynchronize() != hipSuccess) return 1;
    }
    uint64_t o[2] = {0, 0};
    if (hipMemcpy(o, dOut, 16, hipMemcpyDeviceToHost) != hipSuccess) return 1;
    const double ticks = double(o[0]) / iters, ns = double(o[1]) * 10.0 / iters;
    std::printf("%-52s %7.1f ticks %7.1f ns per step (clock %.2f GHz)\n", kNames[m], ticks, ns,
                double(o[0]) / (double(o[1]) * 10.0));
  }
  (void)hipFree(dInit);
  (void)hipFree(dOut);
  return 0;
}
