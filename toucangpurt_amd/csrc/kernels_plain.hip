// Plain cycle kernel: the reference's execution model written for gfx950.
//
// One workgroup of 1024 threads per partition visit, partitions strided over
// the grid per region (src/SimEval.cu:889-965), one on-disk 32-lane
// micropartition per wave pulled from a workgroup work queue
// (SimEval.cu:805-844), the LUT read from global memory, the reference device
// layout (SimEval.cu:69-148,1143-1281), two grid barriers per cycle through
// the cooperative-groups grid sync, and one instance per launch. It is the
// baseline the wave64 and batched kernels are measured against, so it keeps
// the reference's structure rather than the MI355X-specific choices.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include "kernel_args.h"

namespace cg = cooperative_groups;

namespace tgs {

namespace {

template <typename T>
__device__ __forceinline__ T ld(const uint8_t *p) {
  return *reinterpret_cast<const T *>(p);
}

__device__ __forceinline__ uint32_t readOp(const uint8_t *vp, uint32_t s) { return s < 16 ? s : vp[s]; }

__device__ __forceinline__ uint32_t lutG(const uint8_t *lut, uint32_t pos) { return pos < kLutSize ? lut[pos] : 0u; }

// evalSingleMicroPart (SimEval.cu:304-580). Lanes 32..63 of the wave only
// take part in the shuffles.
__device__ void plainMicroPart(const PlainArgs &a, const uint8_t *mp, uint8_t *vp, const uint8_t *consts) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t magic = ld<uint32_t>(mp);
  const uint32_t second = ld<uint32_t>(mp + 4);
  if (magic == kMagicLut) {
    const uint8_t *sizes = mp + 8;
    const uint8_t *ptr = mp + kPlainHeaderBytes;
    uint32_t n = sizes[0];
    uint32_t val = 0;
    if (lane < n) {
      const uint8_t *op = ptr + 8 * lane;
      uint32_t lut = ld<uint16_t>(op), o0 = ld<uint16_t>(op + 2), o1 = ld<uint16_t>(op + 4), o2 = ld<uint16_t>(op + 6);
      uint32_t v0 = readOp(vp, o0), v1 = readOp(vp, o1), v2 = readOp(vp, o2);
      val = lut == 0 ? v2 : lutG(a.c.lut, lutIndex(lut, v0, v1, v2));
    }
    ptr += 8 * n;
    for (uint32_t l = 0; l < second; ++l) {
      n = sizes[1 + l];
      uint32_t w = lane < n ? ld<uint32_t>(ptr + 4 * lane) : 0u;
      uint32_t o2 = w & 63u, o1 = (w >> 6) & 63u, o0 = (w >> 12) & 63u, lut = w >> 18;
      uint32_t s0 = __shfl(val, int(o0 & 31u));
      uint32_t s1 = __shfl(val, int(o1 & 31u));
      uint32_t s2 = __shfl(val, int(o2 & 31u));
      uint32_t v0 = o0 < 32 ? o0 : s0, v1 = o1 < 32 ? o1 : s1, v2 = o2 < 32 ? o2 : s2;
      val = lut == 0 ? (v2 & 0xFFu) : lutG(a.c.lut, lutIndex(lut, v0, v1, v2));
      ptr += 4 * n;
    }
    n = sizes[1 + second];
    uint32_t sid = 0, res = 0;
    if (lane < n) {
      sid = ptr[4 * lane];
      res = ld<uint16_t>(ptr + 4 * lane + 2);
    }
    uint32_t r = __shfl(val, int(sid & 31u));
    if (sid < 32) r = sid;
    if (lane < n) vp[res] = uint8_t(r);
  } else if (magic == kMagicVecRead) {
    if (lane < second) {
      const uint8_t *op = mp + 8 + 20 * lane;
      uint32_t base = ld<uint16_t>(op), len = ld<uint16_t>(op + 2), offset = ld<uint16_t>(op + 4), isConst = op[6];
      uint32_t i0 = readOp(vp, ld<uint16_t>(op + 8)), i1 = readOp(vp, ld<uint16_t>(op + 10));
      uint32_t i2 = readOp(vp, ld<uint16_t>(op + 12)), i3 = readOp(vp, ld<uint16_t>(op + 14));
      uint32_t off = ((i0 << 12) | (i1 << 8) | (i2 << 4) | i3) + offset;
      uint32_t v;
      if (off < len) v = isConst ? consts[base + off] : vp[base + off];
      else v = readOp(vp, ld<uint16_t>(op + 16));
      vp[ld<uint16_t>(op + 18)] = uint8_t(v);
    }
  } else if (magic == kMagicVecOp) {
    if (lane < second) {
      const uint8_t *op = mp + 8 + 10 * lane;
      uint32_t b1 = ld<uint16_t>(op), b2 = ld<uint16_t>(op + 2), len = ld<uint16_t>(op + 4);
      uint32_t name = op[6], cf = op[7], res = ld<uint16_t>(op + 8);
      unsigned __int128 v1 = 0, v2 = 0;
      for (uint32_t k = 0; k < len; ++k) {
        unsigned __int128 s1 = (cf & 2) ? consts[b1 + k] : vp[b1 + k];
        unsigned __int128 s2 = (cf & 1) ? consts[b2 + k] : vp[b2 + k];
        v1 |= s1 << (4 * k);
        v2 |= s2 << (4 * k);
      }
      if (len > 0 && len < 32) {  // sign_extend_128 (SimEval.cu:222-228)
        uint32_t bit = 4 * len - 1;
        unsigned __int128 ext = ~((((unsigned __int128)1) << (bit + 1)) - 1);
        if ((v1 >> bit) & 1) v1 |= ext;
        if ((v2 >> bit) & 1) v2 |= ext;
      }
      __int128 s1 = (__int128)v1, s2 = (__int128)v2;
      if (name <= 2) {
        unsigned __int128 r = name == 0 ? v1 + v2 : name == 1 ? v1 - v2 : v1 * v2;
        for (uint32_t k = 0; k < len; ++k) {
          vp[uint16_t(res + k)] = uint8_t(r & 15u);
          r >>= 4;
        }
      } else {
        uint32_t v = name == 3 ? (s1 == s2) : name == 4 ? (s1 < s2) : name == 5 ? (s1 <= s2) : 0u;
        vp[res] = uint8_t(v);
      }
    }
  } else if (magic == kMagicMemRead) {
    if (lane < second) {
      const uint8_t *op = mp + 8 + 24 * lane;
      uint32_t multi = op[0];
      uint64_t base = ld<uint64_t>(op + 8);
      uint32_t en = readOp(vp, ld<uint16_t>(op + 16)), av = ld<uint16_t>(op + 18), res = ld<uint16_t>(op + 20);
      if (en != 0) {
        uint64_t at = memAddress(vp, av, multi, base);
        vp[res] = at < a.c.memSize ? a.c.mem[at] : uint8_t(0);
      }
    }
  }
}

// evalEachPartition (SimEval.cu:765-864).
__device__ void plainPartition(const PlainArgs &a, uint32_t pid, uint8_t *vp, const uint64_t *draws, uint64_t cycle,
                               uint32_t *queue) {
  const PlainPart P = a.parts[pid];
  const uint8_t *net = a.net + P.net;
  const uint8_t *consts = a.c.consts + P.constBase;
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  for (uint32_t i = tid; i < P.nRegRead; i += nt) {
    const uint8_t *r = net + P.offRegRead + 8 * i;
    vp[ld<uint16_t>(r + 4)] = a.c.reg[ld<uint32_t>(r)];
  }
  for (uint32_t i = tid; i < P.nExgRead; i += nt) {
    const uint8_t *r = net + P.offExgRead + 8 * i;
    vp[ld<uint16_t>(r + 4)] = a.c.exg[ld<uint32_t>(r)];
  }
  for (uint32_t i = tid; i < P.nStimRead; i += nt) {
    const uint8_t *r = net + P.offStimRead + 8 * i;
    vp[ld<uint16_t>(r + 4)] = uint8_t(stimBit(draws, ld<uint32_t>(r)));
  }
  __syncthreads();
  const uint32_t lane = tid & 63u;
  for (uint32_t l = 0; l < P.nLevels; ++l) {
    const uint32_t nM = a.levels[2 * (P.levelBase + l)];
    const uint32_t mBase = a.levels[2 * (P.levelBase + l) + 1];
    if (tid == 0) *queue = 0;
    __syncthreads();
    while (true) {
      uint32_t id = 0;
      if (lane == 0) id = atomicAdd(queue, 1u);
      id = __shfl(id, 0);
      if (id >= nM) break;
      plainMicroPart(a, net + a.mparts[mBase + id], vp, consts);
    }
    __syncthreads();
  }
  spillProbes(a.c, pid, vp, tid, nt);
  // evalLastLevel / evalLastLevelExgWrite (SimEval.cu:625-761)
  if (P.hasExgWrite) {
    const uint8_t *r = net + P.offExgWrite;
    uint32_t exg = ld<uint32_t>(r), dat = ld<uint16_t>(r + 4), cnt = ld<uint16_t>(r + 6);
    uint32_t words = (cnt + 3) / 4;
    for (uint32_t i = tid; i < words; i += nt)
      reinterpret_cast<uint32_t *>(a.c.exg)[(exg >> 2) + i] = reinterpret_cast<const uint32_t *>(vp)[(dat >> 2) + i];
  } else {
    if (P.hasRegWrite) {
      const uint8_t *r = net + P.offRegWrite;
      uint32_t reg = ld<uint32_t>(r), dat = ld<uint16_t>(r + 4), cnt = ld<uint16_t>(r + 6);
      uint32_t words = (cnt + 3) / 4;
      for (uint32_t i = tid; i < words; i += nt)
        reinterpret_cast<uint32_t *>(a.c.reg)[(reg >> 2) + i] = reinterpret_cast<const uint32_t *>(vp)[(dat >> 2) + i];
    }
    for (uint32_t i = tid; i < P.nMemWrite; i += nt) {
      const uint8_t *r = net + P.offMemWrite + 24 * i;
      uint32_t en = readOp(vp, ld<uint16_t>(r + 20));
      if (en != 0) {
        uint32_t dat = readOp(vp, ld<uint16_t>(r + 18));
        uint64_t at = memAddress(vp, ld<uint16_t>(r + 16), r[0], ld<uint64_t>(r + 8));
        if (at < a.c.memSize) a.c.mem[at] = uint8_t(dat);
      }
    }
    if (a.c.printEnable)
      for (uint32_t i = tid; i < P.nPrint; i += nt) {
        const uint8_t *r = net + P.offPrint + 4 * i;
        if (readOp(vp, ld<uint16_t>(r)) != 0) logPrint(a.c, cycle, ld<uint16_t>(r + 2));
      }
    for (uint32_t i = tid; i < P.nStop; i += nt) {
      const uint8_t *r = net + P.offStop + 2 * i;
      if (readOp(vp, ld<uint16_t>(r)) != 0) raiseStop(a.c, cycle);
    }
  }
  __syncthreads();
}

}  // namespace

// evalFreeRunningNCycles (SimEval.cu:889-995), plus the stimulus stage.
__global__ __launch_bounds__(1024) void plainCycles(PlainArgs a) {
  extern __shared__ __align__(16) uint8_t plainVp[];
  __shared__ uint64_t draws[kMaxDraws];
  __shared__ uint32_t queue;
  cg::grid_group grid = cg::this_grid();
  uint64_t xs = *a.c.xs;
  const bool lead = threadIdx.x == 0;
  for (uint32_t c = 0; c < a.c.nCycles; ++c) {
    const uint64_t cycle = a.c.cycleBase + c;
    const bool pc = profCycle(a.c, c) && lead;  // ENABLE_SIM_PROFILE (SimEval.cu:894-983)
    if (threadIdx.x == 0) stimulusDraws(xs, a.c.nIn, draws);
    __syncthreads();
    if (blockIdx.x == 0)
      for (uint32_t i = threadIdx.x; i < a.c.nIn; i += blockDim.x) a.c.reg[a.c.inIds[i]] = uint8_t(stimBit(draws, i));
    uint64_t t0 = pc ? shaderClock() : 0;
    for (uint32_t p = blockIdx.x; p < a.n0; p += gridDim.x) {
      plainPartition(a, p, plainVp, draws, cycle, &queue);
      if (pc) {
        a.c.profUseful[p] += shaderClock() - t0;
        a.c.profOwner[p] = blockIdx.x;
      }
    }
    grid.sync();
    if (pc) {
      a.c.profWg[2 * blockIdx.x] += shaderClock() - t0;
      t0 = shaderClock();
    }
    for (uint32_t p = a.n0 + blockIdx.x; p < a.nTotal; p += gridDim.x) {
      plainPartition(a, p, plainVp, draws, cycle, &queue);
      if (pc) {
        a.c.profUseful[p] += shaderClock() - t0;
        a.c.profOwner[p] = blockIdx.x;
      }
    }
    grid.sync();
    if (pc) a.c.profWg[2 * blockIdx.x + 1] += shaderClock() - t0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.c.xs = xs;
}

// Host launcher: cooperative launch on `stream`.
hipError_t launchPlain(const PlainArgs &args, uint32_t grid, uint32_t block, uint32_t lds, hipStream_t stream) {
  PlainArgs a = args;
  void *params[] = {&a};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(plainCycles), dim3(grid), dim3(block), params, lds,
                                    stream);
}

hipError_t plainMaxBlocksPerCU(int *blocks, uint32_t block, uint32_t lds) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, reinterpret_cast<const void *>(plainCycles), block, lds);
}

hipError_t plainSetLds(uint32_t lds) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(plainCycles), hipFuncAttributeMaxDynamicSharedMemorySize,
                             lds);
}

}  // namespace tgs
