// Batched cycle kernel for gfx950: many stimulus instances of one design in
// one persistent launch, one instance per lane.
//
// Semantics per instance are the reference evaluator's (src/SimEval.cu:304-761,
// :765-864, :889-965). The execution is laid out for 64-wide waves: every
// pool is instance-minor (nibble k of instance i at k * instPad + i), so a
// wave of 64 instances moves 64 contiguous bytes per op, program words are
// uniform (scalar loads shared by all 64 instances), and the value pool of a
// (partition, 64-instance group) work item sits in LDS as [slot][lane].
// A LUT micropartition's level values live in a per-wave LDS buffer
// [2][32][64]; a lane index past the previous level's size reads 0, as the
// reference's idle lanes do (SimEval.cu:372-375,406-408).
//
// Instances run uneven cycle counts: the host orders instance slots by
// decreasing cycle count, so the groups still active in a cycle are a prefix
// and finished groups cost nothing; lanes past their own count never write.
// Stimulus draws for cycle c+1 are produced during region 1 of cycle c into
// the other half of a double buffer, so no extra grid barrier is needed.
#include <hip/hip_runtime.h>

#include "kernel_args.h"

namespace tgs {

namespace {

constexpr uint32_t kLutBytesB = ((kLutSize + 15) / 16) * 16;

struct Ctx {
  const BatchArgs &a;
  uint8_t *vp;       // [slot][64]
  uint8_t *lvl;      // this wave's [2][32][64]
  const uint8_t *lut;
  uint32_t lane, g, slot0;  // group g, first instance slot of the group
  uint64_t cycle;
};

__device__ __forceinline__ uint32_t vpAt(const Ctx &c, uint32_t s) { return c.vp[s * 64u + c.lane]; }
__device__ __forceinline__ uint32_t opnd(const Ctx &c, uint32_t s) { return s < 16 ? s : vpAt(c, s); }
__device__ __forceinline__ void vpPut(const Ctx &c, uint32_t s, uint32_t v) { c.vp[s * 64u + c.lane] = uint8_t(v); }
__device__ __forceinline__ uint32_t lutB(const Ctx &c, uint32_t pos) { return pos < kLutSize ? c.lut[pos] : 0u; }

// One on-disk micropartition (half of a T_LUT2 task), ops sequential, lanes = instances.
__device__ __forceinline__ void lutHalf(const Ctx &c, const uint32_t *S, const uint32_t *top, const uint32_t *ops, uint32_t nTop,
                        uint32_t nMid, uint32_t half, const uint32_t *last, uint32_t nLast) {
  uint8_t *L = c.lvl;
  for (uint32_t i = 0; i < nTop; ++i) {
    const uint32_t x = top[2 * i], y = top[2 * i + 1];
    const uint32_t lt = x & 0xFFFFu, o0 = x >> 16, o1 = y & 0xFFFFu, o2 = y >> 16;
    const uint32_t cc = opnd(c, o2);
    uint32_t v = cc;
    if (lt != 0) v = lutB(c, lutIndex(lt, opnd(c, o0), opnd(c, o1), cc));
    L[i * 64u + c.lane] = uint8_t(v);
  }
  uint32_t nPrev = nTop, cur = 0;
  for (uint32_t m = 0; m < nMid; ++m) {
    const uint32_t sz = (S[m >> 1] >> ((m & 1u) << 4)) & 0xFFFFu;
    const uint32_t na = sz & 0xFFu, nb = sz >> 8;
    const uint32_t n = half ? nb : na;
    const uint32_t *w = ops + (half ? na : 0u);
    const uint8_t *P = L + cur * 2048u;
    uint8_t *Q = L + (cur ^ 1u) * 2048u;
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t op = w[i];
      const uint32_t o2 = op & 63u, o1 = (op >> 6) & 63u, o0 = (op >> 12) & 63u, lt = op >> 18;
      const uint32_t a = o0 < 32 ? o0 : ((o0 & 31u) < nPrev ? P[(o0 & 31u) * 64u + c.lane] : 0u);
      const uint32_t b = o1 < 32 ? o1 : ((o1 & 31u) < nPrev ? P[(o1 & 31u) * 64u + c.lane] : 0u);
      const uint32_t d = o2 < 32 ? o2 : ((o2 & 31u) < nPrev ? P[(o2 & 31u) * 64u + c.lane] : 0u);
      Q[i * 64u + c.lane] = uint8_t(lt == 0 ? d : lutB(c, lutIndex(lt, a, b, d)));
    }
    ops += na + nb;
    nPrev = n;
    cur ^= 1u;
  }
  const uint8_t *P = L + cur * 2048u;
  for (uint32_t i = 0; i < nLast; ++i) {
    const uint32_t w = last[i];
    const uint32_t sid = (w >> 16) & 0xFFu;
    const uint32_t v = sid < 32 ? sid : ((sid & 31u) < nPrev ? P[(sid & 31u) * 64u + c.lane] : 0u);
    vpPut(c, w & 0xFFFFu, v);
  }
}

__device__ __forceinline__ void lutTask(const Ctx &c, const uint32_t *T) {
  const uint32_t w0 = T[0], w1 = T[1];
  const uint32_t nMid = (w0 >> 8) & 0xFFFFu;
  const uint32_t ntA = w1 & 0xFFu, ntB = (w1 >> 8) & 0xFFu, nlA = (w1 >> 16) & 0xFFu, nlB = w1 >> 24;
  const uint32_t *S = T + 2;
  const uint32_t *top = S + ((nMid + 1) >> 1);
  const uint32_t *mids = top + 2 * (ntA + ntB);
  uint32_t midWords = 0;
  for (uint32_t m = 0; m < nMid; ++m) {
    const uint32_t sz = (S[m >> 1] >> ((m & 1u) << 4)) & 0xFFFFu;
    midWords += (sz & 0xFFu) + (sz >> 8);
  }
  const uint32_t *last = mids + midWords;
  lutHalf(c, S, top, mids, ntA, nMid, 0, last, nlA);
  if (ntB | nlB) lutHalf(c, S, top + 2 * ntA, mids, ntB, nMid, 1, last + nlA, nlB);
}

__device__ __forceinline__ void vecReadB(const Ctx &c, const uint32_t *T, const uint8_t *consts) {
  // Ops are evaluated one after another; the loader rejects programs in which
  // one op of an exec level reads a slot another op of that level writes, so
  // the order is unobservable.
  const uint32_t n = T[0] >> 8;
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t *o = T + 1 + 5 * k;
    const uint32_t vb = o[0] & 0xFFFFu, len = o[0] >> 16, offset = o[1] & 0xFFFFu, isConst = o[1] >> 16;
    const uint32_t off = ((opnd(c, o[2] & 0xFFFFu) << 12) | (opnd(c, o[2] >> 16) << 8) |
                          (opnd(c, o[3] & 0xFFFFu) << 4) | opnd(c, o[3] >> 16)) + offset;
    const uint32_t v = off < len ? (isConst ? consts[vb + off] : vpAt(c, vb + off)) : opnd(c, o[4] & 0xFFFFu);
    vpPut(c, o[4] >> 16, v);
  }
}

__device__ __forceinline__ void vecOpB(const Ctx &c, const uint32_t *T, const uint8_t *consts) {
  const uint32_t n = T[0] >> 8;
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t *o = T + 1 + 3 * k;
    const uint32_t b1 = o[0] & 0xFFFFu, b2 = o[0] >> 16;
    const uint32_t len = o[1] & 0xFFFFu, name = (o[1] >> 16) & 0xFFu, cf = o[1] >> 24, res = o[2];
    unsigned __int128 v1 = 0, v2 = 0;
    for (uint32_t j = 0; j < len; ++j) {
      unsigned __int128 s1 = (cf & 2) ? consts[b1 + j] : vpAt(c, b1 + j);
      unsigned __int128 s2 = (cf & 1) ? consts[b2 + j] : vpAt(c, b2 + j);
      v1 |= s1 << (4 * j);
      v2 |= s2 << (4 * j);
    }
    if (len > 0 && len < 32) {
      const uint32_t bit = 4 * len - 1;
      const unsigned __int128 ext = ~((((unsigned __int128)1) << (bit + 1)) - 1);
      if ((v1 >> bit) & 1) v1 |= ext;
      if ((v2 >> bit) & 1) v2 |= ext;
    }
    if (name <= 2) {
      unsigned __int128 r = name == 0 ? v1 + v2 : name == 1 ? v1 - v2 : v1 * v2;
      for (uint32_t j = 0; j < len; ++j) {
        vpPut(c, uint16_t(res + j), uint32_t(r & 15u));
        r >>= 4;
      }
    } else {
      const __int128 s1 = (__int128)v1, s2 = (__int128)v2;
      vpPut(c, res, name == 3 ? (s1 == s2) : name == 4 ? (s1 < s2) : name == 5 ? (s1 <= s2) : 0);
    }
  }
}

__device__ __forceinline__ void memReadB(const Ctx &c, const uint32_t *T) {
  const uint32_t n = T[0] >> 8;
  const BatchArgs &a = c.a;
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t *o = T + 1 + 4 * k;
    const uint32_t en = opnd(c, o[2] & 0xFFFFu);
    if (en == 0) continue;
    uint32_t addr = 0;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) addr |= vpAt(c, (o[2] >> 16) + j) << (4 * j);
    if (o[3] >> 16) addr <<= 2;
    const uint64_t at = (uint64_t(o[0]) | uint64_t(o[1]) << 32) + addr;
    vpPut(c, o[3] & 0xFFFFu, at < a.memSize ? a.mem[at * a.instPad + c.slot0 + c.lane] : 0u);
  }
}

__device__ __forceinline__ void batchItem(const BatchArgs &a, uint32_t pid, uint32_t g, uint8_t *vp, uint8_t *lvlBase,
                          const uint8_t *lut, const uint64_t *draws, uint32_t c) {
  const uint32_t *P = a.words + a.partOff[pid];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6, nWaves = blockDim.x >> 6;
  const uint32_t slot0 = g * 64u, inst = slot0 + lane;
  const bool live = c < a.runCycles[inst];
  const uint64_t cycle = a.cycleBase[inst] + c;
  Ctx cx{a, vp, lvlBase + wave * 4096u, lut, lane, g, slot0, cycle};
  {
    const uint32_t nrr = P[H_NRR], ner = P[H_NER], nsr = P[H_NSR];
    const uint32_t *L = P + P[H_OFF_L0];
    for (uint32_t k = wave; k < nrr; k += nWaves) vpPut(cx, L[2 * k + 1], a.reg[size_t(L[2 * k]) * a.instPad + inst]);
    L += 2 * nrr;
    for (uint32_t k = wave; k < ner; k += nWaves) vpPut(cx, L[2 * k + 1], a.exg[size_t(L[2 * k]) * a.instPad + inst]);
    L += 2 * ner;
    for (uint32_t k = wave; k < nsr; k += nWaves) {
      const uint32_t in = L[2 * k];
      vpPut(cx, L[2 * k + 1], uint32_t((draws[size_t(in >> 6) * a.instPad + inst] >> (in & 63)) & 1u));
    }
  }
  __syncthreads();
  const uint8_t *consts = a.consts + P[H_CONST_BASE];
  const uint32_t nLev = P[H_NLEV];
  const uint32_t *lev = P + P[H_OFF_LEV];
  const uint32_t *tab = P + P[H_OFF_TASK];
  for (uint32_t e = 0; e < nLev; ++e) {
    for (uint32_t t = lev[2 * e] + wave; t < lev[2 * e + 1]; t += nWaves) {
      const uint32_t *T = P + tab[t];
      const uint32_t kind = T[0] & 0xFFu;
      if (kind == T_LUT2) lutTask(cx, T);
      else if (kind == T_VECREAD) vecReadB(cx, T, consts);
      else if (kind == T_VECOP) vecOpB(cx, T, consts);
      else memReadB(cx, T);
    }
    __syncthreads();
  }
  // last level (SimEval.cu:625-761); lanes past their cycle count write nothing
  const uint32_t ewc = P[H_EW_DC] >> 16;
  if (ewc) {
    const uint32_t exg = (P[H_EW_EXG] >> 2) * 4, dat = ((P[H_EW_DC] & 0xFFFFu) >> 2) * 4, n = (ewc + 3) / 4 * 4;
    if (live)
      for (uint32_t k = wave; k < n; k += nWaves) a.exg[size_t(exg + k) * a.instPad + inst] = uint8_t(vpAt(cx, dat + k));
  } else {
    const uint32_t rwc = P[H_RW_DC] >> 16;
    if (rwc && live) {
      const uint32_t reg = (P[H_RW_REG] >> 2) * 4, dat = ((P[H_RW_DC] & 0xFFFFu) >> 2) * 4, n = (rwc + 3) / 4 * 4;
      for (uint32_t k = wave; k < n; k += nWaves) a.reg[size_t(reg + k) * a.instPad + inst] = uint8_t(vpAt(cx, dat + k));
    }
    const uint32_t nmw = P[H_NMW], npr = P[H_NPR], nst = P[H_NST];
    const uint32_t *L = P + P[H_OFF_LAST];
    if (live)
      for (uint32_t k = wave; k < nmw; k += nWaves) {
        const uint32_t *o = L + 4 * k;
        if (opnd(cx, o[3] & 0xFFFFu) == 0) continue;
        const uint32_t dat = opnd(cx, o[2] >> 16);
        uint32_t addr = 0;
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) addr |= vpAt(cx, (o[2] & 0xFFFFu) + j) << (4 * j);
        if (o[3] >> 16) addr <<= 2;
        const uint64_t at = (uint64_t(o[0]) | uint64_t(o[1]) << 32) + addr;
        if (at < a.memSize) a.mem[at * a.instPad + inst] = uint8_t(dat);
      }
    L += 4 * nmw;
    if (live && a.printEnable)
      for (uint32_t k = wave; k < npr; k += nWaves)
        if (opnd(cx, L[k] & 0xFFFFu) != 0) {
          uint32_t q = atomicAdd(a.logCount, 1u);
          if (q < a.logCap) {
            a.logCycle[q] = cycle;
            a.logMsg[q] = uint64_t(L[k] >> 16) | (uint64_t(a.instOf[inst]) << 32);
          }
        }
    L += npr;
    if (live)
      for (uint32_t k = wave; k < nst; k += nWaves)
        if (opnd(cx, L[k] & 0xFFFFu) != 0) {
          atomicOr(reinterpret_cast<unsigned long long *>(&a.stop[inst]), 1ull);
          atomicMin(reinterpret_cast<unsigned long long *>(&a.stopCycle[inst]), (unsigned long long)cycle);
        }
  }
  __syncthreads();
}

// Advance the stimulus of every live instance of the groups this workgroup
// owns and store the draws of cycle `cycle` in buffer `buf`.
__device__ __forceinline__ void stimulusFor(const BatchArgs &a, uint32_t buf, uint32_t c, uint32_t activeGroups) {
  if (a.nIn == 0) return;
  const uint32_t nDraws = (a.nIn + 63) / 64;
  uint64_t *draws = a.draws + size_t(buf) * nDraws * a.instPad;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < activeGroups * 64u; s += gridDim.x * blockDim.x) {
    if (c >= a.runCycles[s]) continue;
    uint64_t x = a.xs[s];
    for (uint32_t d = 0; d < nDraws; ++d) {
      x = xorshift64(x);
      draws[size_t(d) * a.instPad + s] = x;
    }
    a.xs[s] = x;
  }
}

}  // namespace

__global__ __launch_bounds__(1024) void batchCycles(BatchArgs a) {
  extern __shared__ __align__(16) uint8_t bLds[];
  uint8_t *lut = bLds;
  uint8_t *vp = bLds + kLutBytesB;
  uint8_t *lvl = vp + a.vpBytes * 64u;
  for (uint32_t k = threadIdx.x; k < kLutBytesB / 4; k += blockDim.x)
    reinterpret_cast<uint32_t *>(lut)[k] = reinterpret_cast<const uint32_t *>(a.lut)[k];
  const uint32_t nDraws = (a.nIn + 63) / 64;
  uint32_t bar = 0;
  // draws of the first cycle, before anyone reads them
  stimulusFor(a, 0, 0, a.groupsActive[0]);
  bar++;
  gridBarrier(a.barrier, bar * gridDim.x, false);
  for (uint32_t c = 0; c < a.nCycles; ++c) {
    const uint32_t active = a.groupsActive[c];
    if (active == 0) break;
    const uint64_t *draws = a.draws + size_t(c & 1u) * nDraws * a.instPad;
    // input registers of live instances (never read back through normal register reads)
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < active * 64u; s += gridDim.x * blockDim.x)
      if (c < a.runCycles[s])
        for (uint32_t i = 0; i < a.nIn; ++i)
          a.reg[size_t(a.inIds[i]) * a.instPad + s] = uint8_t((draws[size_t(i >> 6) * a.instPad + s] >> (i & 63)) & 1u);
    const uint32_t items0 = a.n0 * active;
    for (uint32_t k = blockIdx.x; k < items0; k += gridDim.x)
      batchItem(a, k / active, k % active, vp, lvl, lut, draws, c);
    bar++;
    gridBarrier(a.barrier, bar * gridDim.x, false);
    if (c + 1 < a.nCycles) stimulusFor(a, (c + 1) & 1u, c + 1, a.groupsActive[c + 1]);
    const uint32_t items1 = (a.nTotal - a.n0) * active;
    for (uint32_t k = blockIdx.x; k < items1; k += gridDim.x)
      batchItem(a, a.n0 + k / active, k % active, vp, lvl, lut, draws, c);
    bar++;
    gridBarrier(a.barrier, bar * gridDim.x, false);
  }
}

hipError_t launchBatch(const BatchArgs &args, uint32_t grid, uint32_t block, uint32_t lds, hipStream_t stream) {
  BatchArgs a = args;
  void *params[] = {&a};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(batchCycles), dim3(grid), dim3(block), params, lds,
                                    stream);
}

hipError_t batchMaxBlocksPerCU(int *blocks, uint32_t block, uint32_t lds) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, reinterpret_cast<const void *>(batchCycles), block, lds);
}

hipError_t batchSetLds(uint32_t lds) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(batchCycles), hipFuncAttributeMaxDynamicSharedMemorySize,
                             lds);
}

}  // namespace tgs

namespace tgs {

// Reorder instance slots of an instance-minor pool: out[r][s] = in[r][src[s]].
__global__ void permuteSlots(const uint8_t *in, uint8_t *out, uint64_t rows, uint32_t instPad, const uint32_t *src) {
  const uint64_t total = rows * instPad;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t r = i / instPad;
    const uint32_t s = uint32_t(i - r * instPad);
    out[i] = in[r * instPad + src[s]];
  }
}

__global__ void permuteSlots64(const uint64_t *in, uint64_t *out, uint64_t rows, uint32_t instPad, const uint32_t *src) {
  const uint64_t total = rows * instPad;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t r = i / instPad;
    const uint32_t s = uint32_t(i - r * instPad);
    out[i] = in[r * instPad + src[s]];
  }
}

hipError_t launchPermute(const uint8_t *in, uint8_t *out, uint64_t rows, uint32_t instPad, const uint32_t *src,
                         hipStream_t stream) {
  if (rows == 0) return hipSuccess;
  permuteSlots<<<dim3(1024), dim3(256), 0, stream>>>(in, out, rows, instPad, src);
  return hipGetLastError();
}

hipError_t launchPermute64(const uint64_t *in, uint64_t *out, uint64_t rows, uint32_t instPad, const uint32_t *src,
                           hipStream_t stream) {
  if (rows == 0) return hipSuccess;
  permuteSlots64<<<dim3(256), dim3(256), 0, stream>>>(in, out, rows, instPad, src);
  return hipGetLastError();
}

}  // namespace tgs
