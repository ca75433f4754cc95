// Batched cycle kernel for gfx950: many stimulus instances of one design in
// one persistent launch, one instance per lane.
//
// Semantics per instance are the reference evaluator's (src/SimEval.cu:304-761,
// :765-864, :889-965). The execution is laid out for 64-wide waves:
//   * every pool is instance-minor (nibble k of slot s at k * instPad + s), so
//     a 64-instance group's nibble k is 64 contiguous bytes; L0 reads and the
//     last-level register / exchange copies move them 16 bytes per lane,
//     16 ops per wave instruction;
//   * a work item is (partition, up to W groups): the partition's program is
//     staged into LDS once and each of the W waves runs the whole partition
//     for its own group with its own value pool [slot][lane] and level buffer
//     [2][32][64] in LDS, so waves never wait on each other inside a region;
//   * ops of one micropartition level are issued four at a time so their LDS
//     latencies overlap; a lane index past the previous level's size reads 0,
//     as the reference's idle lanes do (SimEval.cu:372-375,406-408).
// Instances run uneven cycle counts: the host orders instance slots by
// decreasing cycle count, so the live groups of a cycle are a prefix and the
// live lanes of a group are a prefix; lanes past their own count never write.
// Stimulus draws for cycle c+1 are produced during region 1 of cycle c into
// the other half of a double buffer, so no extra grid barrier is needed.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernel_args.h"

namespace tgs {

namespace {

constexpr uint32_t kLutBytesB = ((kLutSize + 15) / 16) * 16;

struct Ctx {
  const uint8_t *mem;
  uint64_t memSize;
  size_t instPad;
  uint8_t *vp;    // this wave's value pool [slot][64]
  uint8_t *lvl;   // this wave's level buffer [2][32][64]
  const uint8_t *lut;
  uint32_t lane, slot0;
};

__device__ __forceinline__ uint32_t vpAt(const Ctx &c, uint32_t s) { return c.vp[s * 64u + c.lane]; }
__device__ __forceinline__ uint32_t opnd(const Ctx &c, uint32_t s) { return s < 16 ? s : vpAt(c, s); }
__device__ __forceinline__ void vpPut(const Ctx &c, uint32_t s, uint32_t v) { c.vp[s * 64u + c.lane] = uint8_t(v); }

// Middle-level operand: immediate below 32, else lane (o mod 32) of the
// previous level, 0 past that level's size (SimEval.cu:372-404).
__device__ __forceinline__ uint32_t midOp(const uint8_t *__restrict__ P, uint32_t o, uint32_t nPrev, uint32_t lane) {
  return o < 32 ? o : ((o & 31u) < nPrev ? P[(o & 31u) * 64u + lane] : 0u);
}

// One on-disk micropartition, lanes = instances.
__device__ __forceinline__ void lutHalf(const Ctx &c, const uint32_t *S, const uint32_t *top, const uint32_t *ops,
                                        uint32_t nTop, uint32_t nMid, uint32_t half, const uint32_t *last,
                                        uint32_t nLast) {
  uint8_t *__restrict__ L = c.lvl;
  const uint8_t *__restrict__ vp = c.vp;
  const uint8_t *__restrict__ lut = c.lut;
  const uint32_t lane = c.lane;
  auto topOp = [&](uint32_t i) -> uint32_t {
    const uint32_t x = top[2 * i], y = top[2 * i + 1];
    const uint32_t lt = x & 0xFFFFu, o0 = x >> 16, o1 = y & 0xFFFFu, o2 = y >> 16;
    const uint32_t cc = o2 < 16 ? o2 : vp[o2 * 64u + lane];
    const uint32_t aa = o0 < 16 ? o0 : vp[o0 * 64u + lane];
    const uint32_t bb = o1 < 16 ? o1 : vp[o1 * 64u + lane];
    const uint32_t pos = lutIndex(lt, aa, bb, cc);
    const uint32_t lv = pos < kLutSize ? lut[pos] : 0u;
    return lt == 0 ? cc : lv;
  };
  uint32_t i = 0;
  for (; i + 4 <= nTop; i += 4) {
    const uint32_t v0 = topOp(i), v1 = topOp(i + 1), v2 = topOp(i + 2), v3 = topOp(i + 3);
    L[i * 64u + lane] = uint8_t(v0);
    L[(i + 1) * 64u + lane] = uint8_t(v1);
    L[(i + 2) * 64u + lane] = uint8_t(v2);
    L[(i + 3) * 64u + lane] = uint8_t(v3);
  }
  for (; i < nTop; ++i) L[i * 64u + lane] = uint8_t(topOp(i));
  uint32_t nPrev = nTop, cur = 0;
  for (uint32_t m = 0; m < nMid; ++m) {
    const uint32_t sz = (S[m >> 1] >> ((m & 1u) << 4)) & 0xFFFFu;
    const uint32_t na = sz & 0xFFu, nb = sz >> 8;
    const uint32_t n = half ? nb : na;
    const uint32_t *w = ops + (half ? na : 0u);
    const uint8_t *__restrict__ P = L + cur * 2048u;
    uint8_t *__restrict__ Q = L + (cur ^ 1u) * 2048u;
    auto midEval = [&](uint32_t op) -> uint32_t {
      const uint32_t o2 = op & 63u, o1 = (op >> 6) & 63u, o0 = (op >> 12) & 63u, lt = op >> 18;
      const uint32_t aa = midOp(P, o0, nPrev, lane), bb = midOp(P, o1, nPrev, lane), dd = midOp(P, o2, nPrev, lane);
      const uint32_t pos = lutIndex(lt, aa, bb, dd);
      const uint32_t lv = pos < kLutSize ? lut[pos] : 0u;
      return lt == 0 ? dd : lv;
    };
    uint32_t k = 0;
    for (; k + 4 <= n; k += 4) {
      const uint32_t w0 = w[k], w1 = w[k + 1], w2 = w[k + 2], w3 = w[k + 3];
      const uint32_t v0 = midEval(w0), v1 = midEval(w1), v2 = midEval(w2), v3 = midEval(w3);
      Q[k * 64u + lane] = uint8_t(v0);
      Q[(k + 1) * 64u + lane] = uint8_t(v1);
      Q[(k + 2) * 64u + lane] = uint8_t(v2);
      Q[(k + 3) * 64u + lane] = uint8_t(v3);
    }
    for (; k < n; ++k) Q[k * 64u + lane] = uint8_t(midEval(w[k]));
    ops += na + nb;
    nPrev = n;
    cur ^= 1u;
  }
  const uint8_t *__restrict__ P = L + cur * 2048u;
  uint8_t *__restrict__ vpw = c.vp;
  for (uint32_t k = 0; k < nLast; ++k) {
    const uint32_t w = last[k];
    const uint32_t sid = (w >> 16) & 0xFFu;
    const uint32_t v = sid < 32 ? sid : ((sid & 31u) < nPrev ? P[(sid & 31u) * 64u + lane] : 0u);
    vpw[(w & 0xFFFFu) * 64u + lane] = uint8_t(v);
  }
}

__device__ __forceinline__ void lutTask(const Ctx &c, const uint32_t *T) {
  const uint32_t w0 = T[0], w1 = T[1];
  const uint32_t nMid = (w0 >> 8) & 0xFFFFu;
  const uint32_t ntA = w1 & 0xFFu, ntB = (w1 >> 8) & 0xFFu, nlA = (w1 >> 16) & 0xFFu, nlB = w1 >> 24;
  const uint32_t *S = T + 3;
  const uint32_t *top = S + ((nMid + 1) >> 1);
  const uint32_t *mids = top + 2 * (ntA + ntB);
  const uint32_t *last = T + T[2];
  lutHalf(c, S, top, mids, ntA, nMid, 0, last, nlA);
  if (ntB | nlB) lutHalf(c, S, top + 2 * ntA, mids, ntB, nMid, 1, last + nlA, nlB);
}

// VecRead (SimEval.cu:449-481). Ops are evaluated one after another; the
// loader rejects programs in which one op of an exec level reads a slot
// another op of that level writes, so the order is unobservable.
__device__ __forceinline__ void vecReadB(const Ctx &c, const uint32_t *T, const uint8_t *consts) {
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

// VecArithOrLogic (SimEval.cu:482-548).
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

// MemRead (SimEval.cu:550-575).
__device__ __forceinline__ void memReadB(const Ctx &c, const uint32_t *T) {
  const uint32_t n = T[0] >> 8;
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t *o = T + 1 + 4 * k;
    const uint32_t en = opnd(c, o[2] & 0xFFFFu);
    if (en == 0) continue;
    uint32_t addr = 0;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) addr |= vpAt(c, (o[2] >> 16) + j) << (4 * j);
    if (o[3] >> 16) addr <<= 2;
    const uint64_t at = (uint64_t(o[0]) | uint64_t(o[1]) << 32) + addr;
    vpPut(c, o[3] & 0xFFFFu, at < c.memSize ? c.mem[at * c.instPad + c.slot0 + c.lane] : 0u);
  }
}

// L0 reads of one group: op k copies pool[src_k][slot0 .. slot0+64) into
// vp[slot_k][0 .. 64); lane l moves 16 bytes of op (k0 + l/4).
__device__ __forceinline__ void wideRead(uint8_t *__restrict__ vp, const uint8_t *__restrict__ pool, size_t instPad,
                                         uint32_t slot0, const uint32_t *__restrict__ L, uint32_t n, uint32_t lane) {
  const uint32_t q = lane >> 2, j = (lane & 3u) * 16u;
  const uint8_t *base = pool + slot0 + j;
  uint32_t k = q;
  for (; k + 48 < n; k += 64) {
    const uint32_t s0 = L[2 * k], d0 = L[2 * k + 1], s1 = L[2 * k + 32], d1 = L[2 * k + 33];
    const uint32_t s2 = L[2 * k + 64], d2 = L[2 * k + 65], s3 = L[2 * k + 96], d3 = L[2 * k + 97];
    const uint4 v0 = *reinterpret_cast<const uint4 *>(base + size_t(s0) * instPad);
    const uint4 v1 = *reinterpret_cast<const uint4 *>(base + size_t(s1) * instPad);
    const uint4 v2 = *reinterpret_cast<const uint4 *>(base + size_t(s2) * instPad);
    const uint4 v3 = *reinterpret_cast<const uint4 *>(base + size_t(s3) * instPad);
    *reinterpret_cast<uint4 *>(vp + d0 * 64u + j) = v0;
    *reinterpret_cast<uint4 *>(vp + d1 * 64u + j) = v1;
    *reinterpret_cast<uint4 *>(vp + d2 * 64u + j) = v2;
    *reinterpret_cast<uint4 *>(vp + d3 * 64u + j) = v3;
  }
  for (; k < n; k += 16) {
    const uint32_t s0 = L[2 * k], d0 = L[2 * k + 1];
    *reinterpret_cast<uint4 *>(vp + d0 * 64u + j) = *reinterpret_cast<const uint4 *>(base + size_t(s0) * instPad);
  }
}

// Word copy of the last level (SimEval.cu:658-665,751-758) for one group:
// rows [0, n) of vp from `dat` to `pool` from `row0`, live lanes only
// (the live lanes of a group are a prefix of length `live`).
__device__ __forceinline__ void wideWrite(uint8_t *pool, size_t instPad, uint32_t slot0, uint32_t row0,
                                          const uint8_t *vp, uint32_t dat, uint32_t n, uint32_t lane, uint32_t live) {
  const uint32_t q = lane >> 2, j = (lane & 3u) * 16u;
  for (uint32_t k = q; k < n; k += 16) {
    const uint4 v = *reinterpret_cast<const uint4 *>(vp + (dat + k) * 64u + j);
    uint8_t *dst = pool + size_t(row0 + k) * instPad + slot0 + j;
    if (j + 16 <= live) {
      *reinterpret_cast<uint4 *>(dst) = v;
    } else if (j < live) {
      for (uint32_t t = 0; t < live - j; ++t) {
        const uint32_t w = t < 4 ? v.x : t < 8 ? v.y : t < 12 ? v.z : v.w;
        dst[t] = uint8_t(w >> ((t & 3u) * 8u));
      }
    }
  }
}

// One (partition, group) evaluated by one wave, no workgroup barriers.
__device__ __forceinline__ void waveItem(const BatchArgs &a, uint32_t part, const uint32_t *P, uint32_t g, uint8_t *vp,
                                         uint8_t *lvl, const uint8_t *lut, const uint64_t *draws, uint32_t c) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t slot0 = g * 64u, inst = slot0 + lane;
  const bool liveLane = c < a.runCycles[inst];
  const uint32_t live = __popcll(__ballot(liveLane));
  const uint64_t cycle = a.cycleBase[inst] + c;
  const size_t IP = a.instPad;
  Ctx cx{a.mem, a.memSize, IP, vp, lvl, lut, lane, slot0};
  {
    const uint32_t nrr = P[H_NRR], ner = P[H_NER], nsr = P[H_NSR];
    const uint32_t *L = P + P[H_OFF_L0];
    wideRead(vp, a.reg, IP, slot0, L, nrr, lane);
    L += 2 * nrr;
    wideRead(vp, a.exg, IP, slot0, L, ner, lane);
    L += 2 * ner;
    for (uint32_t k = 0; k < nsr; ++k) {
      const uint32_t in = L[2 * k];
      vpPut(cx, L[2 * k + 1], uint32_t((draws[size_t(in >> 6) * IP + inst] >> (in & 63)) & 1u));
    }
  }
  const uint8_t *consts = a.consts + P[H_CONST_BASE];
  const uint32_t nLev = P[H_NLEV];
  const uint32_t *lev = P + P[H_OFF_LEV];
  const uint32_t *tab = P + P[H_OFF_TASK];
  for (uint32_t e = 0; e < nLev; ++e) {
    for (uint32_t t = lev[2 * e]; t < lev[2 * e + 1]; ++t) {
      const uint32_t *T = P + tab[t];
      const uint32_t kind = T[0] & 0xFFu;
      if (kind == T_LUT2) lutTask(cx, T);
      else if (kind == T_VECREAD) vecReadB(cx, T, consts);
      else if (kind == T_VECOP) vecOpB(cx, T, consts);
      else memReadB(cx, T);
    }
  }
  // signal probes of the live lanes (slot-major, [probe][instPad])
  if (a.probeOff != nullptr && liveLane)
    for (uint32_t k = a.probeOff[part]; k < a.probeOff[part + 1]; ++k)
      a.probeVal[size_t(k) * IP + inst] = uint8_t(vpAt(cx, a.probeSlot[k]));
  // last level (SimEval.cu:625-761); lanes past their cycle count write nothing
  const uint32_t ewc = P[H_EW_DC] >> 16;
  if (ewc) {
    wideWrite(a.exg, IP, slot0, (P[H_EW_EXG] >> 2) * 4, vp, ((P[H_EW_DC] & 0xFFFFu) >> 2) * 4, (ewc + 3) / 4 * 4, lane,
              live);
    return;
  }
  const uint32_t rwc = P[H_RW_DC] >> 16;
  if (rwc)
    wideWrite(a.reg, IP, slot0, (P[H_RW_REG] >> 2) * 4, vp, ((P[H_RW_DC] & 0xFFFFu) >> 2) * 4, (rwc + 3) / 4 * 4, lane,
              live);
  const uint32_t nmw = P[H_NMW], npr = P[H_NPR], nst = P[H_NST];
  const uint32_t *L = P + P[H_OFF_LAST];
  if (liveLane)
    for (uint32_t k = 0; k < nmw; ++k) {
      const uint32_t *o = L + 4 * k;
      if (opnd(cx, o[3] & 0xFFFFu) == 0) continue;
      const uint32_t dat = opnd(cx, o[2] >> 16);
      uint32_t addr = 0;
#pragma unroll
      for (uint32_t j = 0; j < 8; ++j) addr |= vpAt(cx, (o[2] & 0xFFFFu) + j) << (4 * j);
      if (o[3] >> 16) addr <<= 2;
      const uint64_t at = (uint64_t(o[0]) | uint64_t(o[1]) << 32) + addr;
      if (at < a.memSize) a.mem[at * IP + inst] = uint8_t(dat);
    }
  L += 4 * nmw;
  if (liveLane && a.printEnable)
    for (uint32_t k = 0; k < npr; ++k)
      if (opnd(cx, L[k] & 0xFFFFu) != 0) {
        const uint64_t q = atomicAdd(reinterpret_cast<unsigned long long *>(a.logCount), 1ull);
        if (q < a.logCap) {
          a.logCycle[q] = cycle;
          a.logMsg[q] = uint64_t(L[k] >> 16) | (uint64_t(a.instOf[inst]) << 32);
        }
      }
  L += npr;
  if (liveLane)
    for (uint32_t k = 0; k < nst; ++k)
      if (opnd(cx, L[k] & 0xFFFFu) != 0) {
        atomicOr(reinterpret_cast<unsigned long long *>(&a.stop[inst]), 1ull);
        atomicMin(reinterpret_cast<unsigned long long *>(&a.stopCycle[inst]), (unsigned long long)cycle);
      }
}

// Advance the stimulus of every live instance of the active groups and store
// the draws of launch cycle `c` in buffer `buf`.
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

// Copy a partition's program words (16-byte aligned, padded to 4 words) into LDS.
__device__ __forceinline__ void stageProgram(uint32_t *dst, const uint32_t *src, uint32_t words) {
  const uint32_t n4 = (words + 3) >> 2;
  for (uint32_t k = threadIdx.x; k < n4; k += blockDim.x)
    reinterpret_cast<uint4 *>(dst)[k] = reinterpret_cast<const uint4 *>(src)[k];
}

// Work items of one region: (partition, W consecutive groups), one group per wave.
__device__ __forceinline__ void batchRegion(const BatchArgs &a, uint32_t pb, uint32_t pe, uint32_t active,
                                            uint8_t *vpAll, uint8_t *lvlAll, uint32_t *prog, const uint8_t *lut,
                                            const uint64_t *draws, uint32_t c) {
  const uint32_t np = pe - pb, W = blockDim.x >> 6, wave = threadIdx.x >> 6;
  const uint32_t nChunks = (active + W - 1) / W;
  const uint32_t items = np * nChunks;
  uint8_t *vp = vpAll + wave * a.vpBytes * 64u;
  uint8_t *lvl = lvlAll + wave * 4096u;
  for (uint32_t k = blockIdx.x; k < items; k += gridDim.x) {
    const uint32_t p = pb + k / nChunks, g = (k % nChunks) * W + wave;
    stageProgram(prog, a.words + a.partOff[p], a.partWords[p]);
    __syncthreads();
    if (g < active) waveItem(a, p, prog, g, vp, lvl, lut, draws, c);
    __syncthreads();
  }
}

}  // namespace

__global__ __launch_bounds__(1024) void batchCycles(BatchArgs a) {
  extern __shared__ __align__(16) uint8_t bLds[];
  const uint32_t W = blockDim.x >> 6;
  uint8_t *lut = bLds;
  uint8_t *vp = bLds + kLutBytesB;
  uint8_t *lvl = vp + W * a.vpBytes * 64u;
  uint32_t *prog = reinterpret_cast<uint32_t *>(lvl + W * 4096u);
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
    batchRegion(a, 0, a.n0, active, vp, lvl, prog, lut, draws, c);
    bar++;
    gridBarrier(a.barrier, bar * gridDim.x, false);
    if (c + 1 < a.nCycles) stimulusFor(a, (c + 1) & 1u, c + 1, a.groupsActive[c + 1]);
    batchRegion(a, a.n0, a.nTotal, active, vp, lvl, prog, lut, draws, c);
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

// One instance's column of an instance-minor pool to or from a dense buffer
// (host register and pool I/O of the batched layout).
__global__ void slotColumn(uint8_t *pool, uint32_t instPad, uint32_t slot, uint64_t rows, uint8_t *dense, int toPool) {
  for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < rows; r += uint64_t(gridDim.x) * blockDim.x) {
    if (toPool)
      pool[r * instPad + slot] = dense[r];
    else
      dense[r] = pool[r * instPad + slot];
  }
}

// Consecutive threads take consecutive slots of one register row, so the
// pool side of an instance-minor layout is coalesced.
__global__ void regMailbox(uint8_t *pool, uint32_t instPad, const uint32_t *instOf, uint32_t nSlots,
                           const uint32_t *ids, uint32_t n, uint8_t *vals, int toPool, const uint32_t *ctlSrc,
                           uint32_t *ctlDst, uint32_t ctlWords) {
  if (blockIdx.x == 0 && threadIdx.x < ctlWords) ctlDst[threadIdx.x] = ctlSrc[threadIdx.x];
  const uint64_t total = uint64_t(nSlots) * n;
  for (uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t k = uint32_t(t / nSlots), s = uint32_t(t - uint64_t(k) * nSlots);
    const uint32_t i = instOf ? instOf[s] : s;
    uint8_t *p = pool + size_t(ids[k]) * instPad + s;
    uint8_t *v = vals + size_t(i) * n + k;
    if (toPool)
      *p = *v;
    else
      *v = *p;
  }
}

hipError_t launchRegMailbox(uint8_t *pool, uint32_t instPad, const uint32_t *instOf, uint32_t nSlots,
                            const uint32_t *ids, uint32_t n, uint8_t *vals, bool toPool, hipStream_t stream,
                            const uint32_t *ctlSrc, uint32_t *ctlDst, uint32_t ctlWords) {
  const uint64_t total = uint64_t(nSlots) * n;
  if (ctlWords > 256) return hipErrorInvalidValue;
  if (total == 0 && ctlWords == 0) return hipSuccess;
  const uint32_t grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>((total + 255) / 256, 1024)));
  regMailbox<<<dim3(grid), dim3(256), 0, stream>>>(pool, instPad, instOf, nSlots, ids, n, vals, toPool ? 1 : 0,
                                                   ctlSrc, ctlDst, ctlWords);
  return hipGetLastError();
}

__global__ void fill32(uint32_t *out, uint64_t n, uint32_t value) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = value;
}

hipError_t launchFill32(uint32_t *out, uint64_t n, uint32_t value, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint32_t grid = uint32_t(std::min<uint64_t>((n + 255) / 256, 1024));
  fill32<<<dim3(grid), dim3(256), 0, stream>>>(out, n, value);
  return hipGetLastError();
}

hipError_t launchSlotColumn(uint8_t *pool, uint32_t instPad, uint32_t slot, uint64_t rows, uint8_t *dense, bool toPool,
                            hipStream_t stream) {
  if (rows == 0) return hipSuccess;
  const uint32_t grid = uint32_t(std::min<uint64_t>((rows + 255) / 256, 1024));
  slotColumn<<<dim3(grid), dim3(256), 0, stream>>>(pool, instPad, slot, rows, dense, toPool ? 1 : 0);
  return hipGetLastError();
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
