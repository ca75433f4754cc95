// Persistent wave64 cycle kernel for gfx950.
//
// Semantics are those of the reference evaluator (src/SimEval.cu:304-761 per
// micropartition, :765-864 per partition, :889-965 per cycle); the execution
// is re-planned for CDNA4:
//   * one persistent workgroup per CU, partitions statically assigned per
//     region by a longest-processing-time schedule (no atomic work queues);
//   * the 4642-entry LUT and the value pool live in LDS; the program words
//     are evaluated from LDS too: either resident for the whole launch (when
//     the workgroup's partitions fit) or staged per partition visit with
//     16-byte coalesced copies, so no op word is a dependent HBM load;
//   * two on-disk 32-lane micropartitions run in the two halves of one
//     64-lane wave; operand shuffles are ds_bpermute within a half, and the
//     next level's op word is loaded while the current level evaluates;
//   * vector/memory ops of an exec level are regrouped 64 per wave;
//   * one grid barrier per region on a two-level arrival tree, launched
//     cooperatively so every workgroup is resident; the register, exchange
//     and memory pools (the only data workgroups exchange) are read and
//     written with agent-coherent accesses, so the barrier needs no L2
//     write-back or invalidate.
#include <hip/hip_runtime.h>

#include "kernel_args.h"

namespace tgs {

namespace {

constexpr uint32_t kLutBytes = ((kLutSize + 15) / 16) * 16;

__device__ __forceinline__ uint32_t rdOp(const uint8_t *vp, uint32_t s) { return s < 16 ? s : vp[s]; }

__device__ __forceinline__ uint32_t lutL(const uint8_t *lut, uint32_t pos) { return pos < kLutSize ? lut[pos] : 0u; }

// Copy `words` u32 words (16-byte aligned source and destination, the source
// padded to a 4-word multiple) with the whole workgroup.
__device__ __forceinline__ void copyWords(uint32_t *dst, const uint32_t *src, uint32_t words) {
  const uint32_t n4 = (words + 3) >> 2;
  const uint4 *s = reinterpret_cast<const uint4 *>(src);
  uint4 *d = reinterpret_cast<uint4 *>(dst);
  for (uint32_t k = threadIdx.x; k < n4; k += blockDim.x) d[k] = s[k];
}

// Two 32-lane LUT micropartitions in one wave (SimEval.cu:319-448).
__device__ __forceinline__ void lutPair(const uint32_t *T, uint8_t *vp, const uint8_t *lut) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t h = lane >> 5, i = lane & 31u, hb = h << 5;
  const uint32_t w0 = T[0], w1 = T[1], lastOff = T[2];
  const uint32_t nMid = (w0 >> 8) & 0xFFFFu;
  const uint32_t ntA = w1 & 0xFFu, ntB = (w1 >> 8) & 0xFFu, nlA = (w1 >> 16) & 0xFFu, nlB = w1 >> 24;
  const uint32_t *S = T + 3;
  const uint32_t *ops = S + ((nMid + 1) >> 1);
  // independent loads first: this lane's top op and last-level op
  const uint32_t nt = h ? ntB : ntA, tb = h ? ntA : 0u;
  uint32_t x = 0, y = 0;
  if (i < nt) {
    x = ops[2 * (tb + i)];
    y = ops[2 * (tb + i) + 1];
  }
  const uint32_t nl = h ? nlB : nlA, lb = h ? nlA : 0u;
  const uint32_t wl = i < nl ? T[lastOff + lb + i] : 0u;
  ops += 2 * (ntA + ntB);
  uint32_t val = 0;
  if (i < nt) {
    const uint32_t lt = x & 0xFFFFu, o0 = x >> 16, o1 = y & 0xFFFFu, o2 = y >> 16;
    const uint32_t c = rdOp(vp, o2);
    if (lt == 0) {
      val = c;
    } else {
      const uint32_t a = rdOp(vp, o0), b = rdOp(vp, o1);
      val = lutL(lut, lutIndex(lt, a, b, c));
    }
  }
  // middle levels, the next level's op word loaded one level ahead
  uint32_t w = 0, na = 0, nb = 0;
  if (nMid) {
    const uint32_t sz = S[0] & 0xFFFFu;
    na = sz & 0xFFu;
    nb = sz >> 8;
    const uint32_t n = h ? nb : na;
    w = i < n ? ops[(h ? na : 0u) + i] : 0u;
  }
  for (uint32_t m = 0; m < nMid; ++m) {
    const uint32_t cur = w;
    ops += na + nb;
    if (m + 1 < nMid) {
      const uint32_t sz = (S[(m + 1) >> 1] >> (((m + 1) & 1u) << 4)) & 0xFFFFu;
      na = sz & 0xFFu;
      nb = sz >> 8;
      const uint32_t n = h ? nb : na;
      w = i < n ? ops[(h ? na : 0u) + i] : 0u;
    }
    const uint32_t o2 = cur & 63u, o1 = (cur >> 6) & 63u, o0 = (cur >> 12) & 63u, lt = cur >> 18;
    const uint32_t s0 = __shfl(val, int((o0 & 31u) | hb));
    const uint32_t s1 = __shfl(val, int((o1 & 31u) | hb));
    const uint32_t s2 = __shfl(val, int((o2 & 31u) | hb));
    const uint32_t a = o0 < 32 ? o0 : s0, b = o1 < 32 ? o1 : s1, c = o2 < 32 ? o2 : s2;
    const uint32_t lv = lutL(lut, lutIndex(lt, a, b, c));
    val = lt == 0 ? c : lv;
  }
  const uint32_t sid = (wl >> 16) & 0xFFu;
  uint32_t r = __shfl(val, int((sid & 31u) | hb));
  if (sid < 32) r = sid;
  if (i < nl) vp[wl & 0xFFFFu] = uint8_t(r);
}

// VecRead (SimEval.cu:449-481), up to 64 ops per wave.
__device__ __forceinline__ void vecReadTask(const uint32_t *T, uint8_t *vp, const uint8_t *consts) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t n = T[0] >> 8;
  if (lane >= n) return;
  const uint32_t *o = T + 1 + 5 * lane;
  const uint32_t w0 = o[0], w1 = o[1], w2 = o[2], w3 = o[3], w4 = o[4];
  const uint32_t base = w0 & 0xFFFFu, len = w0 >> 16, offset = w1 & 0xFFFFu, isConst = w1 >> 16;
  const uint32_t i0 = rdOp(vp, w2 & 0xFFFFu), i1 = rdOp(vp, w2 >> 16), i2 = rdOp(vp, w3 & 0xFFFFu), i3 = rdOp(vp, w3 >> 16);
  const uint32_t off = ((i0 << 12) | (i1 << 8) | (i2 << 4) | i3) + offset;
  uint32_t v;
  if (off < len) v = isConst ? consts[base + off] : vp[base + off];
  else v = rdOp(vp, w4 & 0xFFFFu);
  vp[w4 >> 16] = uint8_t(v);
}

// VecArithOrLogic (SimEval.cu:482-548), up to 64 ops per wave.
__device__ __forceinline__ void vecOpTask(const uint32_t *T, uint8_t *vp, const uint8_t *consts) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t n = T[0] >> 8;
  if (lane >= n) return;
  const uint32_t *o = T + 1 + 3 * lane;
  const uint32_t b1 = o[0] & 0xFFFFu, b2 = o[0] >> 16;
  const uint32_t len = o[1] & 0xFFFFu, name = (o[1] >> 16) & 0xFFu, cf = o[1] >> 24, res = o[2];
  unsigned __int128 v1, v2;
  if (cf == 0) {
    v1 = vpNibbles(vp, b1, len);
    v2 = vpNibbles(vp, b2, len);
  } else {
    v1 = v2 = 0;
    for (uint32_t k = 0; k < len; ++k) {
      unsigned __int128 s1 = (cf & 2) ? consts[b1 + k] : vp[b1 + k];
      unsigned __int128 s2 = (cf & 1) ? consts[b2 + k] : vp[b2 + k];
      v1 |= s1 << (4 * k);
      v2 |= s2 << (4 * k);
    }
  }
  if (len > 0 && len < 32) {  // sign_extend_128 (SimEval.cu:222-228)
    const uint32_t bit = 4 * len - 1;
    const unsigned __int128 ext = ~((((unsigned __int128)1) << (bit + 1)) - 1);
    if ((v1 >> bit) & 1) v1 |= ext;
    if ((v2 >> bit) & 1) v2 |= ext;
  }
  if (name <= 2) {
    const unsigned __int128 r = name == 0 ? v1 + v2 : name == 1 ? v1 - v2 : v1 * v2;
    vpPutNibbles(vp, res, len, r);
  } else {
    const __int128 s1 = (__int128)v1, s2 = (__int128)v2;
    vp[res] = uint8_t(name == 3 ? (s1 == s2) : name == 4 ? (s1 < s2) : name == 5 ? (s1 <= s2) : 0);
  }
}

// MemRead (SimEval.cu:550-575), up to 64 ops per wave.
__device__ __forceinline__ void memReadTask(const uint32_t *T, uint8_t *vp, const CommonArgs &c) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t n = T[0] >> 8;
  if (lane >= n) return;
  const uint32_t *o = T + 1 + 4 * lane;
  const uint64_t base = uint64_t(o[0]) | uint64_t(o[1]) << 32;
  const uint32_t en = rdOp(vp, o[2] & 0xFFFFu);
  if (en == 0) return;
  const uint64_t at = memAddress(vp, o[2] >> 16, o[3] >> 16, base);
  vp[o[3] & 0xFFFFu] = at < c.memSize ? uint8_t(ldShared8(c.mem + at)) : uint8_t(0);
}

__device__ __forceinline__ void w64Partition(const W64Args &a, uint32_t part, const uint32_t *P, uint8_t *vp,
                                             const uint8_t *lut, const uint64_t *draws, uint64_t cycle) {
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  const uint32_t wave = tid >> 6, nWaves = nt >> 6;
  // L0 (SimEval.cu:582-622) plus stimulus reads
  {
    const uint32_t nrr = P[H_NRR], ner = P[H_NER], nsr = P[H_NSR];
    const uint32_t *L = P + P[H_OFF_L0];
    for (uint32_t k = tid; k < nrr; k += nt) vp[L[2 * k + 1]] = uint8_t(ldShared8(a.c.reg + L[2 * k]));
    L += 2 * nrr;
    for (uint32_t k = tid; k < ner; k += nt) vp[L[2 * k + 1]] = uint8_t(ldShared8(a.c.exg + L[2 * k]));
    L += 2 * ner;
    for (uint32_t k = tid; k < nsr; k += nt) vp[L[2 * k + 1]] = uint8_t(stimBit(draws, L[2 * k]));
  }
  __syncthreads();
  const uint8_t *consts = a.c.consts + P[H_CONST_BASE];
  const uint32_t nLev = P[H_NLEV];
  const uint32_t *lev = P + P[H_OFF_LEV];
  const uint32_t *tab = P + P[H_OFF_TASK];
  for (uint32_t e = 0; e < nLev; ++e) {
    const uint32_t tb = lev[2 * e], te = lev[2 * e + 1];
    for (uint32_t t = tb + wave; t < te; t += nWaves) {
      const uint32_t *T = P + tab[t];
      const uint32_t kind = T[0] & 0xFFu;
      if (kind == T_LUT2) lutPair(T, vp, lut);
      else if (kind == T_VECREAD) vecReadTask(T, vp, consts);
      else if (kind == T_VECOP) vecOpTask(T, vp, consts);
      else memReadTask(T, vp, a.c);
    }
    __syncthreads();
  }
  spillProbes(a.c, part, vp, tid, nt);
  // last level (SimEval.cu:625-761)
  const uint32_t ewc = P[H_EW_DC] >> 16;
  if (ewc) {
    const uint32_t exg = P[H_EW_EXG], dat = P[H_EW_DC] & 0xFFFFu, words = (ewc + 3) / 4;
    for (uint32_t k = tid; k < words; k += nt)
      stShared32(reinterpret_cast<uint32_t *>(a.c.exg) + (exg >> 2) + k, reinterpret_cast<const uint32_t *>(vp)[(dat >> 2) + k]);
  } else {
    const uint32_t rwc = P[H_RW_DC] >> 16;
    if (rwc) {
      const uint32_t reg = P[H_RW_REG], dat = P[H_RW_DC] & 0xFFFFu, words = (rwc + 3) / 4;
      for (uint32_t k = tid; k < words; k += nt)
        stShared32(reinterpret_cast<uint32_t *>(a.c.reg) + (reg >> 2) + k, reinterpret_cast<const uint32_t *>(vp)[(dat >> 2) + k]);
    }
    const uint32_t nmw = P[H_NMW], npr = P[H_NPR], nst = P[H_NST];
    const uint32_t *L = P + P[H_OFF_LAST];
    for (uint32_t k = tid; k < nmw; k += nt) {
      const uint32_t *o = L + 4 * k;
      const uint32_t en = rdOp(vp, o[3] & 0xFFFFu);
      if (en != 0) {
        const uint32_t dat = rdOp(vp, o[2] >> 16);
        const uint64_t at = memAddress(vp, o[2] & 0xFFFFu, o[3] >> 16, uint64_t(o[0]) | uint64_t(o[1]) << 32);
        if (at < a.c.memSize) stShared8(a.c.mem + at, dat);
      }
    }
    L += 4 * nmw;
    if (a.c.printEnable)
      for (uint32_t k = tid; k < npr; k += nt)
        if (rdOp(vp, L[k] & 0xFFFFu) != 0) logPrint(a.c, cycle, L[k] >> 16);
    L += npr;
    for (uint32_t k = tid; k < nst; k += nt)
      if (rdOp(vp, L[k] & 0xFFFFu) != 0) raiseStop(a.c, cycle);
  }
  __syncthreads();
}

template <int MODE>
__device__ __forceinline__ const uint32_t *progOf(const W64Args &a, uint32_t *prog, uint32_t s) {
  if (MODE == PM_RESIDENT) return prog + a.schedLds[s];
  const uint32_t *src = a.words + a.partOff[a.schedPart[s]];
  if (MODE == PM_GLOBAL) return src;
  copyWords(prog, src, a.partWords[a.schedPart[s]]);
  __syncthreads();
  return prog;
}

}  // namespace

template <int MODE>
__global__ __launch_bounds__(1024) void w64Cycles(W64Args a) {
  extern __shared__ __align__(16) uint8_t w64Lds[];
  uint8_t *lut = w64Lds;
  uint8_t *vp = w64Lds + kLutBytes;
  uint64_t *draws = reinterpret_cast<uint64_t *>(vp + a.vpBytes);
  uint32_t *prog = reinterpret_cast<uint32_t *>(draws + kMaxDraws);
  const uint32_t g = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  for (uint32_t k = tid; k < kLutBytes / 4; k += nt)
    reinterpret_cast<uint32_t *>(lut)[k] = reinterpret_cast<const uint32_t *>(a.c.lut)[k];
  const uint32_t s0 = a.schedOff[2 * g], s1 = a.schedOff[2 * g + 1], s2 = a.schedOff[2 * g + 2];
  if (MODE == PM_RESIDENT)
    for (uint32_t s = s0; s < s2; ++s)
      copyWords(prog + a.schedLds[s], a.words + a.partOff[a.schedPart[s]], a.partWords[a.schedPart[s]]);
  __syncthreads();
  uint64_t xs = *a.c.xs;
  uint64_t barTicks = 0;
  const uint64_t tStart = __builtin_amdgcn_s_memrealtime();
  uint32_t bar = a.barrierBase;
  if (a.c.mbWrites) {  // testbench register writes, complete before any partition runs
    mailboxWrites(a.c, tid, nt);
    barTicks += gridBarrierShared(a.c.ctl + kCtlTree, ++bar, true);
  }
  // per-partition profile (SimEval.cu:894-983): ticks from the region's
  // start on this workgroup to the end of each partition, and to the end of
  // the region's grid barrier
  auto partDone = [&](uint32_t part, uint64_t t0, bool on) {
    if (on && tid == 0) {
      a.c.profUseful[part] += shaderClock() - t0;
      a.c.profOwner[part] = g;
    }
  };
  for (uint32_t c = 0; c < a.c.nCycles; ++c) {
    const uint64_t cycle = a.c.cycleBase + c;
    const bool pc = profCycle(a.c, c);
    if (tid == 0) stimulusDraws(xs, a.c.nIn, draws);
    __syncthreads();
    if (g == 0)
      for (uint32_t k = tid; k < a.c.nIn; k += nt) stShared8(a.c.reg + a.c.inIds[k], stimBit(draws, k));
    uint64_t t0 = pc ? shaderClock() : 0;
    for (uint32_t s = s0; s < s1; ++s) {
      w64Partition(a, a.schedPart[s], progOf<MODE>(a, prog, s), vp, lut, draws, cycle);
      partDone(a.schedPart[s], t0, pc);
    }
    barTicks += gridBarrierShared(a.c.ctl + kCtlTree, ++bar, true);
    if (pc && tid == 0) a.c.profWg[2 * g] += shaderClock() - t0;
    t0 = pc ? shaderClock() : 0;
    for (uint32_t s = s1; s < s2; ++s) {
      w64Partition(a, a.schedPart[s], progOf<MODE>(a, prog, s), vp, lut, draws, cycle);
      partDone(a.schedPart[s], t0, pc);
    }
    barTicks += gridBarrierShared(a.c.ctl + kCtlTree, ++bar, true);
    if (pc && tid == 0) a.c.profWg[2 * g + 1] += shaderClock() - t0;
  }
  if (tid == 0) {
    if (g == 0) *a.c.xs = xs;
    addBarrierTicks(a.c, barTicks, __builtin_amdgcn_s_memrealtime() - tStart);
  }
  if (tid < 64) finishLaunch(kernargCommon(), bar + 1);
}

namespace {
const void *w64Fn(int mode) {
  return mode == PM_RESIDENT ? reinterpret_cast<const void *>(w64Cycles<PM_RESIDENT>)
         : mode == PM_STAGED ? reinterpret_cast<const void *>(w64Cycles<PM_STAGED>)
                             : reinterpret_cast<const void *>(w64Cycles<PM_GLOBAL>);
}
}  // namespace

hipError_t launchWave64(const W64Args &args, int mode, uint32_t grid, uint32_t block, uint32_t lds,
                        hipStream_t stream) {
  W64Args a = args;
  void *params[] = {&a};
  return hipLaunchCooperativeKernel(w64Fn(mode), dim3(grid), dim3(block), params, lds, stream);
}

hipError_t wave64MaxBlocksPerCU(int *blocks, int mode, uint32_t block, uint32_t lds) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, w64Fn(mode), block, lds);
}

hipError_t wave64SetLds(uint32_t lds) {
  for (int m = 0; m < 3; ++m) {
    hipError_t e = hipFuncSetAttribute(w64Fn(m), hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

uint32_t wave64LutBytes() { return kLutBytes; }

}  // namespace tgs
