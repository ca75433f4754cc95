// Streamed cycle kernel for gfx950.
//
// Semantics are those of the reference evaluator (src/SimEval.cu:304-761 per
// micropartition, :765-864 per partition, :889-965 per cycle). Execution:
//   * persistent workgroups, one per partition slot: the host assigns
//     partitions per region (longest processing time) and the tasks of every
//     exec level to the compute waves the same way, and writes for each
//     workgroup a flat step stream of one cycle (exec levels, partition
//     begin / end, region barriers), every step ending in a workgroup barrier;
//   * each workgroup has NC compute waves and three loader waves. Loader k
//     owns every third step: it loads that step's 64-word program rows from
//     HBM (or the Infinity Cache) into registers three steps ahead (rows 0-5
//     of each compute wave's entry, the only rows the compute waves read
//     from the ring: 16 + 8 bytes per lane, a wave's rows are contiguous)
//     and writes them into a
//     two-slot LDS ring two steps before they are evaluated; the
//     compute waves read the next step's rows from the ring while they
//     evaluate the current step. No compute wave ever waits on a global load
//     for program words, and a loader only ever waits for its own single set
//     of loads;
//   * the LUT, the value pool and the stimulus draws live in LDS, program
//     rows are never resident, so partitions of any size fit and several
//     workgroups share a CU;
//   * one grid barrier per region on a two-level arrival tree; the register,
//     exchange and memory pools (the only data workgroups exchange) are read
//     and written with agent-coherent accesses, so the barrier needs no L2
//     write-back.
#include <hip/hip_runtime.h>

#include "kernel_args.h"

namespace tgs {

namespace {

constexpr uint32_t kLutBytesS = ((kLutSize + 15) / 16) * 16;

// Operand read without a branch: slots below 16 are immediates, but reading
// them is in bounds (the value pool is at least 16 bytes), so the load is
// unconditional and the immediate is selected afterwards.
__device__ __forceinline__ uint32_t rdOpS(const uint8_t *vp, uint32_t s) {
  const uint32_t v = vp[s];
  return s < 16 ? s : v;
}
// LUT read without a branch: positions past the table read the zero padding
// byte at kLutSize.
__device__ __forceinline__ uint32_t lutS(const uint8_t *lut, uint32_t pos) {
  return lut[pos < kLutSize ? pos : kLutSize];
}

// The rows and entry of one step for one compute wave.
struct Pf {
  uint32_t ex, ey;  // the entry (wave-uniform)
  uint32_t r0, r1, r2, r3, r4, r5;
};

// Two 32-lane LUT micropartitions in one wave (SimEval.cu:319-448).
__device__ __forceinline__ void lutPairS(const Pf &p, const uint32_t *__restrict__ W, uint8_t *vp,
                                         const uint8_t *lut) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t hb = lane & 32u;
  const uint32_t nMid = (p.ey >> 8) & 0xFFu;
  const uint32_t x = p.r0, y = p.r1, wl = p.r2;
  uint32_t val;
  {
    const uint32_t lt = x & 0xFFFFu, o0 = x >> 16, o1 = y & 0xFFFFu, o2 = y >> 16;
    const uint32_t c = rdOpS(vp, o2);
    const uint32_t a = rdOpS(vp, o0), b = rdOpS(vp, o1);
    const uint32_t lv = lutS(lut, lutIndex(lt, a, b, c));
    val = lt == 0 ? c : lv;
  }
  auto mid = [&](uint32_t w) {
    const uint32_t o2 = w & 63u, o1 = (w >> 6) & 63u, o0 = (w >> 12) & 63u, lt = w >> 18;
    const uint32_t s0 = __shfl(val, int((o0 & 31u) | hb));
    const uint32_t s1 = __shfl(val, int((o1 & 31u) | hb));
    const uint32_t s2 = __shfl(val, int((o2 & 31u) | hb));
    const uint32_t a = o0 < 32 ? o0 : s0, b = o1 < 32 ? o1 : s1, c = o2 < 32 ? o2 : s2;
    const uint32_t lv = lutS(lut, lutIndex(lt, a, b, c));
    val = lt == 0 ? c : lv;
  };
  if (nMid > 0) mid(p.r3);
  if (nMid > 1) mid(p.r4);
  if (nMid > 2) mid(p.r5);
  for (uint32_t m = 3; m < nMid; ++m) mid(W[p.ex + (3 + m) * kRow + lane]);
  uint32_t r;
  if (__builtin_amdgcn_readfirstlane(p.ey) & kLastDirect) {
    // row 2 is indexed by the lane holding the value: no shuffle
    const uint32_t k = wl >> 16;
    r = k == 0 ? val : k - 1;
  } else {
    const uint32_t sid = (wl >> 16) & 0xFFu;
    r = __shfl(val, int((sid & 31u) | hb));
    if (sid < 32) r = sid;
  }
  if (wl != 0xFFFFFFFFu) vp[wl & 0xFFFFu] = uint8_t(r);
}

// VecRead (SimEval.cu:449-481), one op per lane.
__device__ __forceinline__ void vecReadS(const Pf &p, uint8_t *vp, const uint8_t *consts) {
  const uint32_t lane = threadIdx.x & 63u;
  if (lane >= ((p.ey >> 8) & 0xFFFFu)) return;
  const uint32_t base = p.r0 & 0xFFFFu, len = p.r0 >> 16, offset = p.r1 & 0xFFFFu, isConst = p.r1 >> 16;
  const uint32_t i0 = rdOpS(vp, p.r2 & 0xFFFFu), i1 = rdOpS(vp, p.r2 >> 16);
  const uint32_t i2 = rdOpS(vp, p.r3 & 0xFFFFu), i3 = rdOpS(vp, p.r3 >> 16);
  const uint32_t off = ((i0 << 12) | (i1 << 8) | (i2 << 4) | i3) + offset;
  uint32_t v;
  if (off < len) v = isConst ? consts[base + off] : vp[base + off];
  else v = rdOpS(vp, p.r4 & 0xFFFFu);
  vp[p.r4 >> 16] = uint8_t(v);
}

// VecArithOrLogic (SimEval.cu:482-548), one op per lane.
__device__ __forceinline__ void vecOpS(const Pf &p, uint8_t *vp, const uint8_t *consts) {
  const uint32_t lane = threadIdx.x & 63u;
  const bool act = lane < ((p.ey >> 8) & 0xFFFFu);
  const uint32_t b1 = p.r0 & 0xFFFFu, b2 = p.r0 >> 16;
  const uint32_t len = p.r1 & 0xFFFFu, name = (p.r1 >> 16) & 0xFFu, cf = p.r1 >> 24, res = p.r2;
  if (__ballot(act && (len > 16 || cf != 0)) == 0) {
    // every op of the wave reads at most 16 nibbles from the value pool: the
    // 128-bit value is (high, low) with low its low 64 bits; ADD/SUB/MUL
    // results only keep low bits, compares use both halves
    if (!act) return;
    uint32_t e1, e2;
    uint64_t l1 = vpNibbles64(vp, b1, len, &e1), l2 = vpNibbles64(vp, b2, len, &e2);
    int64_t h1 = int64_t(e1), h2 = int64_t(e2);
    if (len > 0) {  // sign_extend_128 (SimEval.cu:222-228) from bit 4 len - 1 < 64
      const uint32_t bit = 4 * len - 1;
      const uint64_t ext = bit == 63 ? 0 : ~((uint64_t(1) << (bit + 1)) - 1);
      if ((l1 >> bit) & 1) {
        l1 |= ext;
        h1 = -1;
      }
      if ((l2 >> bit) & 1) {
        l2 |= ext;
        h2 = -1;
      }
    }
    if (name <= 2) {
      const uint64_t r = name == 0 ? l1 + l2 : name == 1 ? l1 - l2 : l1 * l2;
      vpPutNibbles(vp, res, len, r);
    } else {
      const bool eq = h1 == h2 && l1 == l2, lt = h1 < h2 || (h1 == h2 && l1 < l2);
      vp[res] = uint8_t(name == 3 ? eq : name == 4 ? lt : name == 5 ? (lt || eq) : 0);
    }
    return;
  }
  if (!act) return;
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

// MemRead (SimEval.cu:550-575), one op per lane.
__device__ __forceinline__ void memReadS(const Pf &p, uint8_t *vp, const CommonArgs &c) {
  const uint32_t lane = threadIdx.x & 63u;
  if (lane >= ((p.ey >> 8) & 0xFFFFu)) return;
  const uint64_t base = uint64_t(p.r0) | uint64_t(p.r1) << 32;
  const uint32_t en = rdOpS(vp, p.r2 & 0xFFFFu);
  if (en == 0) return;
  const uint64_t at = memAddress(vp, p.r2 >> 16, p.r3 >> 16, base);
  vp[p.r3 & 0xFFFFu] = at < c.memSize ? uint8_t(ldShared8(c.mem + at)) : uint8_t(0);
}

// L0 source (register, exchange or stimulus input).
__device__ __forceinline__ uint32_t l0Value(const CommonArgs &c, const uint64_t *draws, uint32_t src) {
  const uint32_t kind = src >> 30, idx = src & 0x3FFFFFFFu;
  if (kind == 0) return ldShared8(c.reg + idx);
  if (kind == 1) return ldShared8(c.exg + idx);
  if (kind == 2) return stimBit(draws, idx);
  return 0;
}

// Partition begin: L0 reads (SimEval.cu:582-622) by the compute threads of
// the partition's slot; each wave reads its own section of the block.
__device__ __forceinline__ void partBegin(const Pf &p, const CommonArgs &c, const uint32_t *__restrict__ W,
                                          uint8_t *vp, const uint64_t *draws, uint32_t &constBase) {
  if (((p.ey >> 4) & 7u) == kKindNone) return;  // no partition in this slot
  const uint32_t tid = threadIdx.x;
  const uint32_t nJ = __builtin_amdgcn_readfirstlane(__builtin_amdgcn_readlane(p.r0, S_NJ));
  constBase = __builtin_amdgcn_readlane(p.r0, S_CONST_BASE);
  const uint32_t lane = tid & 63u;
  if (nJ > 0) {
    // Entries 0 and 1 come from the staged rows; the rows of entries 2..7 are
    // loaded before any gather is issued, so the whole begin costs two global
    // round trips (rows, then values) instead of one per dependent group.
    constexpr uint32_t kE = 6;
    uint32_t src[kE], slot[kE];
#pragma unroll
    for (uint32_t u = 0; u < kE; ++u) {
      const bool in = 2 + u < nJ;
      src[u] = in ? W[p.ex + (1 + 2 * (2 + u)) * kRow + lane] : kL0None;
      slot[u] = in ? W[p.ex + (2 + 2 * (2 + u)) * kRow + lane] : 0u;
    }
    const bool a0 = (p.r1 >> 30) != 3, a1 = nJ > 1 && (p.r3 >> 30) != 3;
    const uint32_t v0 = a0 ? l0Value(c, draws, p.r1) : 0u;
    const uint32_t v1 = a1 ? l0Value(c, draws, p.r3) : 0u;
    uint32_t val[kE];
#pragma unroll
    for (uint32_t u = 0; u < kE; ++u) val[u] = (src[u] >> 30) != 3 ? l0Value(c, draws, src[u]) : 0u;
    if (a0) vp[p.r2] = uint8_t(v0);
    if (a1) vp[p.r4] = uint8_t(v1);
#pragma unroll
    for (uint32_t u = 0; u < kE; ++u)
      if ((src[u] >> 30) != 3) vp[slot[u]] = uint8_t(val[u]);
    // the rest (long L0 lists) four at a time: all row loads, then all gathers
    for (uint32_t j = 2 + kE; j < nJ; j += 4) {
      uint32_t src[4], slot[4], val[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        const bool in = j + u < nJ;
        src[u] = in ? W[p.ex + (1 + 2 * (j + u)) * kRow + lane] : kL0None;
        slot[u] = in ? W[p.ex + (2 + 2 * (j + u)) * kRow + lane] : 0u;
      }
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) val[u] = (src[u] >> 30) != 3 ? l0Value(c, draws, src[u]) : 0u;
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u)
        if ((src[u] >> 30) != 3) vp[slot[u]] = uint8_t(val[u]);
    }
  }
}

// Partition end: last level (SimEval.cu:625-761) by the compute threads of
// the partition's slot (thread `tid` of `nTc`).
__device__ __forceinline__ void partEnd(const Pf &p, const CommonArgs &c, const uint32_t *__restrict__ W,
                                        const uint8_t *vp, uint64_t cycle, uint32_t tid, uint32_t nTc) {
  if (((p.ey >> 4) & 7u) == kKindNone) return;  // no partition in this slot
  spillProbes(c, __builtin_amdgcn_readlane(p.r0, S_PART_ID), vp, tid, nTc);
  const uint32_t ewdc = __builtin_amdgcn_readlane(p.r0, S_EW_DC);
  if (ewdc >> 16) {
    const uint32_t exg = __builtin_amdgcn_readlane(p.r0, S_EW_EXG), dat = ewdc & 0xFFFFu;
    const uint32_t words = ((ewdc >> 16) + 3) / 4;
    for (uint32_t k = tid; k < words; k += nTc)
      stShared32(reinterpret_cast<uint32_t *>(c.exg) + (exg >> 2) + k, reinterpret_cast<const uint32_t *>(vp)[(dat >> 2) + k]);
    return;
  }
  const uint32_t rwdc = __builtin_amdgcn_readlane(p.r0, S_RW_DC);
  if (rwdc >> 16) {
    const uint32_t reg = __builtin_amdgcn_readlane(p.r0, S_RW_REG), dat = rwdc & 0xFFFFu;
    const uint32_t words = ((rwdc >> 16) + 3) / 4;
    for (uint32_t k = tid; k < words; k += nTc)
      stShared32(reinterpret_cast<uint32_t *>(c.reg) + (reg >> 2) + k, reinterpret_cast<const uint32_t *>(vp)[(dat >> 2) + k]);
  }
  const uint32_t nmw = __builtin_amdgcn_readlane(p.r0, S_NMW), npr = __builtin_amdgcn_readlane(p.r0, S_NPR);
  const uint32_t nst = __builtin_amdgcn_readlane(p.r0, S_NST);
  if (nmw | npr | nst) {
    const uint32_t *L = W + p.ex + __builtin_amdgcn_readlane(p.r0, S_OFF_LAST);
    for (uint32_t k = tid; k < nmw; k += nTc) {
      const uint32_t *o = L + 4 * k;
      const uint32_t en = rdOpS(vp, o[3] & 0xFFFFu);
      if (en != 0) {
        const uint32_t dat = rdOpS(vp, o[2] >> 16);
        const uint64_t at = memAddress(vp, o[2] & 0xFFFFu, o[3] >> 16, uint64_t(o[0]) | uint64_t(o[1]) << 32);
        if (at < c.memSize) stShared8(c.mem + at, dat);
      }
    }
    L += 4 * nmw;
    if (c.printEnable)
      for (uint32_t k = tid; k < npr; k += nTc)
        if (rdOpS(vp, L[k] & 0xFFFFu) != 0) logPrint(c, cycle, L[k] >> 16);
    L += npr;
    for (uint32_t k = tid; k < nst; k += nTc)
      if (rdOpS(vp, L[k] & 0xFFFFu) != 0) raiseStop(c, cycle);
  }
}

// Rows 0..5 of each compute wave: rows 0-3 as one 16-byte load per lane and
// rows 4-5 as one 8-byte load per lane (native vector types: HIP's uint4
// wrapper keeps the set out of registers).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
template <int NC>
struct RowSet {
  u32x4 lo[NC];  // rows 0-3
  u32x2 hi[NC];  // rows 4-5
};

// Loader: load the rows of step record `rec` for every compute wave. The
// record is read with one LDS load (lane i holds word i) and broadcast with
// readlane; a wave's rows are contiguous from its entry's x, so each load is
// a uniform base plus the lane offset.
template <int NC>
__device__ __forceinline__ uint32_t recWords(const uint32_t *rec, uint32_t lane) {
  constexpr uint32_t words = 1 + 2 * NC;
  return rec[lane < words ? lane : 0];
}
// rv: the record's words (recWords)
template <int NC>
__device__ __forceinline__ void loadRows(RowSet<NC> &s, uint32_t rv, const uint32_t *__restrict__ W,
                                         uint32_t lane) {
  // unconditional loads: making them depend on the rows an entry needs (exec
  // masks or a uniform skip) measured 20-60% slower per step
#pragma unroll
  for (int w = 0; w < NC; ++w) {
    const uint32_t *x = W + __builtin_amdgcn_readlane(rv, 1 + 2 * w);
    s.lo[w] = reinterpret_cast<const u32x4 *>(x)[lane];
    s.hi[w] = reinterpret_cast<const u32x2 *>(x + 4 * kRow)[lane];
  }
}

template <int NC>
__device__ __forceinline__ void storeRows(const RowSet<NC> &s, uint32_t *slot, uint32_t lane) {
#pragma unroll
  for (int w = 0; w < NC; ++w) {
    uint32_t *rows = slot + w * kStreamRows * 64;
    reinterpret_cast<u32x4 *>(rows)[lane] = s.lo[w];
    reinterpret_cast<u32x2 *>(rows + 4 * kRow)[lane] = s.hi[w];
  }
}

}  // namespace

// Two ring slots: during step i the compute waves read step i + 1 while a
// loader wave writes step i + 2.
constexpr uint32_t kSlots = 2;
// Loader waves: loader k owns the steps i with i % kLoaders == k; it writes
// step i + 2 during step i and then loads step i + 2 + kLoaders, so each of
// its loads has kLoaders steps to arrive and only one load set is ever in
// flight per wave (the wait before the stores is a plain wait for all).
constexpr uint32_t kLoaders = 3;

// At least six waves per SIMD (HIP's second launch bound): two NC = 8 or
// three NC = 4 workgroups per CU, which keeps every partition slot of a
// 1M-LUT4 design co-resident.
// PROF: the per-partition profile's hooks are compiled in (the runtime
// launches this instantiation only while the profile is on).
template <int NC, bool PROF>
__global__ __launch_bounds__((NC + kLoaders) * 64, 6) void streamCycles(StreamArgs a) {
  extern __shared__ __align__(16) uint8_t sLds[];
  constexpr uint32_t nTc = NC * 64;
  constexpr uint32_t rec = 1 + 2 * NC;
  constexpr uint32_t slotWords = NC * kStreamRows * 64;
  __shared__ __align__(16) uint32_t ring[kSlots * slotWords];
  uint8_t *lut = sLds;
  uint8_t *vpAll = sLds + kLutBytesS;  // one value pool per partition slot
  uint64_t *draws = reinterpret_cast<uint64_t *>(vpAll + a.slots * a.vpBytes);
  uint32_t *steps = reinterpret_cast<uint32_t *>(draws + kMaxDraws);
  const uint32_t g = blockIdx.x, tid = threadIdx.x, lane = tid & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  const uint32_t *__restrict__ W = a.words;
  const CommonArgs &c = a.c;
  const uint32_t nSteps = a.stepCount[g];
  {
    const uint32_t *src = a.steps + size_t(a.stepBase[g]) * rec;
    for (uint32_t k = tid; k < nSteps * rec; k += blockDim.x) steps[k] = src[k];
    for (uint32_t k = tid; k < kLutBytesS / 4; k += blockDim.x)
      reinterpret_cast<uint32_t *>(lut)[k] = reinterpret_cast<const uint32_t *>(c.lut)[k];
  }
  __syncthreads();
  auto recOf = [&](uint32_t j) { return steps + (j % nSteps) * rec; };
  if (wave >= NC) {
    const uint32_t k = wave - NC;
    RowSet<NC> rs;
    // steps 0 and 1 go straight into the ring; step 2 + k is this loader's first load
    if (k < kSlots) {
      loadRows<NC>(rs, recWords<NC>(recOf(k), lane), W, lane);
      storeRows<NC>(rs, ring + k * slotWords, lane);
    }
    loadRows<NC>(rs, recWords<NC>(recOf(2 + k), lane), W, lane);
    __syncthreads();  // start barrier
    if (c.mbWrites) {  // the workgroup barriers of the mailbox's grid barrier
      __syncthreads();
      __syncthreads();
    }
    // step counters kept without divisions: jk = j % kLoaders, si = j %
    // nSteps, ln = the record index of this loader's next load
    uint32_t j = 0, si = 0, cyc = 0, jk = 0, ln = (2 + k + kLoaders) % nSteps;
    // each step's type is read a step ahead into a VGPR (an opaque
    // per-lane zero in the address keeps the compiler from waiting for it
    // at once) and made scalar when that step starts
    uint32_t zv;
    asm volatile("v_mov_b32 %0, 0" : "=v"(zv));
    uint32_t typeV = steps[zv];
    for (;;) {
      if (si == 0) {
        if (cyc == c.nCycles) break;
        __syncthreads();  // cycle start (stimulus draws)
      }
      const bool region = __builtin_amdgcn_readfirstlane(typeV) == S_REGION;
      typeV = steps[(si + 1 == nSteps ? 0u : si + 1) * rec + zv];
      if (jk == k) {
        // the next load's record is read before the stores (LDS reads return
        // in order: read after them, it would wait for them)
        const uint32_t rv = recWords<NC>(steps + ln * rec, lane);
        storeRows<NC>(rs, ring + ((j + 2) % kSlots) * slotWords, lane);
        // the set's registers are reused by the next loads: keep the stores first
        __builtin_amdgcn_sched_barrier(0);
        loadRows<NC>(rs, rv, W, lane);
        for (ln += kLoaders; ln >= nSteps;) ln -= nSteps;
      }
      if (++jk == kLoaders) jk = 0;
      if (region) {
        __syncthreads();  // the two workgroup barriers of the grid barrier
        __syncthreads();
      } else {
        __syncthreads();
      }
      ++j;
      if (++si == nSteps) {
        si = 0;
        ++cyc;
      }
    }
    return;
  }
  // Compute waves issue ahead of the loaders when they share a SIMD (wave
  // issue priority; the loaders' work is never on the critical path):
  // 77.7-78.0 against 80.4-81.6 us per cycle on the 1M-LUT4 design.
  __builtin_amdgcn_s_setprio(3);
  // compute waves: in a split step wave `wave` serves partition slot
  // wave / (NC / slots), otherwise every wave serves the one partition
  const uint32_t slotWaves = NC / a.slots;
  const uint32_t splitSlot = wave / slotWaves;
  uint64_t xs = *c.xs;
  uint64_t barTicks = 0;
  const uint64_t tStart = __builtin_amdgcn_s_memrealtime();
  // per-partition profile (SimEval.cu:894-983): region start on this
  // workgroup, the region's index inside the cycle
  uint64_t regionT0 = 0;
  uint32_t region = 0;
  bool pc = false;
  uint32_t bar = a.barrierBase;
  __syncthreads();  // start barrier: steps 0 and 1 are in the ring
  if (c.mbWrites) {  // testbench register writes, complete before any partition runs
    mailboxWrites(c, tid, nTc);
    barTicks += gridBarrierShared(c.ctl + kCtlTree, ++bar, true);
  }
  const uint32_t *mySlot = ring + wave * kStreamRows * 64 + lane;
  // record ri of the cycle's steps, ring slot of step j
  auto readStep = [&](Pf &p, uint32_t ri, uint32_t j) {
    const uint32_t *r = steps + ri * rec;
    const uint32_t *sl = mySlot + (j % kSlots) * slotWords;
    p.ex = r[1 + 2 * wave];
    p.ey = r[2 + 2 * wave] | r[0];  // the step type rides in the entry's unused low nibble
    p.r0 = sl[0];
    p.r1 = sl[64];
    p.r2 = sl[128];
    p.r3 = sl[192];
    p.r4 = sl[256];
    p.r5 = sl[320];
  };
  Pf cur, nxt;
  readStep(cur, 0, 0);
  uint32_t j = 0, si = 0, cyc = 0, constBase = 0;
  for (;;) {
    if (si == 0) {
      if (cyc == c.nCycles) break;
      if (tid == 0) stimulusDraws(xs, c.nIn, draws);
      __syncthreads();
      if (g == 0)
        for (uint32_t k = tid; k < c.nIn; k += nTc) stShared8(c.reg + c.inIds[k], stimBit(draws, k));
      if (PROF) {
        pc = profCycle(c, cyc);
        region = 0;
        if (pc) regionT0 = shaderClock();
      }
    }
    // A run of LUT steps (most steps of a cycle) goes round a loop of its
    // own: the general step dispatch below costs more scalar work per step.
    // Every wave still meets one workgroup barrier per step, as there.
    bool wrapped = false;
    while ((cur.ey & 0x7Fu) == (S_LEVEL | T_LUT2 << 4)) {
      readStep(nxt, si + 1 == nSteps ? 0u : si + 1, j + 1);
      const bool split = (__builtin_amdgcn_readfirstlane(cur.ey) & kSplit) != 0;
      lutPairS(cur, W, vpAll + (split ? splitSlot : 0u) * a.vpBytes, lut);
      __syncthreads();
      cur = nxt;
      ++j;
      if (++si == nSteps) {
        si = 0;
        ++cyc;
        wrapped = true;
        break;
      }
    }
    if (wrapped) continue;
    readStep(nxt, si + 1 == nSteps ? 0u : si + 1, j + 1);
    const uint32_t type = cur.ey & 15u, kind = (cur.ey >> 4) & 7u;
    const bool split = (__builtin_amdgcn_readfirstlane(cur.ey) & kSplit) != 0;
    const uint32_t sw = split ? slotWaves : uint32_t(NC);  // waves serving this wave's partition
    const uint32_t slot = split ? splitSlot : 0u;
    uint8_t *vp = vpAll + slot * a.vpBytes;
    const uint32_t ts = (wave - slot * sw) * 64 + lane, nTs = sw * 64;
    if (type == S_LEVEL) {
      if (kind == T_LUT2) lutPairS(cur, W, vp, lut);
      else if (kind == T_VECREAD) vecReadS(cur, vp, c.consts + constBase);
      else if (kind == T_VECOP) vecOpS(cur, vp, c.consts + constBase);
      else if (kind == T_MEMREAD) memReadS(cur, vp, c);
      __syncthreads();
    } else if (type == S_PART) {
      partBegin(cur, c, W, vp, draws, constBase);
      __syncthreads();
    } else if (type == S_PEND) {
      partEnd(cur, c, W, vp, c.cycleBase + cyc, ts, nTs);
      __syncthreads();
      if (PROF && pc && ts == 0 && kind != kKindNone) {
        const uint32_t part = __builtin_amdgcn_readlane(cur.r0, S_PART_ID);
        c.profUseful[part] += shaderClock() - regionT0;
        c.profOwner[part] = g;
      }
    } else {
      barTicks += gridBarrierShared(c.ctl + kCtlTree, ++bar, true);
      if (PROF && pc) {
        if (tid == 0) c.profWg[2 * g + region] += shaderClock() - regionT0;
        regionT0 = shaderClock();
        region ^= 1u;
      }
    }
    cur = nxt;
    ++j;
    if (++si == nSteps) {
      si = 0;
      ++cyc;
    }
  }
  if (tid == 0) {
    if (g == 0) *c.xs = xs;
    addBarrierTicks(c, barTicks, __builtin_amdgcn_s_memrealtime() - tStart);
  }
  if (wave == 0) finishLaunch(kernargCommon(), bar + 1);
}

namespace {
const void *streamFn(uint32_t nc, bool prof = false) {
  if (prof)
    return nc == 2 ? reinterpret_cast<const void *>(streamCycles<2, true>)
           : nc == 8 ? reinterpret_cast<const void *>(streamCycles<8, true>)
                     : reinterpret_cast<const void *>(streamCycles<4, true>);
  return nc == 2 ? reinterpret_cast<const void *>(streamCycles<2, false>)
         : nc == 8 ? reinterpret_cast<const void *>(streamCycles<8, false>)
                   : reinterpret_cast<const void *>(streamCycles<4, false>);
}
}  // namespace

bool streamSupportsWaves(uint32_t nc) { return nc == 2 || nc == 4 || nc == 8; }

// Dynamic LDS (the ring is static): LUT, value pools, draws, step records.
uint32_t streamLdsBytes(uint32_t nc, uint32_t slots, uint32_t vpBytes, uint32_t maxSteps) {
  return kLutBytesS + slots * vpBytes + kMaxDraws * 8 + maxSteps * (1 + 2 * nc) * 4;
}

uint32_t streamStaticLdsBytes(uint32_t nc) { return kSlots * nc * kStreamRows * 64 * 4; }

uint32_t streamBlock(uint32_t nc) { return (nc + kLoaders) * 64; }

hipError_t launchStream(const StreamArgs &args, uint32_t nc, uint32_t grid, uint32_t lds, hipStream_t stream) {
  StreamArgs a = args;
  void *params[] = {&a};
  return hipLaunchCooperativeKernel(streamFn(nc, args.c.profUseful != nullptr), dim3(grid),
                                    dim3((nc + kLoaders) * 64), params, lds, stream);
}

// The smaller of the two instantiations' figures, so a grid sized by it is
// co-resident with the profile on or off.
hipError_t streamMaxBlocksPerCU(int *blocks, uint32_t nc, uint32_t lds) {
  int b0 = 0, b1 = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b0, streamFn(nc, false), (nc + kLoaders) * 64, lds);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b1, streamFn(nc, true), (nc + kLoaders) * 64, lds);
  *blocks = std::min(b0, b1);
  return e;
}

hipError_t streamSetLds(uint32_t lds) {
  for (uint32_t nc : {2u, 4u, 8u})
    for (bool prof : {false, true}) {
      hipError_t e = hipFuncSetAttribute(streamFn(nc, prof), hipFuncAttributeMaxDynamicSharedMemorySize, lds);
      if (e != hipSuccess) return e;
    }
  return hipSuccess;
}

}  // namespace tgs
