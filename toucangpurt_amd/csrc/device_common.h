// Device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_layout.h"

namespace tgs {

// Control block of one instance (device memory, u32 words).
enum Ctl : uint32_t {
  C_STOP = 0,        // sticky stop flag (SimEval.cu:194 shouldStop)
  C_LOGCOUNT = 2,    // u64: print-log entries appended since the host last drained the log
  C_STOPCYC = 4,     // u64: first cycle a stop op fired (UINT64_MAX if none)
  C_BARTICKS = 6,    // u64: real-time ticks spent in grid barriers, summed over workgroups and launches
  C_TOTTICKS = 8,    // u64: real-time ticks of the cycle loop, summed over workgroups and launches
  C_WORDS = 16
};
// The grid-barrier tree (treeArrive) of an instance's persistent launches
// follows the control words at word kCtlTree of the same allocation (kept
// across launches of one grid size).
constexpr uint32_t kBarGroups = 16;  // first-level groups: workgroup g joins group g % kBarGroups
constexpr uint32_t kBarLine = 64;    // words per counter (one 256-byte line each)
constexpr uint32_t kBarTreeWords = kBarLine * (1 + 2 * kBarGroups);
constexpr uint32_t kCtlTree = 64;
constexpr uint32_t kCtlAllocWords = kCtlTree + kBarTreeWords;
// Control words a launch copies to its mailbox (stop flag through the ticks).
constexpr uint32_t kMbCtlWords = C_TOTTICKS + 2;

constexpr uint32_t kMaxDraws = 64;  // up to 4096 stimulus inputs

// Arguments every cycle kernel takes.
struct CommonArgs {
  const uint8_t *lut;      // 4642-entry table, zero padded to a word multiple
  const uint8_t *consts;   // constant-vector pools of all partitions
  uint8_t *reg, *mem, *exg;
  uint32_t regSize, memSize, exgSize;
  uint64_t *xs;            // xorshift64 stimulus state
  const uint32_t *inIds;   // stimulus input register ids
  uint32_t nIn;
  uint32_t *ctl;
  uint64_t *logCycle;
  uint32_t *logMsg;
  uint32_t logCap;
  uint32_t printEnable;
  uint64_t cycleBase;      // cycles completed before this launch
  uint32_t nCycles;
  // signal probes (null when none): per partition a CSR range of value-pool
  // slots whose end-of-visit values go to probeVal (CSR order)
  const uint32_t *probeOff, *probeSlot;
  uint8_t *probeVal;
  // per-partition profile (null when off): launch-local cycles
  // [profStart, profStart + profCount) add, per partition, the shader-clock
  // ticks from its region's start on its workgroup to its end (profUseful)
  // and, per workgroup and region, the ticks to the end of the region's grid
  // barrier (profWg[2 * workgroup + region]; profOwner names the workgroup)
  uint64_t *profUseful, *profWg;
  uint32_t *profOwner;  // per partition: the workgroup that ran it
  uint32_t profStart, profCount;
  // testbench mailbox (pinned host memory; counts 0 and mbCtl null when
  // none): register mbWriteId[k] takes mbWriteVal[k] before the first cycle,
  // mbReadVal[k] takes register mbReadId[k] after the last, and mbCtl the
  // control words 0 .. kMbCtlWords - 1, so a testbench step is one launch
  const uint32_t *mbWriteId;
  const uint8_t *mbWriteVal;
  uint32_t mbWrites;
  const uint32_t *mbReadId;
  uint8_t *mbReadVal;
  uint32_t mbReads;
  uint32_t *mbCtl;
};

// Whether launch-local cycle `cyc` is inside the profile window.
__device__ __forceinline__ bool profCycle(const CommonArgs &a, uint32_t cyc) {
  return a.profUseful != nullptr && cyc >= a.profStart && cyc - a.profStart < a.profCount;
}

// Shader clock (the counterpart of clock64).
__device__ __forceinline__ uint64_t shaderClock() { return __builtin_amdgcn_s_memtime(); }

// End of workgroup `wg`'s barrier-measured loop: add its barrier-wait and
// loop ticks to the launch totals.
__device__ __forceinline__ void addBarrierTicks(const CommonArgs &a, uint64_t barTicks, uint64_t loopTicks) {
  atomicAdd(reinterpret_cast<unsigned long long *>(&a.ctl[C_BARTICKS]), (unsigned long long)barTicks);
  atomicAdd(reinterpret_cast<unsigned long long *>(&a.ctl[C_TOTTICKS]), (unsigned long long)loopTicks);
}

// xorshift64, Marsaglia (2003), shifts (13, 7, 17).
__device__ __forceinline__ uint64_t xorshift64(uint64_t x) {
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  return x;
}

// Advance the stimulus generator one cycle; every workgroup runs the same
// sequence so no cross-workgroup traffic is needed for the draws.
__device__ __forceinline__ void stimulusDraws(uint64_t &xs, uint32_t nIn, uint64_t *draws) {
  for (uint32_t d = 0; d * 64 < nIn; ++d) {
    xs = xorshift64(xs);
    draws[d] = xs;
  }
}

__device__ __forceinline__ uint32_t stimBit(const uint64_t *draws, uint32_t input) {
  return uint32_t((draws[input >> 6] >> (input & 63)) & 1u);
}

// Print op (SimEval.cu:696-704): append (cycle, message) to the instance log.
__device__ __forceinline__ void logPrint(const CommonArgs &a, uint64_t cycle, uint32_t msg) {
  const uint64_t k = atomicAdd(reinterpret_cast<unsigned long long *>(&a.ctl[C_LOGCOUNT]), 1ull);
  if (k < a.logCap) {
    a.logCycle[k] = cycle;
    a.logMsg[k] = msg;
  }
}

// Stop op (SimEval.cu:706-722): sticky flag plus the first cycle it fired.
__device__ __forceinline__ void raiseStop(const CommonArgs &a, uint64_t cycle) {
  atomicOr(&a.ctl[C_STOP], 1u);
  atomicMin(reinterpret_cast<unsigned long long *>(&a.ctl[C_STOPCYC]), (unsigned long long)cycle);
}

// Signal probes (SimDebugInfo::signalDebugInfo locations): after the exec
// levels of partition `part`, copy the listed value-pool slots out, `tid` of
// `nt` threads. The values are those the partition computed this cycle.
__device__ __forceinline__ void spillProbes(const CommonArgs &a, uint32_t part, const uint8_t *vp, uint32_t tid,
                                            uint32_t nt) {
  if (a.probeOff == nullptr) return;
  const uint32_t b = a.probeOff[part], e = a.probeOff[part + 1];
  for (uint32_t k = b + tid; k < e; k += nt) a.probeVal[k] = vp[a.probeSlot[k]];
}

// Operand of a VecArithOrLogic op read from an LDS value pool (4-byte
// aligned): `len` (<= 32) nibbles from slot `base`, value = OR of byte k
// shifted left by 4k, truncated to 128 bits (SimEval.cu:491-503). The
// covering aligned words are loaded together (one LDS round trip instead of
// one per nibble) and realigned to `base` with byte funnel shifts; each word
// of four bytes b0..b3 then becomes b0 | b1 << 4 | b2 << 8 | b3 << 12 with
// word-wide operations (exact for any byte value, as the reference's OR).
__device__ __forceinline__ unsigned __int128 vpNibbles(const uint8_t *vp, uint32_t base, uint32_t len) {
  // Nine words from the one holding `base`: at most 35 bytes past the
  // operand's last nibble, which stays inside the LDS allocation (the value
  // pool is followed by the stimulus draws); bytes past `len` are masked.
  const uint32_t *w = reinterpret_cast<const uint32_t *>(vp) + (base >> 2);
  const uint32_t skip = base & 3u;
  uint32_t x[9];
#pragma unroll
  for (uint32_t i = 0; i < 9; ++i) x[i] = w[i];
  uint32_t c[8];
#pragma unroll
  for (uint32_t i = 0; i < 8; ++i) {
    uint32_t y = __builtin_amdgcn_alignbyte(x[i + 1], x[i], skip);
    const int rem = int(len) - int(4 * i);  // bytes of this word inside the operand
    y &= rem >= 4 ? ~0u : rem <= 0 ? 0u : (1u << (8 * rem)) - 1u;
    const uint32_t even = (y & 0xFFu) | ((y >> 8) & 0xFF00u);        // b0 | b2 << 8
    const uint32_t odd = ((y >> 8) & 0xFFu) | ((y >> 16) & 0xFF00u);  // b1 | b3 << 8
    c[i] = even | (odd << 4);  // 20 bits at bit 16 i of the value
  }
  const uint32_t z0 = c[0] | (c[1] << 16);
  const uint32_t z1 = (c[1] >> 16) | c[2] | (c[3] << 16);
  const uint32_t z2 = (c[3] >> 16) | c[4] | (c[5] << 16);
  const uint32_t z3 = (c[5] >> 16) | c[6] | (c[7] << 16);
  return ((unsigned __int128)z3 << 96) | ((unsigned __int128)z2 << 64) | ((unsigned __int128)z1 << 32) | z0;
}

// vpNibbles for len <= 16: the low 64 bits of the same value, and in *high
// bits 64-67, which only an out-of-range byte 15 (value above 15) reaches.
__device__ __forceinline__ uint64_t vpNibbles64(const uint8_t *vp, uint32_t base, uint32_t len, uint32_t *high) {
  const uint32_t *w = reinterpret_cast<const uint32_t *>(vp) + (base >> 2);
  const uint32_t skip = base & 3u;
  uint32_t x[5];
#pragma unroll
  for (uint32_t i = 0; i < 5; ++i) x[i] = w[i];
  uint32_t c[4];
#pragma unroll
  for (uint32_t i = 0; i < 4; ++i) {
    uint32_t y = __builtin_amdgcn_alignbyte(x[i + 1], x[i], skip);
    const int rem = int(len) - int(4 * i);
    y &= rem >= 4 ? ~0u : rem <= 0 ? 0u : (1u << (8 * rem)) - 1u;
    const uint32_t even = (y & 0xFFu) | ((y >> 8) & 0xFF00u);
    const uint32_t odd = ((y >> 8) & 0xFFu) | ((y >> 16) & 0xFF00u);
    c[i] = even | (odd << 4);
  }
  *high = c[3] >> 16;
  const uint32_t z0 = c[0] | (c[1] << 16);
  const uint32_t z1 = (c[1] >> 16) | c[2] | (c[3] << 16);
  return (uint64_t(z1) << 32) | z0;
}

// Result of a VecArithOrLogic ADD/SUB/MUL: nibbles 0..len-1 of r to value
// slots res..res+len-1 (SimEval.cu:527-534).
__device__ __forceinline__ void vpPutNibbles(uint8_t *vp, uint32_t res, uint32_t len, unsigned __int128 r) {
  for (uint32_t k = 0; k < len; k += 8) {  // eight nibbles per 32-bit word of r
    const uint32_t w = uint32_t(r);
    r >>= 32;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j)
      if (k + j < len) vp[uint16_t(res + k + j)] = uint8_t((w >> (4 * j)) & 15u);
  }
}

// The same for results that fit 64 bits (len <= 16). When every active lane
// of the wave has a multiple of four nibbles, four slots go under one lane
// mask at immediate offsets from the first slot; otherwise one slot at a
// time. Slots res..res+len-1 lie inside the value pool (checked when the
// image is built), so no slot index wraps.
__device__ __forceinline__ void vpPutNibbles(uint8_t *vp, uint32_t res, uint32_t len, uint64_t r) {
  uint8_t *d = vp + res;
  if (__ballot((len & 3u) != 0) == 0) {
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q)
      if (4 * q < len) {
        const uint32_t w = uint32_t(r >> (16 * q));
        d[4 * q] = uint8_t(w & 15u);
        d[4 * q + 1] = uint8_t((w >> 4) & 15u);
        d[4 * q + 2] = uint8_t((w >> 8) & 15u);
        d[4 * q + 3] = uint8_t((w >> 12) & 15u);
      }
    return;
  }
#pragma unroll
  for (uint32_t j = 0; j < 16; ++j)
    if (j < len) d[j] = uint8_t((r >> (4 * j)) & 15u);
}

// LUT lookup (SimEval.cu:349-356,406-413) with the u16 index arithmetic of the
// reference; positions past the 4642-entry table read 0.
__device__ __forceinline__ uint32_t lutIndex(uint32_t lut, uint32_t a, uint32_t b, uint32_t c) {
  return (lut + (((a & 0xFFFFu) << 8) | (b << 4) | c)) & 0xFFFFu;
}

// Memory address of SimEval.cu:561-570 / 680-690 from 8 value-pool nibbles.
__device__ __forceinline__ uint64_t memAddress(const uint8_t *vp, uint32_t addrVec, uint32_t multi, uint64_t base) {
  uint32_t addr = 0;
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) addr |= uint32_t(vp[addrVec + j]) << (4 * j);
  if (multi) addr <<= 2;
  return base + addr;
}

// Two-level arrival on a grid-barrier tree (kBarTreeWords words, 256-byte
// aligned, all zero before its first barrier): workgroup g adds to the
// counter of group g % kBarGroups, the last arrival of a group adds to the
// top counter, and the top's last arrival writes the barrier index b to
// every group's release word, which that group's workgroups poll
// (treeWait). Each counter and word has a 256-byte line of its own. On
// MI355X with no work between barriers this measured 1.39 us per barrier
// for 256 workgroups and 1.58 us for 512, against 3.64 and 8.30 us for one
// counter that every workgroup adds to and polls. Counters only grow
// (barrier b completes a group at b * its members and the top at b *
// groups), so barriers 1, 2, ... need no reset in between, also across
// launches that carry b on with the same grid size. Returns whether this
// arrival completed barrier b (the last workgroup).
__device__ __forceinline__ bool treeArrive(uint32_t *tree, uint32_t b) {
  const uint32_t n = gridDim.x, groups = n < kBarGroups ? n : kBarGroups, grp = blockIdx.x % kBarGroups;
  const uint32_t members = (n - grp + kBarGroups - 1) / kBarGroups;
  if (__hip_atomic_fetch_add(tree + kBarLine * (1 + grp), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u !=
      b * members)
    return false;
  if (__hip_atomic_fetch_add(tree, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u != b * groups) return false;
  for (uint32_t k = 0; k < groups; ++k)
    __hip_atomic_store(tree + kBarLine * (1 + kBarGroups + k), b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}
__device__ __forceinline__ void treeWait(const uint32_t *tree, uint32_t b) {
  const uint32_t *rw = tree + kBarLine * (1 + kBarGroups + blockIdx.x % kBarGroups);
  while (__hip_atomic_load(rw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < b) __builtin_amdgcn_s_sleep(1);
}

// Every wave first waits for its own outstanding global stores (`drain`; a
// wave that has none, such as a pure loader, may skip it); after the
// workgroup barrier one wave publishes them device-wide (agent-scope release,
// an L2 write-back), arrives, waits, and re-acquires once (agent-scope
// acquire, an L1/L2 invalidate that every wave of the CU then sees).
// Register, exchange and memory pools written before the barrier are
// visible to every XCD after it. `b` = the barrier's index on `tree`.
__device__ __forceinline__ uint64_t gridBarrier(uint32_t *tree, uint32_t b, bool measure, bool drain = true) {
  // every wave completes its own global stores first (s_waitcnt vmcnt(0))
  if (drain) __builtin_amdgcn_s_waitcnt(0x0F70);
  __syncthreads();
  uint64_t waited = 0;
  if (threadIdx.x == 0) {
    const uint64_t t0 = measure ? __builtin_amdgcn_s_memrealtime() : 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    // The fence's L2 write-back (buffer_wbl2) must complete before the
    // arrival: the compiler follows it with an LDS-only wait when it sees no
    // vector memory operation outstanding, so the arrival's atomic could be
    // performed first and a workgroup released by it read pool lines the
    // write-back had not yet reached memory with.
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    treeArrive(tree, b);
    treeWait(tree, b);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    // The fence's L2 invalidate (buffer_inv) orders this wave's later loads
    // only; the other waves load after the workgroup barrier below, so it
    // must have completed before this wave reaches that barrier (without
    // the wait they could read pool lines the invalidate had not yet
    // dropped: intermittently wrong registers with several partitions per
    // workgroup).
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (measure) waited = __builtin_amdgcn_s_memrealtime() - t0;
  }
  __syncthreads();
  return waited;
}

// This workgroup's share of the mailbox's register writes (tid of nt
// threads). The caller passes a grid barrier before any partition runs: a
// region-0 partition may write a register the mailbox writes, and no
// workgroup may still be writing the mailbox's value then.
__device__ __forceinline__ void mailboxWrites(const CommonArgs &a, uint32_t tid, uint32_t nt) {
  for (uint32_t k = blockIdx.x * nt + tid; k < a.mbWrites; k += gridDim.x * nt)
    __hip_atomic_store(a.reg + a.mbWriteId[k], a.mbWriteVal[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The CommonArgs at the start of the running kernel's argument block (every
// persistent kernel's argument struct begins with one), through a pointer
// the compiler cannot relate to its entry loads: what a kernel reads through
// it after its cycle loop is loaded there, not held in scalar registers
// through the loop.
__device__ __forceinline__ CommonArgs kernargCommon() {
  const CommonArgs *p = (const CommonArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(p));
  return *p;
}

// End of a launch, called by one whole wave of every workgroup after the
// workgroup has added its ticks: one more arrival on the instance's barrier
// tree, as barrier b (one past the launch's last grid barrier), with no
// wait. The workgroup whose arrival completes it is the last: it sees every
// workgroup's stores and ticks and copies the mailbox reads and the control
// words to the mailbox. Everything it orders is an agent-scope atomic or an
// agent-coherent access (the tick sums, the pools after the last grid
// barrier), so waiting for the wave's own atomics to complete (vmcnt(0))
// stands in for a release fence, which would write back the XCD's L2.
__device__ __forceinline__ void finishLaunch(const CommonArgs &a, uint32_t b) {
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t last = 0;
  if (lane == 0) {
    __builtin_amdgcn_s_waitcnt(0x0F70);
    last = treeArrive(a.ctl + kCtlTree, b);
  }
  if (!__builtin_amdgcn_readfirstlane(last)) return;
  for (uint32_t k = lane; k < a.mbReads; k += 64)
    a.mbReadVal[k] = __hip_atomic_load(a.reg + a.mbReadId[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (a.mbCtl != nullptr && lane < kMbCtlWords)
    a.mbCtl[lane] = __hip_atomic_load(&a.ctl[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Cross-workgroup pools (register, exchange and memory pools) accessed with
// agent-coherent loads and stores: relaxed agent-scope atomics, which gfx950
// performs with sc1 (a store is written through to the device coherence
// point, a load does not hit another XCD's stale L2 line). A kernel that
// touches those pools only this way can use gridBarrierShared.
__device__ __forceinline__ uint32_t ldShared8(const uint8_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stShared8(uint8_t *p, uint32_t v) {
  __hip_atomic_store(p, uint8_t(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stShared32(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier for a kernel whose data shared between workgroups moves only
// through ldShared8 / stShared8 / stShared32: every wave waits for its own
// stores to complete (vmcnt(0); a completed sc1 store is visible device-wide),
// the workgroup synchronizes, one thread arrives on the tree and waits. No
// L2 write-back before the arrival and no invalidate after it: gridBarrier's
// agent-scope release writes back every dirty line of the XCD's L2, which on
// the 1M-LUT4 design held the last arrival back by 5-10 us per region.
// `b` = the barrier's index on `tree`.
__device__ __forceinline__ uint64_t gridBarrierShared(uint32_t *tree, uint32_t b, bool measure) {
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __syncthreads();
  uint64_t waited = 0;
  if (threadIdx.x == 0) {
    const uint64_t t0 = measure ? __builtin_amdgcn_s_memrealtime() : 0;
    treeArrive(tree, b);
    treeWait(tree, b);
    if (measure) waited = __builtin_amdgcn_s_memrealtime() - t0;
  }
  __syncthreads();
  return waited;
}

}  // namespace tgs
