// Batched cycle kernel for gfx950: many stimulus instances of one design in
// one persistent launch.
//
// Semantics per instance are the reference evaluator's (src/SimEval.cu:304-761,
// :765-864, :889-965). The execution is laid out for 64-wide waves:
//   * a wave serves a group of I instances (I = 8, 16, 32 or 64) and W = 64 / I
//     ops at once: lane l evaluates op sub = l / I of the current round for
//     instance ii = l % I. A round of W ops costs what one op costs when
//     every lane is an instance, so small I cuts the latency of one
//     (partition, group) item by W and its LDS by W, which lets W times more
//     waves share a CU. The LUT tasks of nibble programs (T_LUTFAST) go
//     further: a lane holds four instances in the bytes of 32-bit LDS words,
//     so a round is 256 / I ops and an op's program words, operand reads and
//     write are shared by four instances;
//   * every pool is instance-minor (nibble k of slot s at k * instPad + s),
//     so a group's nibble k is I contiguous bytes, moved 16 bytes per lane by
//     the L0 reads and the last-level word copies;
//   * a work item is (partition, up to NW groups): the workgroup stages the
//     partition's program words in LDS once and each of its NW waves runs the
//     whole partition for its own group, with its own value pool [slot][I]
//     and micropartition level buffers [2][half][32][I] in LDS, so waves
//     never wait on each other inside a partition;
//   * an op index past the end of a level is clamped to the level's last op:
//     the lane recomputes that op and rewrites the same value, so no lane
//     needs a branch;
//   * a reference lane index past the previous level's size (which reads 0,
//     SimEval.cu:372-375,406-408) is stored as the immediate 0 by the image
//     builder.
// Instances run uneven cycle counts: the host orders instance slots by
// decreasing cycle count, so the live groups of a cycle are a prefix and the
// live lanes of a group are a prefix; lanes past their own count never write.
// Stimulus draws for cycle c+1 are produced during region 1 of cycle c into
// the other half of a double buffer, so no extra grid barrier is needed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "kernel_args.h"

namespace tgs {

namespace {

constexpr uint32_t kLutBytesB = ((kLutSize + 15) / 16) * 16;

// The batched kernel's dynamic LDS: the LUT region at offset 0, then the
// waves' rows and the program buffers (a file-scope name, so LUT positions
// become immediate offsets of the lookups).
extern __shared__ __align__(16) uint8_t bLds[];

// One wave's view of its group of I instances. Rows hold a byte per
// instance, or (NIB, nibble programs) a nibble per instance: instances 2j
// and 2j + 1 in the low and high nibble of byte j. Per-instance tasks give a
// lane one byte of a row: one instance, or with NIB the pair of instances
// that share it, so a lane's writes never share a byte with another lane's.
template <int I, bool NIB>
struct Grp {
  static constexpr uint32_t RB = NIB ? I / 2 : I;  // bytes per row
  static constexpr uint32_t H = NIB ? 2 : 1;       // instances per lane in per-instance tasks
  static constexpr uint32_t W = 64 / RB;           // ops per round of per-instance tasks
  uint8_t *vp;         // value pool rows [slot][RB], then the level rows [kLevelRows][RB]
  uint8_t *lvl;        // row 0 of the level rows
  const uint8_t *lut;  // the LUT (4,642 entries + zero padding)
  const uint8_t *mem;
  uint64_t memSize;
  size_t instPad;
  uint32_t ii, sub;    // byte within the row, op within the round
  uint32_t slot0;      // first instance slot of the group
  // instance slot of the lane's instance h
  __device__ __forceinline__ uint32_t inst(uint32_t h) const { return slot0 + (NIB ? 2 * ii + h : ii); }
  __device__ __forceinline__ uint32_t raw(uint32_t s) const { return vp[s * RB + ii]; }
  __device__ __forceinline__ void putRaw(uint32_t s, uint32_t b) const { vp[s * RB + ii] = uint8_t(b); }
  // value of row s for the lane's instance h
  __device__ __forceinline__ uint32_t at(uint32_t s, uint32_t h = 0) const {
    const uint32_t b = raw(s);
    return NIB ? (b >> (4 * h)) & 15u : b;
  }
  // Operand: slots below 16 are immediates. The read is unconditional (every
  // slot a program names is inside the value pool, which has at least 16
  // slots), so operand reads need no branch and are issued together.
  __device__ __forceinline__ uint32_t opnd(uint32_t s, uint32_t h = 0) const {
    const uint32_t v = at(s, h);
    return s < 16 ? s : v;
  }
  // the lane's byte of row s from the values of its instances (v[h], nibbles with NIB)
  __device__ __forceinline__ void put(uint32_t s, const uint32_t *v) const {
    putRaw(s, NIB ? (v[0] | v[1] << 4) : v[0]);
  }
  // op index of this lane in round `r` of a level of n >= 1 ops
  __device__ __forceinline__ uint32_t opOf(uint32_t r, uint32_t n) const { return min(r + sub, n - 1); }
};

// A wave-uniform value read from LDS, kept in a scalar register.
__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ uint32_t lutAt(const uint8_t *lut, uint32_t lt, uint32_t a, uint32_t b, uint32_t c) {
  const uint32_t pos = lutIndex(lt, a, b, c);
  const uint32_t v = lut[pos < kLutSize ? pos : kLutSize];  // the padding byte at kLutSize is 0
  return lt == 0 ? c : v;
}

// A T_LUTROWS task: two on-disk micropartitions with the same number of
// middle levels (SimEval.cu:319-448), every operand and destination already
// resolved by the image builder to a row of the wave's LDS area (value-pool
// rows, level buffers, constant rows), so an op costs its row reads, one LUT
// lookup and one write. Two rounds of W ops are in flight per iteration, and
// the program words of each level's first iteration are read while the
// level before it is evaluated, so a level's dependent chain is its operand
// reads and its LUT lookups.
template <int I>
__device__ __forceinline__ void lutTask(const Grp<I, false> &g, const uint32_t *T) {
  constexpr uint32_t W = Grp<I, false>::W;
  // counts are wave-uniform: scalar registers, scalar loop control
  const uint32_t w0 = uni(T[0]), w1 = uni(T[1]);
  const uint32_t nMid = (w0 >> 8) & 0xFFFFu;
  const uint32_t nt = w1 & 0xFFFFu, nl = w1 >> 16;
  const uint32_t *S = T + 4;
  const uint2 *top = reinterpret_cast<const uint2 *>(S + (((nMid + 3) >> 2) + 1 & ~1u));
  const uint2 *ops = reinterpret_cast<const uint2 *>(T + uni(T[2]));
  const uint32_t *last = T + uni(T[3]);
  const uint8_t *lut = g.lut;
  uint8_t *vb = g.vp + g.ii;   // value-pool row 0 of this lane's instance
  uint8_t *lb = g.lvl + g.ii;  // level row 0
  auto sizeOf = [&](uint32_t m) -> uint32_t { return uni((S[m >> 2] >> ((m & 3u) * 8)) & 0xFFu); };
  // top level (SimEval.cu:338-359): value-pool or constant rows in, a level
  // row out. Both ops of an iteration read before either writes (ops of one
  // level are independent), so their LDS round trips overlap.
  auto topPair = [&](uint2 a, uint2 b) {
    const uint32_t va = lutAt(lut, a.x & 0xFFFFu, vb[(a.x >> 16) * I], vb[(a.y & 0x1FFFu) * I],
                              vb[((a.y >> 13) & 0x1FFFu) * I]);
    const uint32_t vbb = lutAt(lut, b.x & 0xFFFFu, vb[(b.x >> 16) * I], vb[(b.y & 0x1FFFu) * I],
                               vb[((b.y >> 13) & 0x1FFFu) * I]);
    lb[(a.y >> 26) * I] = uint8_t(va);
    lb[(b.y >> 26) * I] = uint8_t(vbb);
  };
  // middle level (SimEval.cu:362-421): level or constant rows in, an absolute row out
  auto midPair = [&](uint2 a, uint2 b) {
    const uint32_t va = lutAt(lut, a.x & 0x3FFFu, lb[((a.x >> 14) & 0xFFu) * I], lb[(a.x >> 22) * I],
                              lb[(a.y & 0xFFu) * I]);
    const uint32_t vbb = lutAt(lut, b.x & 0x3FFFu, lb[((b.x >> 14) & 0xFFu) * I], lb[(b.x >> 22) * I],
                               lb[(b.y & 0xFFu) * I]);
    vb[(a.y >> 8) * I] = uint8_t(va);
    vb[(b.y >> 8) * I] = uint8_t(vbb);
  };
  const uint2 zero = make_uint2(0, 0);
  uint2 c0 = zero, c1 = zero;  // the current level's first-iteration words
  if (nt) {
    c0 = top[g.opOf(0, nt)];
    c1 = top[g.opOf(W, nt)];
  }
  uint32_t n = nMid ? sizeOf(0) : 0;
  uint2 d0 = zero, d1 = zero;
  if (n) {
    d0 = ops[g.opOf(0, n)];
    d1 = ops[g.opOf(W, n)];
  }
  if (nt) {
    topPair(c0, c1);
    for (uint32_t r = 2 * W; r < nt; r += 2 * W) topPair(top[g.opOf(r, nt)], top[g.opOf(r + W, nt)]);
  }
  for (uint32_t m = 0; m < nMid; ++m) {
    c0 = d0;
    c1 = d1;
    const uint2 *cur = ops;
    ops += n;
    const uint32_t cn = n;
    n = m + 1 < nMid ? sizeOf(m + 1) : 0;
    if (n) {  // next level's first words, read before this level's chain
      d0 = ops[g.opOf(0, n)];
      d1 = ops[g.opOf(W, n)];
    }
    if (cn) {
      midPair(c0, c1);
      for (uint32_t r = 2 * W; r < cn; r += 2 * W) midPair(cur[g.opOf(r, cn)], cur[g.opOf(r + W, cn)]);
    }
  }
  // last level (SimEval.cu:425-447): the copies not folded into the last
  // middle level (fan-outs, immediates, a last level with no middle levels)
  for (uint32_t r = 0; r < nl; r += W) {
    const uint32_t w = last[g.opOf(r, nl)];
    vb[(w & 0xFFFFu) * I] = lb[((w >> 16) & 0xFFu) * I];
  }
}

// The four LUT indices of a T_LUTFAST op for four instances whose operand
// nibbles sit in the bytes of A, B and C (byte k = instance k): index k =
// a_k << 8 | b_k << 4 | c_k. B << 4 | C puts b_k << 4 | c_k in byte k (the
// operands are nibbles, so no byte carries into the next); two byte
// permutes then interleave those bytes with A's, which gives index k in a
// 16-bit half: i01 = {index 0, index 1}, i23 = {index 2, index 3}.
struct QuadIdx {
  uint32_t i01, i23;
};
__device__ __forceinline__ QuadIdx quadIndex(uint32_t A, uint32_t B, uint32_t C) {
  uint32_t bc;
  asm("v_lshl_or_b32 %0, %1, 4, %2" : "=v"(bc) : "v"(B), "v"(C));
  // v_perm_b32 byte selectors: 0-3 pick bytes of the second source (bc),
  // 4-7 bytes of the first (A)
  return {__builtin_amdgcn_perm(A, bc, 0x05010400u), __builtin_amdgcn_perm(A, bc, 0x07030602u)};
}

// LUT position = one 16-bit half of an index pair plus the low half of the
// op's first word (its LUT position of index 0), in one SDWA add each.
__device__ __forceinline__ uint32_t posLo(uint32_t pair, uint32_t w) {
  uint32_t pos;
  asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:WORD_0 src1_sel:WORD_0"
      : "=v"(pos)
      : "v"(pair), "v"(w));
  return pos;
}
__device__ __forceinline__ uint32_t posHi(uint32_t pair, uint32_t w) {
  uint32_t pos;
  asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:WORD_1 src1_sel:WORD_0"
      : "=v"(pos)
      : "v"(pair), "v"(w));
  return pos;
}

__device__ __forceinline__ uint32_t ldsByte(uint32_t pos) {
  return *reinterpret_cast<const __attribute__((address_space(3))) uint8_t *>(pos);
}

// The eight LUT indices of a T_LUTFAST op for eight instances whose operand
// nibbles sit in the nibbles of A, B and C (nibble k = instance k: even
// instances in the low nibbles of the bytes, odd ones in the high nibbles).
// Masked shifts spread them into bytes (even: a_k, b_k << 4 | c_k in byte
// k / 2; odd the same from the high nibbles) and quadIndex's byte permutes
// give the indices in 16-bit halves: e01 = {0, 2}, e23 = {4, 6}, o01 =
// {1, 3}, o23 = {5, 7}.
struct OctIdx {
  QuadIdx e, o;
};
__device__ __forceinline__ OctIdx octIndex(uint32_t A, uint32_t B, uint32_t C) {
  const uint32_t bce = ((B << 4) & 0xF0F0F0F0u) | (C & 0x0F0F0F0Fu);
  const uint32_t bco = (B & 0xF0F0F0F0u) | ((C >> 4) & 0x0F0F0F0Fu);
  const uint32_t ae = A & 0x0F0F0F0Fu, ao = (A >> 4) & 0x0F0F0F0Fu;
  return {{__builtin_amdgcn_perm(ae, bce, 0x05010400u), __builtin_amdgcn_perm(ae, bce, 0x07030602u)},
          {__builtin_amdgcn_perm(ao, bco, 0x05010400u), __builtin_amdgcn_perm(ao, bco, 0x07030602u)}};
}

// A T_LUTFAST task: the same two micropartitions as T_LUTROWS, with every
// operand and destination a byte offset from the lane's column of the wave's
// LDS area and every LUT index a position in the LDS LUT region that needs
// no bounds check (the image builder emits it only where every value is a
// nibble). A lane holds one 32-bit word of the group's rows: four
// consecutive instances (a byte each), or with NIB eight (a nibble each).
// Lane l evaluates op sub = l / P of a round for word q = l % P of the rows,
// P = RB / 4 (RB bytes per row), so a round is W = 64 / P ops (256 / I
// bytes, 512 / I nibbles). An op is its program words, three 32-bit operand
// reads, the indices computed from those words, a lookup per instance and
// one 32-bit write; top and middle ops share that form. A lane past the end of a level takes the
// level's last op: it rewrites that op's destination with the same value on
// the last task level, and elsewhere writes level row r + sub, a slack row
// past the level's ops that no op reads (r + sub < 64: a level has at most
// 64 ops and W divides 64). Level sizes, prefetch and the two rounds per
// iteration are as in lutTask.
template <int I, bool NIB>
__device__ __forceinline__ void lutFastTask(const Grp<I, NIB> &g, const uint32_t *T) {
  constexpr uint32_t RB = Grp<I, NIB>::RB;
  static_assert(RB % 4 == 0, "whole 32-bit words per row");
  constexpr uint32_t P = RB / 4, W = 64 / P;
  const uint32_t lane = threadIdx.x & 63u, q = lane % P, sub = lane / P;
  const uint32_t w0 = uni(T[0]), w1 = uni(T[1]);
  const uint32_t nMid = (w0 >> 8) & 0xFFFFu;
  const uint32_t nt = w1 & 0xFFFFu, nl = w1 >> 16;
  const uint32_t *S = T + 4;
  const uint2 *cur = reinterpret_cast<const uint2 *>(S + (((nMid + 3) >> 2) + 1 & ~1u));
  const uint2 *mids = reinterpret_cast<const uint2 *>(T + uni(T[2]));
  const uint32_t *last = T + uni(T[3]);
  uint8_t *vb = g.vp + 4 * q;                        // offsets are relative to this lane's column
  uint8_t *lb = g.lvl + 4 * q + sub * RB;  // level row `sub` of buffer X
  auto sizeOf = [&](uint32_t m) -> uint32_t { return uni((S[m >> 2] >> ((m & 3u) * 8)) & 0xFFu); };
  // this lane's op in round r of a level of n >= 1 ops
  auto opOf = [&](uint32_t r, uint32_t n) -> uint32_t { return min(r + sub, n - 1); };
  // offsets are multiples of RB (a multiple of 4), so every access is an
  // aligned 32-bit word
  auto rd = [&](uint32_t off) -> uint32_t { return *reinterpret_cast<const uint32_t *>(vb + off); };
  auto wr = [&](uint8_t *p, uint32_t v) { *reinterpret_cast<uint32_t *>(p) = v; };
  // the LUT region starts at LDS address 0 (the kernel has no static LDS,
  // which the runtime checks before the first launch)
  auto eval = [&](uint2 w) -> uint32_t {
    if constexpr (NIB) {
      const OctIdx x = octIndex(rd(w.x >> 16), rd(w.y & 0xFFFFu), rd(w.y >> 16));
      const uint32_t r0 = ldsByte(posLo(x.e.i01, w.x)), r2 = ldsByte(posHi(x.e.i01, w.x));
      const uint32_t r4 = ldsByte(posLo(x.e.i23, w.x)), r6 = ldsByte(posHi(x.e.i23, w.x));
      const uint32_t r1 = ldsByte(posLo(x.o.i01, w.x)), r3 = ldsByte(posHi(x.o.i01, w.x));
      const uint32_t r5 = ldsByte(posLo(x.o.i23, w.x)), r7 = ldsByte(posHi(x.o.i23, w.x));
      const uint32_t lo = (r0 | r1 << 4) | (r2 | r3 << 4) << 8, hi = (r4 | r5 << 4) | (r6 | r7 << 4) << 8;
      return lo | hi << 16;
    } else {
      const QuadIdx x = quadIndex(rd(w.x >> 16), rd(w.y & 0xFFFFu), rd(w.y >> 16));
      const uint32_t r0 = ldsByte(posLo(x.i01, w.x)), r1 = ldsByte(posHi(x.i01, w.x));
      const uint32_t r2 = ldsByte(posLo(x.i23, w.x)), r3 = ldsByte(posHi(x.i23, w.x));
      return r0 | r1 << 8 | r2 << 16 | r3 << 24;
    }
  };
  const uint2 zero = make_uint2(0, 0);
  uint2 c0 = zero, c1 = zero;
  uint32_t n = nt;
  if (n) c0 = cur[opOf(0, n)];
  if (n > W) c1 = cur[opOf(W, n)];
  // level l: 0 = the first task level (writes X), then Y, X, ... (the last
  // one writes through its destination list)
  for (uint32_t l = 0; l <= nMid; ++l) {
    const uint32_t nn = l < nMid ? sizeOf(l) : 0;
    const uint2 *next = l == 0 ? mids : cur + n;
    uint2 d0 = zero, d1 = zero;
    if (nn) d0 = next[opOf(0, nn)];  // next level's first words, read before this level's chain
    if (nn > W) d1 = next[opOf(W, nn)];
    // rounds of W ops: two per iteration while a second round remains, then
    // at most one more
    if (n) {
      if (l > 0 && l == nMid) {
        const uint16_t *dst = reinterpret_cast<const uint16_t *>(cur + n);
        uint32_t r = 0;
        for (; r + W < n; r += 2 * W) {
          const uint2 a = r ? cur[opOf(r, n)] : c0, b = r ? cur[opOf(r + W, n)] : c1;
          const uint32_t o0 = dst[opOf(r, n)], o1 = dst[opOf(r + W, n)];
          const uint32_t va = eval(a), vbb = eval(b);
          wr(vb + o0, va);
          wr(vb + o1, vbb);
        }
        if (r < n) {
          const uint32_t o0 = dst[opOf(r, n)];
          wr(vb + o0, eval(r ? cur[opOf(r, n)] : c0));
        }
      } else {
        uint8_t *wl = lb + ((l == 0 || ((l - 1) & 1u)) ? 0u : kFastBufY * RB);
        uint32_t r = 0;
        for (; r + W < n; r += 2 * W) {
          const uint2 a = r ? cur[opOf(r, n)] : c0, b = r ? cur[opOf(r + W, n)] : c1;
          const uint32_t va = eval(a), vbb = eval(b);
          wr(wl + r * RB, va);
          wr(wl + (r + W) * RB, vbb);
        }
        if (r < n) wr(wl + r * RB, eval(r ? cur[opOf(r, n)] : c0));
      }
    }
    cur = next;
    n = nn;
    c0 = d0;
    c1 = d1;
  }
  // last level: the copies not folded into the last middle level
  for (uint32_t r = 0; r < nl; r += W) {
    const uint32_t w = last[opOf(r, nl)];
    wr(vb + (w & 0xFFFFu), rd(w >> 16));
  }
}

// VecRead (SimEval.cu:449-481), W ops per round, each lane for its H
// instances. The loader rejects programs in which one op of an exec level
// reads a slot another op of that level writes, so the order of ops is
// unobservable.
template <int I, bool NIB>
__device__ __forceinline__ void vecReadB(const Grp<I, NIB> &g, const uint32_t *T, const uint8_t *consts) {
  using G = Grp<I, NIB>;
  const uint32_t n = T[0] >> 8;
  for (uint32_t r = 0; r < n; r += G::W) {
    const uint32_t *o = T + 1 + 5 * g.opOf(r, n);
    const uint32_t vb = o[0] & 0xFFFFu, len = o[0] >> 16, offset = o[1] & 0xFFFFu, isConst = o[1] >> 16;
    uint32_t v[G::H];
#pragma unroll
    for (uint32_t h = 0; h < G::H; ++h) {
      const uint32_t off = ((g.opnd(o[2] & 0xFFFFu, h) << 12) | (g.opnd(o[2] >> 16, h) << 8) |
                            (g.opnd(o[3] & 0xFFFFu, h) << 4) | g.opnd(o[3] >> 16, h)) + offset;
      v[h] = off < len ? (isConst ? consts[vb + off] : g.at(vb + off, h)) : g.opnd(o[4] & 0xFFFFu, h);
    }
    g.put(o[4] >> 16, v);
  }
}

// VecArithOrLogic (SimEval.cu:482-548), W ops per round, each lane for its
// H instances.
template <int I, bool NIB>
__device__ __forceinline__ void vecOpB(const Grp<I, NIB> &g, const uint32_t *T, const uint8_t *consts) {
  using G = Grp<I, NIB>;
  const uint32_t n = T[0] >> 8;
  for (uint32_t r = 0; r < n; r += G::W) {
    const uint32_t *o = T + 1 + 3 * g.opOf(r, n);
    const uint32_t b1 = o[0] & 0xFFFFu, b2 = o[0] >> 16;
    const uint32_t len = o[1] & 0xFFFFu, name = (o[1] >> 16) & 0xFFu, cf = o[1] >> 24, res = o[2];
    unsigned __int128 rr[G::H];
    uint32_t cmp[G::H];
#pragma unroll
    for (uint32_t h = 0; h < G::H; ++h) {
      unsigned __int128 v1 = 0, v2 = 0;
      for (uint32_t j = 0; j < len; ++j) {
        unsigned __int128 s1 = (cf & 2) ? consts[b1 + j] : g.at(b1 + j, h);
        unsigned __int128 s2 = (cf & 1) ? consts[b2 + j] : g.at(b2 + j, h);
        v1 |= s1 << (4 * j);
        v2 |= s2 << (4 * j);
      }
      if (len > 0 && len < 32) {  // sign_extend_128 (SimEval.cu:222-228)
        const uint32_t bit = 4 * len - 1;
        const unsigned __int128 ext = ~((((unsigned __int128)1) << (bit + 1)) - 1);
        if ((v1 >> bit) & 1) v1 |= ext;
        if ((v2 >> bit) & 1) v2 |= ext;
      }
      rr[h] = name == 0 ? v1 + v2 : name == 1 ? v1 - v2 : v1 * v2;
      const __int128 s1 = (__int128)v1, s2 = (__int128)v2;
      cmp[h] = name == 3 ? (s1 == s2) : name == 4 ? (s1 < s2) : name == 5 ? (s1 <= s2) : 0;
    }
    if (name <= 2) {
      for (uint32_t j = 0; j < len; ++j) {
        uint32_t v[G::H];
#pragma unroll
        for (uint32_t h = 0; h < G::H; ++h) {
          v[h] = uint32_t(rr[h] & 15u);
          rr[h] >>= 4;
        }
        g.put(uint16_t(res + j), v);
      }
    } else {
      g.put(res, cmp);
    }
  }
}

// MemRead (SimEval.cu:550-575), W ops per round, each lane for its H
// instances; an instance whose enable is 0 keeps its value.
template <int I, bool NIB>
__device__ __forceinline__ void memReadB(const Grp<I, NIB> &g, const uint32_t *T) {
  using G = Grp<I, NIB>;
  const uint32_t n = T[0] >> 8;
  for (uint32_t r = 0; r < n; r += G::W) {
    const uint32_t *o = T + 1 + 4 * g.opOf(r, n);
    const uint32_t res = o[3] & 0xFFFFu;
    uint32_t v[G::H];
    bool any = false;
#pragma unroll
    for (uint32_t h = 0; h < G::H; ++h) {
      v[h] = g.at(res, h);
      if (g.opnd(o[2] & 0xFFFFu, h) == 0) continue;
      uint32_t addr = 0;
#pragma unroll
      for (uint32_t j = 0; j < 8; ++j) addr |= g.at((o[2] >> 16) + j, h) << (4 * j);
      if (o[3] >> 16) addr <<= 2;
      const uint64_t at = (uint64_t(o[0]) | uint64_t(o[1]) << 32) + addr;
      v[h] = at < g.memSize ? g.mem[at * g.instPad + g.inst(h)] : 0u;
      any = true;
    }
    if (any) g.put(res, v);
  }
}

// A lane's share of one row of a group in the instance-minor pools: 16
// bytes (I / 16 lanes per row) for I >= 16, the whole 8-byte row for I = 8;
// half as many bytes in the LDS rows with NIB.
template <int I>
struct RowPart {
  static constexpr uint32_t B = I >= 16 ? 16u : uint32_t(I);  // pool bytes per lane
  static constexpr uint32_t LPO = I / B, OPR = 64 / LPO;       // lanes per row, rows per instruction
  using T = typename std::conditional<(I >= 16), uint4, uint2>::type;
  using N = typename std::conditional<(I >= 16), uint2, uint32_t>::type;  // the same instances as nibbles
};

// Eight pool bytes (nibbles) d0 | d1 << 32 as four bytes of nibble pairs:
// byte j = byte 2j | byte 2j + 1 << 4. d | d >> 4 puts each pair in bytes 0
// and 2 of its word (the bytes are nibbles, so no bits cross), one byte
// permute gathers them.
__device__ __forceinline__ uint32_t packNib(uint32_t d0, uint32_t d1) {
  return __builtin_amdgcn_perm(d1 | (d1 >> 4), d0 | (d0 >> 4), 0x06040200u);
}
// The inverse: four bytes of nibble pairs as eight bytes.
__device__ __forceinline__ uint2 unpackNib(uint32_t p) {
  const uint32_t lo = __builtin_amdgcn_perm(p, p, 0x01010000u), hi = __builtin_amdgcn_perm(p, p, 0x03030202u);
  auto split = [](uint32_t t) { return (t & 0x000F000Fu) | ((t >> 4) & 0x0F000F00u); };
  return make_uint2(split(lo), split(hi));
}

template <int I>
__device__ __forceinline__ typename RowPart<I>::N packRow(const typename RowPart<I>::T &v) {
  if constexpr (I >= 16) return make_uint2(packNib(v.x, v.y), packNib(v.z, v.w));
  else return packNib(v.x, v.y);
}
template <int I>
__device__ __forceinline__ typename RowPart<I>::T unpackRow(const typename RowPart<I>::N &v) {
  if constexpr (I >= 16) {
    const uint2 a = unpackNib(v.x), b = unpackNib(v.y);
    return make_uint4(a.x, a.y, b.x, b.y);
  } else {
    return unpackNib(v);
  }
}

// L0 reads of one group: op k copies pool[src_k][slot0 .. slot0 + I) into
// row slot_k of vp, RowPart<I>::B pool bytes per lane (nibble-packed with
// NIB: the bytes are nibbles there).
template <int I, bool NIB>
__device__ __forceinline__ void wideRead(uint8_t *__restrict__ vp, const uint8_t *__restrict__ pool, size_t instPad,
                                         uint32_t slot0, const uint32_t *__restrict__ L, uint32_t n, uint32_t lane) {
  using R = RowPart<I>;
  using V = typename R::T;
  constexpr uint32_t opr = R::OPR, RB = NIB ? I / 2 : I;
  const uint32_t q = lane / R::LPO, j = (lane % R::LPO) * R::B, jr = NIB ? j / 2 : j;
  const uint8_t *base = pool + slot0 + j;
  auto put = [&](uint32_t d, const V &v) {
    if constexpr (NIB) *reinterpret_cast<typename R::N *>(vp + d * RB + jr) = packRow<I>(v);
    else *reinterpret_cast<V *>(vp + d * RB + jr) = v;
  };
  uint32_t k = q;
  for (; k + 3 * opr < n; k += 4 * opr) {
    const uint32_t s0 = L[2 * k], d0 = L[2 * k + 1], s1 = L[2 * (k + opr)], d1 = L[2 * (k + opr) + 1];
    const uint32_t s2 = L[2 * (k + 2 * opr)], d2 = L[2 * (k + 2 * opr) + 1];
    const uint32_t s3 = L[2 * (k + 3 * opr)], d3 = L[2 * (k + 3 * opr) + 1];
    const V v0 = *reinterpret_cast<const V *>(base + size_t(s0) * instPad);
    const V v1 = *reinterpret_cast<const V *>(base + size_t(s1) * instPad);
    const V v2 = *reinterpret_cast<const V *>(base + size_t(s2) * instPad);
    const V v3 = *reinterpret_cast<const V *>(base + size_t(s3) * instPad);
    put(d0, v0);
    put(d1, v1);
    put(d2, v2);
    put(d3, v3);
  }
  for (; k < n; k += opr) put(L[2 * k + 1], *reinterpret_cast<const V *>(base + size_t(L[2 * k]) * instPad));
}

// Word copy of the last level (SimEval.cu:658-665,751-758) for one group:
// rows [0, n) of vp from `dat` to `pool` from `row0`, live instances only
// (the live instances of a group are a prefix of length `live`).
template <int I, bool NIB>
__device__ __forceinline__ void wideWrite(uint8_t *pool, size_t instPad, uint32_t slot0, uint32_t row0,
                                          const uint8_t *vp, uint32_t dat, uint32_t n, uint32_t lane, uint32_t live) {
  using R = RowPart<I>;
  using V = typename R::T;
  constexpr uint32_t RB = NIB ? I / 2 : I;
  const uint32_t q = lane / R::LPO, j = (lane % R::LPO) * R::B, jr = NIB ? j / 2 : j;
  for (uint32_t k = q; k < n; k += R::OPR) {
    const uint8_t *src = vp + (dat + k) * RB + jr;
    uint8_t *dst = pool + size_t(row0 + k) * instPad + slot0 + j;
    V v;
    if constexpr (NIB) v = unpackRow<I>(*reinterpret_cast<const typename R::N *>(src));
    else v = *reinterpret_cast<const V *>(src);
    if (j + R::B <= live) {
      *reinterpret_cast<V *>(dst) = v;
    } else if (j < live) {
      const uint8_t *b = reinterpret_cast<const uint8_t *>(&v);
      for (uint32_t t = 0; t < live - j; ++t) dst[t] = b[t];
    }
  }
}

// One (partition, group) evaluated by one wave, no workgroup barriers. In
// per-instance work a lane serves its byte of the group's rows: instance
// slot0 + ii, or with NIB the pair slot0 + 2 ii + h, h = 0, 1.
template <int I, bool NIB>
__device__ __forceinline__ void waveItem(const BatchArgs &a, uint32_t part, const uint32_t *P, uint32_t grp, uint8_t *vp,
                                         uint8_t *lvl, const uint8_t *lut, const uint64_t *draws, uint32_t c) {
  using G = Grp<I, NIB>;
  constexpr uint32_t W = G::W, H = G::H, RB = G::RB;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t ii = lane % RB, sub = lane / RB;
  const uint32_t slot0 = grp * I;
  const size_t IP = a.instPad;
  G g{vp, lvl, lut, a.mem, a.memSize, IP, ii, sub, slot0};
  bool liveH[H];
  uint64_t cycleH[H];
  uint32_t live = 0;  // live instances of the group (a prefix): counted over the lanes of op sub-lane 0
#pragma unroll
  for (uint32_t h = 0; h < H; ++h) {
    liveH[h] = c < a.runCycles[g.inst(h)];
    cycleH[h] = a.cycleBase[g.inst(h)] + c;
    live += __popcll(__ballot(liveH[h] && sub == 0));
  }
  {
    const uint32_t nrr = P[H_NRR], ner = P[H_NER], nsr = P[H_NSR];
    const uint32_t *L = P + P[H_OFF_L0];
    wideRead<I, NIB>(vp, a.reg, IP, slot0, L, nrr, lane);
    L += 2 * nrr;
    wideRead<I, NIB>(vp, a.exg, IP, slot0, L, ner, lane);
    L += 2 * ner;
    for (uint32_t k = sub; k < nsr; k += W) {
      const uint32_t in = L[2 * k];
      uint32_t v[H];
#pragma unroll
      for (uint32_t h = 0; h < H; ++h) v[h] = uint32_t((draws[size_t(in >> 6) * IP + g.inst(h)] >> (in & 63)) & 1u);
      g.put(L[2 * k + 1], v);
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
      if (kind == T_LUTFAST) {
        lutFastTask<I, NIB>(g, T);
      } else if (kind == T_LUTROWS) {
        if constexpr (!NIB) lutTask<I>(g, T);  // nibble rows are used only when every LUT task is T_LUTFAST
      } else if (kind == T_VECREAD) {
        vecReadB<I, NIB>(g, T, consts);
      } else if (kind == T_VECOP) {
        vecOpB<I, NIB>(g, T, consts);
      } else {
        memReadB<I, NIB>(g, T);
      }
    }
  }
  // signal probes of the live instances (slot-major, [probe][instPad])
  if (a.probeOff != nullptr)
    for (uint32_t k = a.probeOff[part] + sub; k < a.probeOff[part + 1]; k += W)
#pragma unroll
      for (uint32_t h = 0; h < H; ++h)
        if (liveH[h]) a.probeVal[size_t(k) * IP + g.inst(h)] = uint8_t(g.at(a.probeSlot[k], h));
  // last level (SimEval.cu:625-761); instances past their cycle count write nothing
  const uint32_t ewc = P[H_EW_DC] >> 16;
  if (ewc) {
    wideWrite<I, NIB>(a.exg, IP, slot0, (P[H_EW_EXG] >> 2) * 4, vp, ((P[H_EW_DC] & 0xFFFFu) >> 2) * 4,
                      (ewc + 3) / 4 * 4, lane, live);
    return;
  }
  const uint32_t rwc = P[H_RW_DC] >> 16;
  if (rwc)
    wideWrite<I, NIB>(a.reg, IP, slot0, (P[H_RW_REG] >> 2) * 4, vp, ((P[H_RW_DC] & 0xFFFFu) >> 2) * 4,
                      (rwc + 3) / 4 * 4, lane, live);
  const uint32_t nmw = P[H_NMW], npr = P[H_NPR], nst = P[H_NST];
  const uint32_t *L = P + P[H_OFF_LAST];
  // memory writes one op at a time in program order (a later write to the
  // same address wins, as in the oracle)
  if (sub == 0)
#pragma unroll
    for (uint32_t h = 0; h < H; ++h) {
      if (!liveH[h]) continue;
      for (uint32_t k = 0; k < nmw; ++k) {
        const uint32_t *o = L + 4 * k;
        if (g.opnd(o[3] & 0xFFFFu, h) == 0) continue;
        const uint32_t dat = g.opnd(o[2] >> 16, h);
        uint32_t addr = 0;
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) addr |= g.at((o[2] & 0xFFFFu) + j, h) << (4 * j);
        if (o[3] >> 16) addr <<= 2;
        const uint64_t at = (uint64_t(o[0]) | uint64_t(o[1]) << 32) + addr;
        if (at < a.memSize) a.mem[at * IP + g.inst(h)] = uint8_t(dat);
      }
    }
  L += 4 * nmw;
  if (a.printEnable)
#pragma unroll
    for (uint32_t h = 0; h < H; ++h) {
      if (!liveH[h]) continue;
      for (uint32_t k = sub; k < npr; k += W)
        if (g.opnd(L[k] & 0xFFFFu, h) != 0) {
          const uint64_t q = atomicAdd(reinterpret_cast<unsigned long long *>(a.logCount), 1ull);
          if (q < a.logCap) {
            a.logCycle[q] = cycleH[h];
            a.logMsg[q] = uint64_t(L[k] >> 16) | (uint64_t(a.instOf[g.inst(h)]) << 32);
          }
        }
    }
  L += npr;
#pragma unroll
  for (uint32_t h = 0; h < H; ++h) {
    if (!liveH[h]) continue;
    for (uint32_t k = sub; k < nst; k += W)
      if (g.opnd(L[k] & 0xFFFFu, h) != 0) {
        atomicOr(reinterpret_cast<unsigned long long *>(&a.stop[g.inst(h)]), 1ull);
        atomicMin(reinterpret_cast<unsigned long long *>(&a.stopCycle[g.inst(h)]), (unsigned long long)cycleH[h]);
      }
  }
}

// Advance the stimulus of every live instance of the active groups and store
// the draws of launch cycle `c` in buffer `buf`.
__device__ __forceinline__ void stimulusFor(const BatchArgs &a, uint32_t buf, uint32_t c, uint32_t activeSlots) {
  if (a.nIn == 0) return;
  const uint32_t nDraws = (a.nIn + 63) / 64;
  uint64_t *draws = a.draws + size_t(buf) * nDraws * a.instPad;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < activeSlots; s += gridDim.x * blockDim.x) {
    if (c >= a.runCycles[s]) continue;
    uint64_t x = a.xs[s];
    for (uint32_t d = 0; d < nDraws; ++d) {
      x = xorshift64(x);
      draws[size_t(d) * a.instPad + s] = x;
    }
    a.xs[s] = x;
  }
}

// Start copying a partition's program words (16-byte aligned, padded to 4
// words) into LDS with global-to-LDS loads: each wave instruction moves 64
// consecutive 16-byte chunks (lane k's chunk lands at the wave's LDS base +
// 16 k), so nothing passes through registers and the copy completes while
// the workgroup works on something else; s_waitcnt vmcnt(0) and a
// workgroup barrier precede any read of it.
__device__ __forceinline__ void stageProgramAsync(uint32_t *dst, const uint32_t *src, uint32_t words) {
  const uint32_t n4 = (words + 3) >> 2;
  const uint32_t wave = threadIdx.x & ~63u, lane = threadIdx.x & 63u;
  for (uint32_t k0 = wave; k0 < n4; k0 += blockDim.x) {
    if (k0 + lane < n4)
      __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint4 *>(src) + k0 + lane,
                                       (__attribute__((address_space(3))) void *)(reinterpret_cast<uint4 *>(dst) + k0),
                                       16, 0, 0);
  }
}

// Instance groups of context `a` with a live instance in launch cycle c.
__device__ __forceinline__ uint32_t activeOf(const BatchArgs &a, uint32_t c) {
  return c < a.nCycles ? a.groupsActive[c] : 0u;
}

// A wave's rows of the LDS area for context `a` (whose image puts its level
// rows after its own value pool): all zero, then the constant rows (row c
// holds c for every instance; with NIB only rows 0..15 are read).
template <int I, bool NIB>
__device__ __forceinline__ void resetRows(uint8_t *vp, uint32_t rowBytes, const BatchArgs &a) {
  constexpr uint32_t RB = NIB ? I / 2 : I;
  uint32_t *wv = reinterpret_cast<uint32_t *>(vp);
  for (uint32_t k = threadIdx.x & 63u; k < rowBytes / 4; k += 64) wv[k] = 0;
  uint8_t *cr = vp + (a.vpBytes + kConstRow) * RB;
  for (uint32_t k = threadIdx.x & 63u; k < (kLevelRows - kConstRow) * RB; k += 64) {
    const uint32_t v = k / RB;
    cr[k] = uint8_t(NIB ? (v & 15u) * 0x11u : v);
  }
}

// Work items of region r over every context of the launch, context by
// context: (context, partition, NW consecutive groups), one group per wave.
// A workgroup that moves to another context's items loads that context's
// LUT and resets its waves' rows first.
template <int I, bool NIB>
__device__ __forceinline__ void batchRegion(const BatchLaunch &L, uint32_t r, uint32_t c, uint8_t *vp, uint32_t rowBytes,
                                            uint32_t *prog, uint32_t &curCtx) {
  const uint32_t NW = blockDim.x >> 6, wave = threadIdx.x >> 6;
  auto partsOf = [&](const BatchArgs &a) -> uint32_t { return r == 0 ? a.n0 : a.nTotal - a.n0; };
  auto itemsOf = [&](const BatchArgs &a) -> uint32_t { return partsOf(a) * ((activeOf(a, c) + NW - 1) / NW); };
  uint32_t total = 0;
  for (uint32_t j = 0; j < L.nCtx; ++j) total += itemsOf(L.ctx[j]);
  // item k -> (context j, partition p, group)
  auto locate = [&](uint32_t k, uint32_t &j, uint32_t &p, uint32_t &grp) {
    j = 0;
    for (uint32_t n; k >= (n = itemsOf(L.ctx[j])); ++j) k -= n;
    const BatchArgs &a = L.ctx[j];
    const uint32_t chunks = (activeOf(a, c) + NW - 1) / NW;
    p = (r == 0 ? 0u : a.n0) + k / chunks;
    grp = (k % chunks) * NW + wave;
  };
  // two program buffers: the next item's program streams into one while
  // the current item runs from the other
  uint32_t k = blockIdx.x;
  if (k < total) {
    uint32_t j, p, grp;
    locate(k, j, p, grp);
    const BatchArgs &a = L.ctx[j];
    stageProgramAsync(prog, a.words + a.partOff[p], a.partWords[p]);
  }
  for (uint32_t buf = 0; k < total; k += gridDim.x, buf ^= 1u) {
    uint32_t j, p, grp;
    locate(k, j, p, grp);
    const BatchArgs &a = L.ctx[j];
    uint32_t *cur = prog + buf * L.progWords;
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's share of the copy has landed
    __syncthreads();
    if (j != curCtx) {  // every wave is past its last item: switch the LUT and rows
      for (uint32_t q = threadIdx.x; q < kLutBytesB / 4; q += blockDim.x)
        reinterpret_cast<uint32_t *>(bLds)[q] = reinterpret_cast<const uint32_t *>(a.lut)[q];
      resetRows<I, NIB>(vp, rowBytes, a);
      curCtx = j;
      __syncthreads();
    }
    const uint32_t kn = k + gridDim.x;
    if (kn < total) {
      uint32_t jn, pn, gn;
      locate(kn, jn, pn, gn);
      const BatchArgs &an = L.ctx[jn];
      stageProgramAsync(prog + (buf ^ 1u) * L.progWords, an.words + an.partOff[pn], an.partWords[pn]);
    }
    if (grp < activeOf(a, c)) {
      const uint64_t *draws = a.draws + size_t(c & 1u) * ((a.nIn + 63) / 64) * a.instPad;
      waveItem<I, NIB>(a, p, cur, grp, vp, vp + a.vpBytes * (NIB ? I / 2 : I), bLds, draws, c);
    }
    __syncthreads();  // every wave is done with `cur` before it is overwritten
  }
}

}  // namespace

// One persistent launch over the launch's contexts: every launch cycle runs
// region 0 of every context, a grid barrier, region 1, a grid barrier.
// groupsActive[c] of a context counts its groups of I instances.
template <int I, bool NIB>
__global__ __launch_bounds__(1024) void batchCycles(BatchLaunch L) {
  const uint32_t NW = blockDim.x >> 6;
  const uint32_t rowBytes = (L.vpRows + kLevelRows) * (NIB ? I / 2 : I);
  uint8_t *vpAll = bLds + kLutRegionB;  // per wave: value-pool rows, then level rows
  uint8_t *vp = vpAll + (threadIdx.x >> 6) * rowBytes;
  uint32_t *prog = reinterpret_cast<uint32_t *>(vpAll + NW * rowBytes);
  // LUT region: the first context's table (zero-padded to kLutBytesB), zeros
  // to kLutIdBase, then the identity bytes 0..31
  for (uint32_t k = threadIdx.x; k < kLutRegionB / 4; k += blockDim.x) {
    uint32_t v = 0;
    if (k < kLutBytesB / 4) v = reinterpret_cast<const uint32_t *>(L.ctx[0].lut)[k];
    else if (k >= kLutIdBase / 4) {
      const uint32_t b = 4 * k - kLutIdBase;
      v = b | (b + 1) << 8 | (b + 2) << 16 | (b + 3) << 24;
    }
    reinterpret_cast<uint32_t *>(bLds)[k] = v;
  }
  // the wave's rows start at 0 (a slot read before any write reads a nibble)
  resetRows<I, NIB>(vp, rowBytes, L.ctx[0]);
  uint32_t curCtx = 0;
  uint32_t bar = 0;
  auto stimulus = [&](uint32_t c) {  // draws of launch cycle c, before anyone reads them
    for (uint32_t j = 0; j < L.nCtx; ++j) {
      const BatchArgs &a = L.ctx[j];
      stimulusFor(a, c & 1u, c, activeOf(a, c) * uint32_t(I));
    }
  };
  stimulus(0);
  gridBarrier(L.barrier, ++bar, false);
  for (uint32_t c = 0; c < L.nCycles; ++c) {
    uint32_t any = 0;
    for (uint32_t j = 0; j < L.nCtx; ++j) any |= activeOf(L.ctx[j], c);
    if (any == 0) break;
    // input registers of live instances (never read back through normal register reads)
    for (uint32_t j = 0; j < L.nCtx; ++j) {
      const BatchArgs &a = L.ctx[j];
      const uint32_t activeSlots = activeOf(a, c) * uint32_t(I);
      const uint64_t *draws = a.draws + size_t(c & 1u) * ((a.nIn + 63) / 64) * a.instPad;
      for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < activeSlots; s += gridDim.x * blockDim.x)
        if (c < a.runCycles[s])
          for (uint32_t i = 0; i < a.nIn; ++i)
            a.reg[size_t(a.inIds[i]) * a.instPad + s] = uint8_t((draws[size_t(i >> 6) * a.instPad + s] >> (i & 63)) & 1u);
    }
    batchRegion<I, NIB>(L, 0, c, vp, rowBytes, prog, curCtx);
    gridBarrier(L.barrier, ++bar, false);
    if (c + 1 < L.nCycles) stimulus(c + 1);
    batchRegion<I, NIB>(L, 1, c, vp, rowBytes, prog, curCtx);
    gridBarrier(L.barrier, ++bar, false);
  }
}

namespace {
const void *batchFn(uint32_t inst, bool nib) {
  if (nib)
    return inst == 8    ? reinterpret_cast<const void *>(batchCycles<8, true>)
           : inst == 16 ? reinterpret_cast<const void *>(batchCycles<16, true>)
           : inst == 32 ? reinterpret_cast<const void *>(batchCycles<32, true>)
                        : reinterpret_cast<const void *>(batchCycles<64, true>);
  return inst == 8    ? reinterpret_cast<const void *>(batchCycles<8, false>)
         : inst == 16 ? reinterpret_cast<const void *>(batchCycles<16, false>)
         : inst == 32 ? reinterpret_cast<const void *>(batchCycles<32, false>)
                      : reinterpret_cast<const void *>(batchCycles<64, false>);
}
}  // namespace

bool batchSupportsGroup(uint32_t inst) { return inst == 8 || inst == 16 || inst == 32 || inst == 64; }

// LUT region, per wave value-pool and level rows (a byte, or with nib a
// nibble, per instance), two program buffers.
uint32_t batchLdsBytes(uint32_t inst, uint32_t waves, uint32_t vpBytes, uint32_t progWords, bool nib) {
  return kLutRegionB + waves * (vpBytes + kLevelRows) * (nib ? inst / 2 : inst) + 2u * progWords * 4u;
}

hipError_t launchBatch(const BatchLaunch &launch, uint32_t inst, bool nib, uint32_t grid, uint32_t block, uint32_t lds,
                       hipStream_t stream) {
  BatchLaunch a = launch;
  void *params[] = {&a};
  return hipLaunchCooperativeKernel(batchFn(inst, nib), dim3(grid), dim3(block), params, lds, stream);
}

hipError_t batchMaxBlocksPerCU(int *blocks, uint32_t inst, bool nib, uint32_t block, uint32_t lds) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, batchFn(inst, nib), block, lds);
}

// The kernels address the LUT region at LDS address 0: no static LDS.
hipError_t batchCheckNoStaticLds() {
  for (uint32_t inst : {8u, 16u, 32u, 64u})
    for (bool nib : {false, true}) {
      hipFuncAttributes fa;
      hipError_t e = hipFuncGetAttributes(&fa, batchFn(inst, nib));
      if (e != hipSuccess) return e;
      if (fa.sharedSizeBytes != 0) return hipErrorInvalidKernelFile;
    }
  return hipSuccess;
}

hipError_t batchSetLds(uint32_t lds) {
  for (uint32_t inst : {8u, 16u, 32u, 64u})
    for (bool nib : {false, true}) {
      hipError_t e = hipFuncSetAttribute(batchFn(inst, nib), hipFuncAttributeMaxDynamicSharedMemorySize, lds);
      if (e != hipSuccess) return e;
    }
  return hipSuccess;
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
