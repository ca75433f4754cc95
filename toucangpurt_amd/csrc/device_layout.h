// Device image layouts shared by the host image builder (image.cpp) and the
// gfx950 kernels (kernels_*.hip). Plain C structs, 4-byte fields only.
#pragma once
#include <stdint.h>

namespace tgs {

constexpr uint32_t kLutSize = 4642;       // include/ToucanGPUGenDataTypes.h LUT, SimEval.cu:33
constexpr uint32_t kLutWords = (kLutSize + 3) / 4;
constexpr uint32_t kMaxMidLevels = 158;   // SimEval.cu:92 (+2 <= 160)

// ---------------------------------------------------------------------------
// Plain image: the reference's device layout (SimEval.cu:69-148,1143-1281).
// Per partition one byte stream of 16-byte aligned chunks:
//   [L0 regRead | L0 exgRead | stimulus reads][microparts ...][last level]
// A LUT micropart is {u32 magic, u32 nMid, u8 levelSizes[160], top, mid, last}
// and the other kinds are {u32 magic, u32 n, records}; records keep the
// on-disk F1 layouts. Stimulus reads are a chunk of {u32 input, u16 slot, pad}.
constexpr uint32_t kMagicLut = 0xfb709394u;      // SimEval.cu:36
constexpr uint32_t kMagicVecRead = 0xfb809917u;  // SimEval.cu:37
constexpr uint32_t kMagicMemRead = 0xccccccccu;  // SimEval.cu:38
constexpr uint32_t kMagicVecOp = 0xbeef1234u;    // SimEval.cu:39
constexpr uint32_t kPlainHeaderBytes = 8 + 160;

struct PlainPart {
  uint32_t net;          // byte offset of the partition stream in the image
  uint32_t nRegRead;     // chunk offsets below are relative to `net`
  uint32_t nExgRead;
  uint32_t nStimRead;
  uint32_t offRegRead, offExgRead, offStimRead;
  uint32_t nLevels;
  uint32_t levelBase;    // index into the level table (pairs {nMParts, mpartBase})
  uint32_t offLast;      // regWrite/exgWrite/memWrite/print/stop chunks
  uint32_t hasRegWrite, hasExgWrite, nMemWrite, nPrint, nStop;
  uint32_t offRegWrite, offExgWrite, offMemWrite, offPrint, offStop;
  uint32_t constBase;    // byte offset of the constant-vector pool in the const image
  uint32_t nConst;
};

// ---------------------------------------------------------------------------
// Wave64 image: one u32 word stream; all offsets are word offsets relative to
// the partition's first word. The per-partition header:
enum W64Hdr : uint32_t {
  H_VP = 0,         // value-pool size (slots)
  H_NRR, H_NER, H_NSR,  // L0 register / exchange / stimulus reads
  H_OFF_L0,         // 2 words per read: {source, slot}
  H_NLEV,           // exec levels
  H_OFF_LEV,        // per level {task begin, task end} indices into the task table
  H_OFF_TASK,       // task table: word offset of each task stream
  H_RW_REG, H_RW_DC,  // register write: reg, dat | count << 16 (count 0 = none)
  H_EW_EXG, H_EW_DC,  // exchange write
  H_NMW, H_NPR, H_NST,  // last-level memory writes, prints, stops
  H_OFF_LAST,       // memwrite: 4 words each, print: 1, stop: 1
  H_CONST_BASE,     // byte offset of the constant-vector pool in the const image
  H_NCONST,
  H_WORDS           // header size
};

// Task kinds (low byte of a task's first word).
enum W64Task : uint32_t { T_LUT2 = 0, T_VECREAD = 1, T_VECOP = 2, T_MEMREAD = 3, T_LUTROWS = 4, T_LUTFAST = 5 };
// T_LUTROWS (batched image only): the ops of both micropartitions of a
// T_LUT2 pair with every operand and destination resolved to a row of the
// wave's LDS area: value-pool rows [0, V) then level rows: buffer X
// [0, 64), buffer Y [64, 128) (halves A and B at +0 / +32), and constant
// rows [kConstRow, kConstRow + 32) holding 0..31 (level rows are relative
// to row V).
//   w0  kind | nMid << 8
//   w1  nTop (A + B) | nLast << 16
//   w2  word offset (from w0) of the middle ops, w3 of the last ops
//   ceil(nMid / 4) words of level sizes, 8 bits each (A + B ops)
//   top ops (even offset), 2 words each: lut | r0 << 16,
//       r1 | r2 << 13 | dst << 26 (r*: absolute rows < 8192, dst a level row)
//   middle ops per level, 2 words each: lut | r0 << 14 | r1 << 22
//       (level rows), r2 (a level row) | dst << 8 (an absolute row: a level
//       row + V, or, on the last middle level, the value-pool slot of the
//       one last-level op that copies the lane, which is then dropped)
//   last ops, 1 word each: value-pool slot | source level row << 16
constexpr uint32_t kLevelRows = 176, kConstRow = 144;
// T_LUTFAST's level buffers: X at level row 0, Y at kFastBufY, each with
// room for 64 ops (a kernel lane past a level's end writes a slack row below
// 64: rounds are 256 / I ops, which divides 64)
constexpr uint32_t kFastBufY = 72;
// T_LUTFAST (batched image only, built for one instance count I per wave):
// LUT micropartitions of one exec level packed into one task, with every
// operand and destination resolved to a byte offset row * I from the lane's
// column of the wave's LDS area (rows as in T_LUTROWS), and every LUT index
// resolved to a position in the LDS LUT region below that needs no bounds
// check. Parts with fewer middle levels are aligned at the end: with nMid
// the largest count, a part with m middle levels has its top level on task
// level nMid - m and its last middle level on task level nMid; at most 64
// ops on a task level (levels are not padded: the kernel clamps a lane's op
// index to the level's last op). Only for programs in which every value is a nibble
// (nibbleProgram), and only for parts whose ops satisfy: top-op LUT index 0
// or <= 65536 - 4096 (no 16-bit wrap), no middle-op LUT operand an
// immediate >= 16 (operands <= 15 keep the index <= 0xFFF).
//   w0  kind | nMid << 8
//   w1  ops on task level 0 | nLast << 16
//   w2  word offset (from w0) of task level 1's ops, w3 of the last ops
//   ceil(nMid / 4) words of the sizes of task levels 1..nMid, 8 bits each
//   per task level (level 0 at an even offset), 2 words per op:
//       lut' | r0 << 16, r1 | r2 << 16 (a part's top ops read value-pool or
//       constant rows, its middle ops the rows of its own ops one task level
//       up); op k of task level L writes level row k of buffer X (L = 0),
//       Y (L odd, row kFastBufY + k) or X (L even), except on task level
//       nMid >= 1, whose ops
//       are followed by one u16 destination offset each (padded to a word):
//       a level row or, when exactly one last-level op copies the lane, that
//       op's value-pool slot
//   last ops, 1 word each: destination | source << 16
// lut' = LUT position of index 0: the op's own LUT index when it is below
// 4642, 4642 (a zero window) when it is not, kLutIdBase for index 0 (pass
// op2 through: op0 and op1 read the constant row of 0, and the identity
// bytes return op2).
constexpr uint32_t kLutZeroEnd = kLutSize + 4096;  // zeros [kLutSize, kLutZeroEnd)
constexpr uint32_t kLutIdBase = (kLutZeroEnd + 15) / 16 * 16;  // 32 identity bytes 0..31
constexpr uint32_t kLutRegionB = kLutIdBase + 32;              // LDS bytes of the batched kernel's LUT region
// T_LUT2: two on-disk 32-lane micropartitions in one wave (lanes 0-31 = A,
//   lanes 32-63 = B). Words:
//   w0  kind | nMid << 8
//   w1  nTopA | nTopB << 8 | nLastA << 16 | nLastB << 24
//   w2  word offset (from w0) of the last-level ops
//   ceil(nMid/2) words of middle sizes, (nA | nB << 8) per 16 bits
//   top ops   (nTopA + nTopB) x 2 words: lut | op0 << 16, op1 | op2 << 16
//   middle    per level (nA + nB) words: the packed on-disk op
//   last ops  (nLastA + nLastB) words: result | shuffleId << 16
// T_VECREAD: w0 kind | n << 8, then n x 5 words:
//   vecBase | vecLength << 16, offset | isConst << 16, idx0 | idx1 << 16,
//   idx2 | idx3 << 16, outRange | result << 16
// T_VECOP: w0 kind | n << 8, then n x 3 words:
//   v1 | v2 << 16, len | op << 16 | constFlags << 24, result
// T_MEMREAD: w0 kind | n << 8, then n x 4 words:
//   memBase lo, memBase hi, en | addrVec << 16, result | multi << 16
// Last level: memwrite 4 words {base lo, base hi, addrVec | dat << 16,
//   en | multi << 16}; print 1 word {en | msg << 16}; stop 1 word {en}.

// Per-launch schedule of the persistent wave64 kernel: for each workgroup and
// region, the partitions it evaluates (indices into the partition table).
// sched[g * 2 + r] .. sched[g * 2 + r + 1] index `parts`.

// ---------------------------------------------------------------------------
// Stream image (streamed kernel). The program is a u32 word array of 64-word
// rows. Each workgroup has nC compute waves and three loader waves and walks
// a flat step stream for one cycle; every step ends with a workgroup
// barrier. The loader waves copy the rows of upcoming steps from HBM into an
// LDS ring so the compute waves never wait on a global load for program
// words.
// Word 0..63 of the image is a zero row (the rows of "no task" entries).
//
// Step record: 1 + 2 * nC words: type, then per compute wave an entry
// {x = word offset (a multiple of 64) of the task or partition block,
//  y = kind << 4 | split << 7 | n << 8 | rows << 24}.
// A workgroup runs its partitions of a region in rounds. In a round with
// partitions in several of its `slots`, the compute waves split: slot s
// uses waves [s nC/slots, (s+1) nC/slots) and its own value pool, and every
// entry of the round's steps carries kSplit. A round with one partition
// gives it all nC waves (kSplit clear, value pool of slot 0).
enum StreamStep : uint32_t {
  S_LEVEL = 0,   // exec level (or a slice of one): one task per compute wave
  S_PART = 1,    // partition begin: L0 reads by the compute waves
  S_PEND = 2,    // partition end: last level by the compute waves
  S_REGION = 3   // grid barrier (end of a region)
};
constexpr uint32_t kKindNone = 7;      // entry kind of a wave with no task in the step
constexpr uint32_t kSplit = 1u << 7;   // entry flag: the waves are split over the slots
constexpr uint32_t kLastDirect = 1u << 23;  // T_LUT2 entry flag (n = nMid uses bits 8-15): direct last level
constexpr uint32_t kRow = 64;
// Task rows, lane l of a row belongs to lane l of the wave:
//   T_LUT2   (n = nMid, rows 3 + nMid): row 0 top word {lut | op0 << 16},
//            row 1 top word {op1 | op2 << 16}, row 2 last {result | shuffleId
//            << 16} (0xFFFFFFFF = no op), row 3 + m middle level m (packed
//            on-disk op, 0 = no op); lanes 0-31 micropart A, 32-63 B.
//            With kLastDirect, row 2 is indexed by the lane that holds the
//            value instead: {slot | 0 << 16} writes the lane's own final value
//            to slot, {slot | (1 + c) << 16} writes the constant c (an
//            immediate shuffle id, or 0 for a lane past the last level's
//            size), 0xFFFFFFFF nothing. The builder uses it when no lane's
//            value goes to more than one slot; it saves the last level's
//            shuffle.
//   T_VECREAD (n ops, 5 rows), T_VECOP (3 rows), T_MEMREAD (4 rows): row k
//            lane i = word k of op i in the wave64 layout above.
// Partition block: one section of 1 + 2 * nJ rows per compute wave (row 0 =
// the header, word index below; then for L0 entry j of the wave's lane l,
// thread t = wave * 64 + l handling entry j * nC * 64 + t, its source at row
// 1 + 2j and its value-pool slot at row 2 + 2j; source = index | kind << 30,
// kind 0 register, 1 exchange, 2 stimulus input, 3 none), then the
// last-level lists. A partition-begin entry points at the wave's section, so
// for every entry the rows a wave needs are contiguous from x.
enum StreamHdr : uint32_t {
  S_VP = 0, S_NJ, S_RW_REG, S_RW_DC, S_EW_EXG, S_EW_DC, S_NMW, S_NPR, S_NST,
  S_OFF_LAST,       // word offset (from the first section) of memwrite/print/stop lists
  S_CONST_BASE,     // byte offset of the constant-vector pool in the const image
  S_PART_ID,        // partition index (signal probes)
  S_WORDS
};
constexpr uint32_t kL0None = 3u << 30;
constexpr uint32_t kStreamRows = 6;   // rows staged per wave and step (the loader copies rows 0..5 from x)

}  // namespace tgs
