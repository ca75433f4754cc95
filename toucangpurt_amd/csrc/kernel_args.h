// Kernel argument blocks and host launchers of the gfx950 cycle kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "device_common.h"

namespace tgs {

struct PlainArgs {
  CommonArgs c;
  const uint8_t *net;
  const PlainPart *parts;
  const uint32_t *levels;
  const uint32_t *mparts;
  uint32_t n0, nTotal;
};

struct W64Args {
  CommonArgs c;
  const uint32_t *words;     // wave64 image
  const uint32_t *partOff;   // word offset per partition
  const uint32_t *partWords; // words per partition
  const uint32_t *schedOff;  // groups * 2 + 1
  const uint32_t *schedPart;
  const uint32_t *schedLds;
  uint32_t vpBytes;          // value-pool bytes in LDS (16-byte multiple)
  uint32_t barrierBase;      // index of the last barrier the instance's barrier tree completed before the launch
};

struct StreamArgs {
  CommonArgs c;
  const uint32_t *words;      // stream image rows
  const uint32_t *steps;      // step records, 1 + 2 * NC words each
  const uint32_t *stepBase;   // per workgroup: first record
  const uint32_t *stepCount;  // per workgroup: records per cycle
  uint32_t vpBytes;           // value-pool bytes in LDS per slot (16-byte multiple)
  uint32_t slots;             // partitions side by side per workgroup (NC / slots waves each)
  uint32_t barrierBase;       // index of the last barrier the instance's barrier tree completed before the launch
};

// kernargCommon() reads the CommonArgs at offset 0 of a kernel's argument block.
static_assert(offsetof(W64Args, c) == 0 && offsetof(StreamArgs, c) == 0 && offsetof(PlainArgs, c) == 0,
              "every persistent kernel's argument struct begins with its CommonArgs");

// Batched kernel: pools are instance-minor (nibble k of slot s at k * instPad + s).
struct BatchArgs {
  const uint8_t *lut, *consts;
  uint8_t *reg, *mem, *exg;
  uint32_t regSize, memSize, exgSize;
  uint32_t instPad;               // slots, a multiple of 64
  const uint32_t *words, *partOff, *partWords;  // wave64 image
  uint32_t n0, nTotal;
  uint32_t vpBytes;               // value-pool slots per instance (16-multiple); level rows follow
  const uint32_t *runCycles;      // per slot: cycles to run in this launch
  const uint64_t *cycleBase;      // per slot: cycles completed before this launch
  const uint32_t *groupsActive;   // per launch cycle: instance groups with a live slot (a prefix)
  uint32_t nCycles;               // longest run in this launch
  uint64_t *xs;                   // per slot xorshift64 state
  uint64_t *draws;                // [2][nDraws][instPad]
  const uint32_t *inIds;
  uint32_t nIn;
  uint64_t *stop;                 // per slot
  uint64_t *stopCycle;            // per slot
  uint64_t *logCount;             // entries appended since the host last drained the log
  uint64_t *logCycle;
  uint64_t *logMsg;               // msg | instance << 32
  const uint32_t *instOf;         // slot -> instance
  uint32_t logCap;
  uint32_t printEnable;
  const uint32_t *probeOff, *probeSlot;  // signal probes (null when none), CSR by partition
  uint8_t *probeVal;                     // [probe][instPad]
};

// One persistent batched launch over one or more contexts of one device
// (same instances per wave): their BatchArgs in device memory, the grid
// barrier, and the LDS layout sized for the largest of them.
struct BatchLaunch {
  const BatchArgs *ctx;  // nCtx entries
  uint32_t nCtx;
  uint32_t nCycles;      // longest run of any context in this launch
  uint32_t *barrier;
  uint32_t vpRows;       // value-pool rows per wave (largest vpBytes)
  uint32_t progWords;    // words of one LDS program buffer (largest partition, 4-multiple)
};

hipError_t launchPlain(const PlainArgs &args, uint32_t grid, uint32_t block, uint32_t lds, hipStream_t stream);
hipError_t plainMaxBlocksPerCU(int *blocks, uint32_t block, uint32_t lds);
hipError_t plainSetLds(uint32_t lds);

// Where the wave64 kernel reads program words from: global memory, an LDS
// copy staged per partition visit, or LDS-resident for the whole launch.
enum ProgMode : int { PM_GLOBAL = 0, PM_STAGED = 1, PM_RESIDENT = 2 };

hipError_t launchWave64(const W64Args &args, int mode, uint32_t grid, uint32_t block, uint32_t lds,
                        hipStream_t stream);
hipError_t wave64MaxBlocksPerCU(int *blocks, int mode, uint32_t block, uint32_t lds);
hipError_t wave64SetLds(uint32_t lds);
uint32_t wave64LutBytes();

// Streamed kernel with `nc` compute waves (2, 4 or 8) plus three loader waves.
bool streamSupportsWaves(uint32_t nc);
uint32_t streamBlock(uint32_t nc);
uint32_t streamLdsBytes(uint32_t nc, uint32_t slots, uint32_t vpBytes, uint32_t maxSteps);  // dynamic part
uint32_t streamStaticLdsBytes(uint32_t nc);
hipError_t launchStream(const StreamArgs &args, uint32_t nc, uint32_t grid, uint32_t lds, hipStream_t stream);
hipError_t streamMaxBlocksPerCU(int *blocks, uint32_t nc, uint32_t lds);
hipError_t streamSetLds(uint32_t lds);

// Batched kernel with `inst` instances per wave (8, 16, 32 or 64; 64 / inst ops
// per round in per-instance tasks) and block / 64 waves per workgroup; `nib`:
// rows hold a nibble per instance (nibble programs whose LUT tasks are all
// T_LUTFAST), else a byte.
bool batchSupportsGroup(uint32_t inst);
uint32_t batchLdsBytes(uint32_t inst, uint32_t waves, uint32_t vpBytes, uint32_t progWords, bool nib);
hipError_t launchBatch(const BatchLaunch &launch, uint32_t inst, bool nib, uint32_t grid, uint32_t block, uint32_t lds,
                       hipStream_t stream);
hipError_t batchMaxBlocksPerCU(int *blocks, uint32_t inst, bool nib, uint32_t block, uint32_t lds);
hipError_t batchSetLds(uint32_t lds);
// hipErrorInvalidKernelFile when a batched kernel was built with static LDS
// (its LUT region must sit at LDS address 0).
hipError_t batchCheckNoStaticLds();
// out[0 .. n) = value (a kernel, so callers that record their work into a
// HIP graph get no memset node).
hipError_t launchFill32(uint32_t *out, uint64_t n, uint32_t value, hipStream_t stream);
hipError_t launchSlotColumn(uint8_t *pool, uint32_t instPad, uint32_t slot, uint64_t rows, uint8_t *dense, bool toPool,
                            hipStream_t stream);
hipError_t launchPermute(const uint8_t *in, uint8_t *out, uint64_t rows, uint32_t instPad, const uint32_t *src,
                         hipStream_t stream);
hipError_t launchPermute64(const uint64_t *in, uint64_t *out, uint64_t rows, uint32_t instPad, const uint32_t *src,
                           hipStream_t stream);
// Testbench mailbox: register nibbles ids[0..n) of nSlots instance slots to
// (toPool) or from a dense buffer vals[instance * n + k]; nibble r of slot s
// lives at pool[r * instPad + s], and slot s holds instance instOf[s]
// (instOf null: slot s holds instance s). With ctlWords > 0, thread 0 also
// copies ctlWords words from ctlSrc to ctlDst (a single instance's control
// words read back with its registers).
hipError_t launchRegMailbox(uint8_t *pool, uint32_t instPad, const uint32_t *instOf, uint32_t nSlots,
                            const uint32_t *ids, uint32_t n, uint8_t *vals, bool toPool, hipStream_t stream,
                            const uint32_t *ctlSrc = nullptr, uint32_t *ctlDst = nullptr, uint32_t ctlWords = 0);

}  // namespace tgs
