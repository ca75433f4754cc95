// C-ABI runtime of the MI355X Toucan cycle simulator (include/toucan_rt.h).
//
// Replaces the reference's process-global device state and launch code
// (src/SimEval.cu:190-215,1058-1343; src/ToucanSim.cu:16-267) with per-context
// state: one context holds one loaded design, its device images and any
// number of resident instances (register / memory / exchange pools each).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "kernel_args.h"
#include "image.hpp"
#include "toucan_rt.h"


using namespace tgs;
using toucanGPUSim::SimDebugInfo;
using toucanGPUSim::SimDesignInfo;

// Scratch memory of tgs_run_batch_noalloc: one static array per GPU, held
// until the process ends. Layout: per-slot cycle counts (u32) and cycle bases
// (u64) of up to kScratchSlots instance slots, the grid-barrier tree, the
// launch's BatchArgs, then the active-group table of up to kScratchCycles
// launch-cycles. In host code
// `scratch` names a host-side stand-in; its device address on each GPU comes
// from hipGetSymbolAddress (scratchFor) and only that address is used.
constexpr uint32_t kScratchSlots = 4096, kScratchCycles = 65536;
constexpr uint32_t kMaxBatchContexts = 64;  // contexts per tgs_run_batch_multi call
constexpr size_t kScratchRunOff = 0, kScratchBaseOff = 4ull * kScratchSlots,
                 kScratchBarrierOff = kScratchBaseOff + 8ull * kScratchSlots, kScratchArgsOff = kScratchBarrierOff + 4ull * kBarTreeWords,
                 kScratchActiveOff = kScratchArgsOff + 512, kScratchBytes = kScratchActiveOff + 4ull * kScratchCycles;
static_assert(sizeof(BatchArgs) <= 512, "the scratch array holds one context's BatchArgs");
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

// Device print-log entries per instance (single-instance layouts) and the
// floor of the shared log of the batched layout. Launches are cut short so
// that one launch can never fire more print events than the log holds; the
// host drains the log into HostLog after every launch that may have printed.
constexpr uint32_t kLogCap = 1u << 16;
constexpr uint64_t kBatchLogMax = 1ull << 24;

// Device buffer owned by a context.
struct DBuf {
  void *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t bytes) {
    reset();
    n = bytes;
    return hipMalloc(&p, bytes ? bytes : 16);
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

// Pinned host memory that kernels read and write directly (coherent, mapped
// into the device's address space); `d` is its device-side address.
struct PBuf {
  void *h = nullptr, *d = nullptr;
  size_t n = 0;
  PBuf() = default;
  PBuf(const PBuf &) = delete;
  PBuf &operator=(const PBuf &) = delete;
  ~PBuf() { reset(); }
  void reset() {
    if (h) (void)hipHostFree(h);
    h = d = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t bytes) {
    reset();
    hipError_t e = hipHostMalloc(&h, bytes ? bytes : 16, hipHostMallocCoherent);
    if (e != hipSuccess) {
      h = nullptr;
      return e;
    }
    std::memset(h, 0, bytes ? bytes : 16);
    n = bytes;
    return hipHostGetDevicePointer(&d, h, 0);
  }
  template <typename T>
  T *host() const {
    return static_cast<T *>(h);
  }
  template <typename T>
  T *dev() const {
    return static_cast<T *>(d);
  }
};

// Waveform snapshots (SimVCD): the register pool of one instance is copied
// to a device staging slot on the context stream, ordered after the cycle
// launches already queued, and from there to pinned host memory on a side
// stream, so the device-to-host copy overlaps the next launches.
struct SnapRing {
  hipStream_t side = nullptr;
  uint32_t slots = 0;
  uint64_t bytes = 0;
  std::vector<void *> host;  // pinned (hipHostMalloc)
  std::vector<hipEvent_t> staged, done;
  std::vector<uint64_t> ticketOf, cycleOf;
  std::vector<uint32_t> probes;  // probe values staged after the register pool
  std::vector<bool> ordered;     // probe values already in probe-list order
  uint64_t next = 0;
  void *staging = nullptr;   // slots * bytes of device memory
  void release() {
    if (side) (void)hipStreamSynchronize(side);
    for (void *h : host) (void)hipHostFree(h);
    for (auto e : staged) (void)hipEventDestroy(e);
    for (auto e : done) (void)hipEventDestroy(e);
    if (staging) (void)hipFree(staging);
    if (side) (void)hipStreamDestroy(side);
    *this = SnapRing();
  }
};

struct Instance {
  DBuf reg, mem, exg, xs, ctl, logCycle, logMsg, probe;
  PBuf ctlPin;  // the control words as the last launch left them (written by the launch's last workgroup)
  uint64_t cycles = 0;
  // The grid-barrier tree (after the control words) is not cleared between
  // launches of one grid size: the next launch starts after barrier index
  // barNext; barGrid 0 means the tree is unknown and the next launch clears
  // it first. ticksSeen: the control block's barrier and loop tick totals
  // already added to the profile.
  uint32_t barNext = 0, barGrid = 0;
  uint64_t ticksSeen[2] = {0, 0};
  // the host rewrote the control words (tick totals zero; the tree is cleared before the next launch)
  void ctlRewritten() {
    barNext = 0;
    barGrid = 0;
    ticksSeen[0] = ticksSeen[1] = 0;
  }
};

// Instance-minor state of the batched kernel: pool nibble k of slot s at
// k * instPad + s. Slots are instances reordered by decreasing cycle count.
struct BatchState {
  uint32_t n = 0, instPad = 0;
  uint64_t logCap = 0;  // entries of the shared print log
  DBuf reg, mem, exg, xs, draws, stop, stopCycle, logCount, logCycle, logMsg, barrier, probe;
  DBuf runCycles, cycleBase, groupsActive, instOfD, perm, scratch, argsDev;
  std::vector<uint32_t> slotOf, instOf;  // instance -> slot, slot -> instance
  std::vector<uint64_t> cycles;          // per instance
};

// Print events of one instance moved off the device (tgs_get_print_log).
struct HostLog {
  std::vector<uint64_t> cyc;
  std::vector<uint32_t> msg;
  uint64_t dropped = 0;  // events a launch fired past the device log's capacity
  void clear() {
    cyc.clear();
    msg.clear();
    dropped = 0;
  }
};

// Host copy of one instance's architectural state (layout conversions).
struct HostInst {
  std::vector<uint8_t> reg, mem, exg, probe;
  uint64_t xs = 0, stopCycle = UINT64_MAX, cycles = 0;
  uint32_t stop = 0;
};

uint32_t up4(uint32_t x) { return (x + 3u) & ~3u; }
uint32_t up16(uint32_t x) { return (x + 15u) & ~15u; }

uint64_t seedFor(uint64_t seed, uint32_t instance) {
  uint64_t z = seed + instance + 0x9E3779B97F4A7C15ull;  // splitmix64(seed + instance)
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return z ? z : 0x9E3779B97F4A7C15ull;
}

}  // namespace

struct tgs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool ownStream = true;  // false: the caller's stream (tgs_create_on_stream), not destroyed here
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipDeviceProp_t prop{};
  std::string err;

  bool loaded = false;
  SimDesignInfo design;
  SimDebugInfo symbols;
  std::vector<uint8_t> initReg, initMem;
  uint64_t designHash = 0;  // FNV-1a of the serialized design (checkpoint identity)
  tgs_program_stats stats{};
  int kernel = TGS_KERNEL_STREAM;
  bool kernelChosen = false;  // tgs_set_kernel was called
  int printEnable = 0;
  StimMap stim;
  uint64_t stimSeed = 0;
  std::vector<std::unique_ptr<Instance>> inst;
  bool batched = false;  // instance state lives in `bs` (batched kernel)
  BatchState bs;
  // Batched launch shape: instances per wave (8, 16, 32 or 64; 0 = chosen at
  // image build), waves per workgroup (0 = chosen) and the effective values.
  uint32_t batchInst = 0, batchWaves = 0, batchInstEff = 64, batchBlockEff = 128;
  uint32_t batchGrid = 0, batchLds = 0, batchGridOverride = 0, batchProgWords = 0;
  uint32_t batchShapeN = 0;  // instance count the batched launch shape was chosen for
  bool batchNib = false;     // the batched rows hold a nibble per instance
  bool batchReady = false;
  // Every byte of every pool is a nibble (true after a load whose initial
  // pools are nibbles, false from the first wider byte written through the
  // API); with a nibble program the batched image then uses T_LUTFAST tasks.
  bool nibbleState = true, batchFast = false;
  // A write of bytes `p`: a byte above 15 ends the nibble state, and a
  // T_LUTFAST image is rebuilt without those tasks before the next run.
  void noteWrite(const uint8_t *p, size_t n) {
    if (!nibbleState) return;
    for (size_t i = 0; i < n; ++i)
      if (p[i] > 15) {
        nibbleState = false;
        if (batchFast) batchReady = false;
        return;
      }
  }

  DBuf lut, inIds;
  bool plainReady = false, w64Ready = false;
  PlainImage plain;
  DBuf pNet, pParts, pLevels, pMparts, pConsts;
  Wave64Image w64;
  Wave64Schedule sched;
  DBuf wWords, wPartOff, wPartWords, wSchedOff, wSchedPart, wSchedLds, wConsts;
  uint32_t w64Block = 512, w64Groups = 0, w64Lds = 0;
  int w64Mode = PM_GLOBAL;
  int w64ModeForce = -1;
  uint32_t plainGrid = 0, plainBlock = 1024, plainLds = 0;
  bool streamReady = false;
  StreamImage simg;
  DBuf sWords, sSteps, sStepBase, sStepCount, sConsts;
  uint32_t streamNC = 0, streamNCUsed = 4, streamSlots = 1, streamGroups = 0, streamGrid = 0, streamLds = 0;

  tgs_profile prof{};
  // per-partition profile (tgs_set_part_profile) of instance 0
  bool partProf = false;
  uint32_t profStart = 100, profCount = 100, profCollected = 0;
  DBuf profUseful, profWg, profOwner;
  // print events moved off the device, per instance; logDirty: a launch may
  // have printed since the last drain
  std::vector<HostLog> logs;
  bool logDirty = false;
  bool stepFailed = false;  // the last tgs_step returned an error after queueing work
  // host bookkeeping of batched runs (kept so steady-state runs allocate nothing)
  std::vector<uint64_t> bhRemaining, bhBase, bhStop;
  std::vector<uint32_t> bhRun, bhActive;
  std::vector<BatchArgs> bhArgs;
  void *scratchDev = nullptr;  // this device's static scratch array (tgs_run_batch_noalloc)
  SnapRing snap;
  bool timingPending = false;  // the last launch's events are not accounted yet
  // testbench mailbox (tgs_step): pinned host memory (read and written by
  // the streamed and wave64 kernels directly) and a device twin (the copy
  // path of the other kernels)
  PBuf box;
  DBuf boxDev;
  // signal probes: request order -> CSR position (CSR = grouped by partition)
  std::vector<uint32_t> probeCsrOf;
  DBuf prOff, prSlot;
  uint32_t nProbe() const { return uint32_t(probeCsrOf.size()); }

  int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }
  int hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return TGS_OK;
    (void)hipGetLastError();
    return fail(TGS_E_HIP, "%s: %s", what, hipGetErrorString(e));
  }
};

#define HIPCHK(ctx, call)                          \
  do {                                             \
    int _rc = (ctx)->hip((call), #call);           \
    if (_rc != TGS_OK) return _rc;                 \
  } while (0)

namespace {

int setDevice(tgs_ctx *c) { return c->hip(hipSetDevice(c->device), "hipSetDevice"); }

// Workgroups per CU that are certain to be co-resident, for the persistent
// kernels whose grid barrier needs every workgroup of the grid resident.
// The hardware allocates a workgroup's LDS in whole granules while the
// occupancy query can count the requested bytes exactly, so a grid packed to
// within a granule of the CU's LDS can exceed what is actually resident and
// the grid barrier then never completes. Each workgroup is charged one more
// granule's worth (kLdsGranuleMargin, at least the allocation granule) on top
// of the query's answer.
constexpr uint32_t kLdsGranuleMargin = 1280;
int residentPerCU(const tgs_ctx *c, int queried, uint32_t ldsPerGroup) {
  const uint64_t perCU = c->prop.maxSharedMemoryPerMultiProcessor ? c->prop.maxSharedMemoryPerMultiProcessor
                                                                  : c->prop.sharedMemPerBlock;
  const uint64_t byLds = ldsPerGroup ? perCU / (uint64_t(ldsPerGroup) + kLdsGranuleMargin) : uint64_t(queried);
  return int(std::min<uint64_t>(uint64_t(std::max(queried, 0)), byLds));
}

int upload(tgs_ctx *c, DBuf &b, const void *src, size_t n) {
  HIPCHK(c, b.alloc(n));
  if (n) HIPCHK(c, hipMemcpyAsync(b.p, src, n, hipMemcpyHostToDevice, c->stream));
  return TGS_OK;
}

// Reset one instance to the loaded initial state (copy_netlist_to_gpu,
// SimEval.cu:1114-1139: stop flag and cycle counter cleared, exchange zeroed).
int resetInstance(tgs_ctx *c, Instance &in, uint32_t index, uint64_t seed) {
  const auto &d = c->design;
  HIPCHK(c, in.reg.alloc(up4(d.regPoolSize) + 16));
  HIPCHK(c, in.mem.alloc(size_t(d.memPoolSize) + 16));
  HIPCHK(c, in.exg.alloc(up4(d.exchangePoolSize) + 16));
  HIPCHK(c, in.xs.alloc(8));
  HIPCHK(c, in.ctl.alloc(kCtlAllocWords * 4));
  if (!in.ctlPin.h) HIPCHK(c, in.ctlPin.alloc(C_WORDS * 4));
  HIPCHK(c, in.logCycle.alloc(kLogCap * 8));
  HIPCHK(c, in.logMsg.alloc(kLogCap * 4));
  HIPCHK(c, in.probe.alloc(std::max<uint32_t>(c->nProbe(), 1)));
  HIPCHK(c, hipMemsetAsync(in.probe.p, 0, in.probe.n, c->stream));
  HIPCHK(c, hipMemsetAsync(in.reg.p, 0, in.reg.n, c->stream));
  HIPCHK(c, hipMemsetAsync(in.mem.p, 0, in.mem.n, c->stream));
  HIPCHK(c, hipMemsetAsync(in.exg.p, 0, in.exg.n, c->stream));
  if (!c->initReg.empty())
    HIPCHK(c, hipMemcpyAsync(in.reg.p, c->initReg.data(), c->initReg.size(), hipMemcpyHostToDevice, c->stream));
  if (!c->initMem.empty())
    HIPCHK(c, hipMemcpyAsync(in.mem.p, c->initMem.data(), c->initMem.size(), hipMemcpyHostToDevice, c->stream));
  uint32_t ctl[C_WORDS] = {0};
  uint64_t never = UINT64_MAX;
  std::memcpy(&ctl[C_STOPCYC], &never, 8);
  HIPCHK(c, hipMemcpyAsync(in.ctl.p, ctl, sizeof ctl, hipMemcpyHostToDevice, c->stream));
  std::vector<uint64_t> xs(1, seedFor(seed, index));
  HIPCHK(c, hipMemcpyAsync(in.xs.p, xs.data(), 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  in.cycles = 0;
  in.ctlRewritten();
  return TGS_OK;
}

int buildPlain(tgs_ctx *c) {
  if (c->plainReady) return TGS_OK;
  c->plain = buildPlainImage(c->design, c->stim);
  auto &im = c->plain;
  int rc;
  if ((rc = upload(c, c->pNet, im.net.data(), im.net.size()))) return rc;
  if ((rc = upload(c, c->pParts, im.parts.data(), im.parts.size() * sizeof(PlainPart)))) return rc;
  if ((rc = upload(c, c->pLevels, im.levels.data(), im.levels.size() * 4))) return rc;
  if ((rc = upload(c, c->pMparts, im.mparts.data(), im.mparts.size() * 4))) return rc;
  if ((rc = upload(c, c->pConsts, im.consts.data(), im.consts.size()))) return rc;
  c->plainLds = up16(std::max<uint32_t>(im.maxVp, 16));
  if (c->plainLds > c->prop.sharedMemPerBlock)
    return c->fail(TGS_E_LIMIT, "value pool needs %u B of LDS, device allows %zu B (ToucanSim.cu:194-197)",
                   c->plainLds, size_t(c->prop.sharedMemPerBlock));
  HIPCHK(c, plainSetLds(c->plainLds));
  int perCU = 0;
  HIPCHK(c, plainMaxBlocksPerCU(&perCU, c->plainBlock, c->plainLds));
  perCU = residentPerCU(c, perCU, c->plainLds);
  if (perCU < 1) return c->fail(TGS_E_LIMIT, "plain kernel cannot be resident with %u B LDS", c->plainLds);
  // The reference forces one block per SM for the multi-cycle kernel (ToucanSim.cu:80,189).
  c->plainGrid = std::max<uint32_t>(1, std::min<uint32_t>(im.maxRegion, c->prop.multiProcessorCount));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->plainReady = true;
  return TGS_OK;
}

int buildW64(tgs_ctx *c) {
  if (c->w64Ready) return TGS_OK;
  const uint32_t nWaves = c->w64Block / 64;
  c->w64 = buildWave64Image(c->design, c->stim, nWaves);
  auto &im = c->w64;
  uint32_t maxRegion = std::max<uint32_t>(im.n0, im.nTotal - im.n0);
  uint32_t groups = c->w64Groups ? c->w64Groups : std::min<uint32_t>(maxRegion, c->prop.multiProcessorCount);
  groups = std::max<uint32_t>(groups, 1);
  c->sched = scheduleWave64(im, groups);
  const uint32_t base = wave64LutBytes() + up16(std::max<uint32_t>(im.maxVp, 16)) + kMaxDraws * 8;
  const uint32_t resLds = base + c->sched.maxGroupWords * 4;
  const size_t limit = c->prop.sharedMemPerBlock - kLdsGranuleMargin;
  uint32_t maxPart = 0;
  for (auto w : im.partWords) maxPart = std::max(maxPart, (w + 3) & ~3u);
  const uint32_t stagedLds = base + maxPart * 4;
  c->w64Mode = resLds <= limit ? PM_RESIDENT : stagedLds <= limit ? PM_STAGED : PM_GLOBAL;
  if (c->w64ModeForce >= 0 && c->w64ModeForce < c->w64Mode) c->w64Mode = c->w64ModeForce;
  c->w64Lds = c->w64Mode == PM_RESIDENT ? resLds : c->w64Mode == PM_STAGED ? stagedLds : base;
  if (base > limit) return c->fail(TGS_E_LIMIT, "value pool needs %u B of LDS, device allows %zu B", base, limit);
  int rc;
  if ((rc = upload(c, c->wWords, im.words.data(), im.words.size() * 4))) return rc;
  if ((rc = upload(c, c->wPartOff, im.partOff.data(), im.partOff.size() * 4))) return rc;
  if ((rc = upload(c, c->wPartWords, im.partWords.data(), im.partWords.size() * 4))) return rc;
  if ((rc = upload(c, c->wSchedOff, c->sched.off.data(), c->sched.off.size() * 4))) return rc;
  if ((rc = upload(c, c->wSchedPart, c->sched.parts.data(), c->sched.parts.size() * 4))) return rc;
  if ((rc = upload(c, c->wSchedLds, c->sched.ldsOff.data(), c->sched.ldsOff.size() * 4))) return rc;
  if ((rc = upload(c, c->wConsts, im.consts.data(), im.consts.size()))) return rc;
  HIPCHK(c, wave64SetLds(c->w64Lds));
  int perCU = 0;
  HIPCHK(c, wave64MaxBlocksPerCU(&perCU, c->w64Mode, c->w64Block, c->w64Lds));
  perCU = residentPerCU(c, perCU, c->w64Lds);
  if (perCU < 1) return c->fail(TGS_E_LIMIT, "wave64 kernel cannot be resident with %u B LDS", c->w64Lds);
  if (uint64_t(perCU) * c->prop.multiProcessorCount < groups)
    return c->fail(TGS_E_LIMIT, "%u workgroups exceed the co-resident limit", groups);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->w64Ready = true;
  return TGS_OK;
}

// Streamed kernel: one workgroup per partition slot (up to the co-resident
// limit), NC compute waves plus a loader wave each.
int buildStream(tgs_ctx *c) {
  if (c->streamReady) return TGS_OK;
  const uint32_t maxRegion = std::max<uint32_t>(c->design.regionPartitionIds[0].size(),
                                                c->design.parts.size() - c->design.regionPartitionIds[0].size());
  const uint32_t cus = c->prop.multiProcessorCount;
  uint32_t nc = c->streamNC;
  if (!nc) {
    // compute waves per workgroup (LUT micropartitions pair up in a wave)
    std::vector<uint32_t> per;
    for (const auto &p : c->design.parts)
      for (const auto &lv : p.exec_mParts) {
        uint32_t luts = 0, other = 0;
        for (const auto &m : lv) (m.isLUTPart ? luts : other)++;
        per.push_back((luts + 1) / 2 + other);
      }
    // the widest exec level decides: levels wider than the compute waves split
    // into extra barrier-separated steps
    uint32_t widest = 1;
    for (uint32_t v : per) widest = std::max(widest, v);
    nc = widest > 4 ? 8 : widest > 2 ? 4 : 2;
  }
  c->streamNCUsed = nc;
  // Launch shape: partition slots per workgroup (up to 1, 2 or 4 side by
  // side, chosen per region by the image builder) and workgroups (one or two
  // per CU). Each candidate's image estimates its cycle time (per region the
  // slowest workgroup's steps, weighted by step type and by the cost of
  // split steps); a step costs about 1.4x when two workgroups share a CU (the
  // dependent LDS chains of both contend, measured on the 10M-LUT4 design).
  // The cheapest candidate that fits runs.
  constexpr double kTwoPerCUStep = 1.4;
  // slot counts worth building for `g` workgroups: more than one only when
  // some workgroup holds several partitions of a region
  const char *forced = std::getenv("TGS_STREAM_SLOTS");  // every region laid out for this many slots
  auto slotsFor = [&](uint32_t g) {
    std::vector<uint32_t> l{1};
    if (forced) {
      const uint32_t sl = uint32_t(std::max(1, std::min(std::atoi(forced), int(nc))));
      l.assign(1, nc % sl ? 1u : sl);
      return l;
    }
    if (maxRegion > g)
      for (uint32_t sl : {2u, 4u})
        if (nc % sl == 0 && nc / sl >= 2) l.push_back(sl);
    return l;
  };
  std::vector<uint32_t> groupList;
  if (c->streamGroups) {
    groupList.push_back(c->streamGroups);
  } else {
    groupList.push_back(std::max<uint32_t>(std::min<uint32_t>(maxRegion, cus), 1));
    if (maxRegion > cus) groupList.push_back(std::min<uint32_t>(maxRegion, 2 * cus));
  }
  const size_t limit = c->prop.sharedMemPerBlock - kLdsGranuleMargin;
  double best = 0;
  bool have = false;
  uint32_t groups = 0, slots = 1, lds = 0;
  std::string why;
  for (uint32_t g : groupList)
    for (uint32_t sl : slotsFor(g)) {
      StreamImage img = buildStreamImage(c->design, c->stim, nc, g, sl, forced != nullptr);
      const uint32_t vpBytes = up16(std::max<uint32_t>(img.maxVp, 16));
      const uint32_t l = streamLdsBytes(nc, sl, vpBytes, img.maxSteps);
      if (l + streamStaticLdsBytes(nc) > limit) {
        char buf[256];
        snprintf(buf, sizeof buf, "streamed kernel needs %u B of LDS (value pool %u, %u steps per cycle); device allows %zu B",
                 l + streamStaticLdsBytes(nc), img.maxVp, img.maxSteps, limit);
        why = buf;
        continue;
      }
      HIPCHK(c, streamSetLds(l));
      int perCU = 0;
      HIPCHK(c, streamMaxBlocksPerCU(&perCU, nc, l));
      perCU = residentPerCU(c, perCU, l + streamStaticLdsBytes(nc));
      if (uint64_t(perCU) * cus < g) {
        char buf[256];
        snprintf(buf, sizeof buf, "%u workgroups exceed the co-resident limit %llu", g,
                 (unsigned long long)(uint64_t(std::max(perCU, 0)) * cus));
        why = buf;
        continue;
      }
      const double score = img.estCycle * (g > cus ? kTwoPerCUStep : 1.0);
      if (!have || score < best) {
        have = true;
        best = score;
        groups = g;
        slots = sl;
        lds = l;
        c->simg = std::move(img);
      }
    }
  if (!have) return c->fail(TGS_E_LIMIT, "%s", why.c_str());
  c->streamSlots = slots;
  HIPCHK(c, streamSetLds(lds));
  const auto &im = c->simg;
  int rc;
  if ((rc = upload(c, c->sWords, im.words.data(), im.words.size() * 4))) return rc;
  if ((rc = upload(c, c->sSteps, im.steps.data(), im.steps.size() * 4))) return rc;
  if ((rc = upload(c, c->sStepBase, im.stepBase.data(), im.stepBase.size() * 4))) return rc;
  if ((rc = upload(c, c->sStepCount, im.stepCount.data(), im.stepCount.size() * 4))) return rc;
  if ((rc = upload(c, c->sConsts, im.consts.data(), im.consts.size()))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->streamGrid = groups;
  c->streamLds = lds;
  c->streamReady = true;
  return TGS_OK;
}

CommonArgs commonArgs(tgs_ctx *c, Instance &in, uint32_t cycles) {
  CommonArgs a{};
  a.lut = c->lut.as<uint8_t>();
  a.reg = in.reg.as<uint8_t>();
  a.mem = in.mem.as<uint8_t>();
  a.exg = in.exg.as<uint8_t>();
  a.regSize = c->design.regPoolSize;
  a.memSize = c->design.memPoolSize;
  a.exgSize = c->design.exchangePoolSize;
  a.xs = in.xs.as<uint64_t>();
  a.inIds = c->inIds.as<uint32_t>();
  a.nIn = c->stim.ids.size();
  a.ctl = in.ctl.as<uint32_t>();
  a.logCycle = in.logCycle.as<uint64_t>();
  a.logMsg = in.logMsg.as<uint32_t>();
  a.logCap = kLogCap;
  a.printEnable = c->printEnable;
  a.cycleBase = in.cycles;
  a.nCycles = cycles;
  if (c->nProbe()) {
    a.probeOff = c->prOff.as<uint32_t>();
    a.probeSlot = c->prSlot.as<uint32_t>();
    a.probeVal = in.probe.as<uint8_t>();
  }
  if (c->partProf && !c->inst.empty() && &in == c->inst[0].get()) {
    a.profUseful = c->profUseful.as<uint64_t>();
    a.profWg = c->profWg.as<uint64_t>();
    a.profOwner = c->profOwner.as<uint32_t>();
    a.profStart = c->profStart;
    a.profCount = c->profCount;
  }
  return a;
}

// Largest launch (in cycles) whose print events always fit the device log:
// every print op fires at most once per cycle and instance.
uint32_t printCycleCap(const tgs_ctx *c, uint64_t logCap, uint64_t instances) {
  if (!c->printEnable || c->stats.n_print == 0) return UINT32_MAX;
  return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(UINT32_MAX, logCap / (c->stats.n_print * instances))));
}

// Move the device print logs to the instances' host logs and empty them.
int drainLogs(tgs_ctx *c) {
  if (!c->logDirty) return TGS_OK;
  c->logDirty = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const uint32_t n = c->batched ? c->bs.n : uint32_t(c->inst.size());
  if (c->logs.size() < n) c->logs.resize(n);
  const uint64_t zero = 0;
  if (!c->batched) {
    for (uint32_t i = 0; i < c->inst.size(); ++i) {
      Instance &in = *c->inst[i];
      uint64_t count = 0;
      HIPCHK(c, hipMemcpy(&count, in.ctl.as<uint32_t>() + C_LOGCOUNT, 8, hipMemcpyDeviceToHost));
      if (count == 0) continue;
      const uint64_t k = std::min<uint64_t>(count, kLogCap);
      HostLog &L = c->logs[i];
      const size_t at = L.cyc.size();
      L.cyc.resize(at + k);
      L.msg.resize(at + k);
      HIPCHK(c, hipMemcpy(L.cyc.data() + at, in.logCycle.p, k * 8, hipMemcpyDeviceToHost));
      HIPCHK(c, hipMemcpy(L.msg.data() + at, in.logMsg.p, k * 4, hipMemcpyDeviceToHost));
      L.dropped += count - k;
      HIPCHK(c, hipMemcpy(in.ctl.as<uint32_t>() + C_LOGCOUNT, &zero, 8, hipMemcpyHostToDevice));
    }
    return TGS_OK;
  }
  auto &b = c->bs;
  uint64_t count = 0;
  HIPCHK(c, hipMemcpy(&count, b.logCount.p, 8, hipMemcpyDeviceToHost));
  if (count == 0) return TGS_OK;
  const uint64_t k = std::min<uint64_t>(count, b.logCap);
  std::vector<uint64_t> cy(k), mg(k);
  HIPCHK(c, hipMemcpy(cy.data(), b.logCycle.p, k * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(mg.data(), b.logMsg.p, k * 8, hipMemcpyDeviceToHost));
  for (uint64_t e = 0; e < k; ++e) {
    const uint32_t inst = uint32_t(mg[e] >> 32);
    if (inst >= n) continue;
    c->logs[inst].cyc.push_back(cy[e]);
    c->logs[inst].msg.push_back(uint32_t(mg[e]));
  }
  if (count > k) c->logs[0].dropped += count - k;  // instance unknown: reported on instance 0
  HIPCHK(c, hipMemcpy(b.logCount.p, &zero, 8, hipMemcpyHostToDevice));
  return TGS_OK;
}

// Before a launch of instance 0 that reaches the profile window: clear the
// sums the kernel adds to.
int profBegin(tgs_ctx *c, Instance &in, uint32_t n) {
  if (!c->partProf || c->inst.empty() || &in != c->inst[0].get() || n <= c->profStart) return TGS_OK;
  HIPCHK(c, hipMemsetAsync(c->profUseful.p, 0, c->profUseful.n, c->stream));
  HIPCHK(c, hipMemsetAsync(c->profWg.p, 0, c->profWg.n, c->stream));
  c->profCollected = std::min(c->profCount, n - c->profStart);
  return TGS_OK;
}

// Largest cycle count per launch that keeps the barrier counter in range
// (two grid barriers per cycle, plus the mailbox's and the closing arrival).
uint32_t maxCyclesPerLaunch(uint32_t groups) {
  uint64_t m = (0xF0000000ull / (2ull * std::max<uint32_t>(groups, 1))) - 2;
  return uint32_t(std::min<uint64_t>(m, 1u << 30));
}

// Barrier base of a launch of `grid` workgroups that waits at `barriers`
// grid barriers and then arrives once more (finishLaunch). The tree is
// cleared first, by a fill kernel, when it counted for another grid or is
// unknown, or when this launch would carry a counter out of range (no
// counter exceeds the barrier index times the grid); otherwise it goes on
// from the last launch and no clear precedes the launch.
int barrierBaseFor(tgs_ctx *c, Instance &in, uint32_t grid, uint64_t barriers, uint32_t *base) {
  if (in.barGrid != grid || (uint64_t(in.barNext) + barriers + 1) * grid >= 0xF0000000ull) {
    HIPCHK(c, launchFill32(in.ctl.as<uint32_t>() + kCtlTree, kBarTreeWords, 0, c->stream));
    in.barNext = 0;
    in.barGrid = grid;
  }
  *base = in.barNext;
  in.barNext += uint32_t(barriers + 1);
  return TGS_OK;
}

// Barrier and loop ticks of the instance's launches so far: `ctlWords` are
// the control words its last completed launch copied out (cumulative
// totals); the profile takes what was not yet counted.
void accountTicks(tgs_ctx *c, Instance &in, const uint32_t *ctlWords) {
  uint64_t t[2];
  std::memcpy(t, ctlWords + C_BARTICKS, 16);
  c->prof.barrier_ticks += t[0] - in.ticksSeen[0];
  c->prof.total_ticks += t[1] - in.ticksSeen[1];
  in.ticksSeen[0] = t[0];
  in.ticksSeen[1] = t[1];
}

// A testbench step's mailbox (device addresses of pinned host memory): the
// register writes go with the step's first launch, the register reads with
// its last (streamed and wave64 kernels).
struct StepBox {
  const uint32_t *writeId;
  const uint8_t *writeVal;
  uint32_t nWrite;
  const uint32_t *readId;
  uint8_t *readVal;
  uint32_t nRead;
};

// Launch `cycles` cycles of one instance. With `wait` each launch is waited
// for and timed at once; without it the last launch's events stay pending
// (c->timingPending) and finishTiming() accounts them after the caller's own
// synchronization, so a testbench step costs one round trip. Every streamed
// or wave64 launch leaves its control words in in.ctlPin.
int runInstance(tgs_ctx *c, Instance &in, uint32_t cycles, bool wait = true, const StepBox *mb = nullptr) {
  if (c->timingPending && cycles > 0) {  // a step left its launch untimed: account it before reusing the events
    c->timingPending = false;
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->prof.kernel_ms += ms;
  }
  const uint32_t printCap = printCycleCap(c, kLogCap, 1);
  bool firstLaunch = true;
  // mailbox and control-word arguments of a streamed or wave64 launch of n cycles
  auto boxArgs = [&](CommonArgs &a, uint32_t n) {
    a.mbCtl = in.ctlPin.dev<uint32_t>();
    if (!mb) return;
    if (firstLaunch) {
      a.mbWriteId = mb->writeId;
      a.mbWriteVal = mb->writeVal;
      a.mbWrites = mb->nWrite;
    }
    if (n == cycles) {
      a.mbReadId = mb->readId;
      a.mbReadVal = mb->readVal;
      a.mbReads = mb->nRead;
    }
  };
  while (cycles > 0) {
    int rc = drainLogs(c);  // every launch starts with an empty device log
    if (rc) return rc;
    if (c->kernel == TGS_KERNEL_PLAIN) {
      if ((rc = buildPlain(c))) return rc;
      uint32_t n = std::min({cycles, maxCyclesPerLaunch(c->plainGrid), printCap});
      if ((rc = profBegin(c, in, n))) return rc;
      PlainArgs a{};
      a.c = commonArgs(c, in, n);
      a.c.consts = c->pConsts.as<uint8_t>();
      a.net = c->pNet.as<uint8_t>();
      a.parts = c->pParts.as<PlainPart>();
      a.levels = c->pLevels.as<uint32_t>();
      a.mparts = c->pMparts.as<uint32_t>();
      a.n0 = c->plain.n0;
      a.nTotal = c->plain.nTotal;
      HIPCHK(c, hipEventRecord(c->ev0, c->stream));
      HIPCHK(c, launchPlain(a, c->plainGrid, c->plainBlock, c->plainLds, c->stream));
      HIPCHK(c, hipEventRecord(c->ev1, c->stream));
      c->prof.grid = c->plainGrid;
      c->prof.block = c->plainBlock;
      c->prof.lds_bytes = c->plainLds;
      c->prof.resident = 0;
      c->prof.prog_bytes = c->plain.net.size();
      cycles -= n;
      in.cycles += n;
      c->prof.cycles += n;
    } else if (c->kernel == TGS_KERNEL_STREAM) {
      if ((rc = buildStream(c))) {
        if (rc != TGS_E_LIMIT || c->kernelChosen) return rc;
        c->kernel = TGS_KERNEL_WAVE64;  // the default family does not fit: the wave64 kernel runs it
        continue;
      }
      uint32_t n = std::min({cycles, maxCyclesPerLaunch(c->streamGrid), printCap});
      if ((rc = profBegin(c, in, n))) return rc;
      StreamArgs a{};
      a.c = commonArgs(c, in, n);
      a.c.consts = c->sConsts.as<uint8_t>();
      a.words = c->sWords.as<uint32_t>();
      a.steps = c->sSteps.as<uint32_t>();
      a.stepBase = c->sStepBase.as<uint32_t>();
      a.stepCount = c->sStepCount.as<uint32_t>();
      a.vpBytes = up16(std::max<uint32_t>(c->simg.maxVp, 16));
      a.slots = c->streamSlots;
      boxArgs(a.c, n);
      if ((rc = barrierBaseFor(c, in, c->streamGrid, 2ull * n + (a.c.mbWrites ? 1 : 0), &a.barrierBase))) return rc;
      HIPCHK(c, hipEventRecord(c->ev0, c->stream));
      if (hipError_t e = launchStream(a, c->streamNCUsed, c->streamGrid, c->streamLds, c->stream); e != hipSuccess) {
        in.barGrid = 0;  // the counter is unknown after a failed launch
        HIPCHK(c, e);
      }
      HIPCHK(c, hipEventRecord(c->ev1, c->stream));
      c->prof.grid = c->streamGrid;
      c->prof.block = streamBlock(c->streamNCUsed);
      c->prof.lds_bytes = c->streamLds + streamStaticLdsBytes(c->streamNCUsed);
      c->prof.resident = 3;
      c->prof.prog_bytes = c->simg.words.size() * 4;
      cycles -= n;
      in.cycles += n;
      c->prof.cycles += n;
    } else {
      if ((rc = buildW64(c))) return rc;
      uint32_t n = std::min({cycles, maxCyclesPerLaunch(c->sched.groups), printCap});
      if ((rc = profBegin(c, in, n))) return rc;
      W64Args a{};
      a.c = commonArgs(c, in, n);
      a.c.consts = c->wConsts.as<uint8_t>();
      a.words = c->wWords.as<uint32_t>();
      a.partOff = c->wPartOff.as<uint32_t>();
      a.partWords = c->wPartWords.as<uint32_t>();
      a.schedOff = c->wSchedOff.as<uint32_t>();
      a.schedPart = c->wSchedPart.as<uint32_t>();
      a.schedLds = c->wSchedLds.as<uint32_t>();
      a.vpBytes = up16(std::max<uint32_t>(c->w64.maxVp, 16));
      boxArgs(a.c, n);
      if ((rc = barrierBaseFor(c, in, c->sched.groups, 2ull * n + (a.c.mbWrites ? 1 : 0), &a.barrierBase)))
        return rc;
      HIPCHK(c, hipEventRecord(c->ev0, c->stream));
      if (hipError_t e = launchWave64(a, c->w64Mode, c->sched.groups, c->w64Block, c->w64Lds, c->stream);
          e != hipSuccess) {
        in.barGrid = 0;  // the counter is unknown after a failed launch
        HIPCHK(c, e);
      }
      HIPCHK(c, hipEventRecord(c->ev1, c->stream));
      c->prof.grid = c->sched.groups;
      c->prof.block = c->w64Block;
      c->prof.lds_bytes = c->w64Lds;
      c->prof.resident = uint32_t(c->w64Mode);
      c->prof.prog_bytes = c->w64.words.size() * 4;
      cycles -= n;
      in.cycles += n;
      c->prof.cycles += n;
    }
    firstLaunch = false;
    c->prof.launches += 1;
    c->prof.kernel = c->kernel;
    if (c->printEnable && c->stats.n_print) c->logDirty = true;
    if (!wait && cycles == 0) {
      c->timingPending = true;
      break;
    }
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->prof.kernel_ms += ms;
    if (c->kernel == TGS_KERNEL_WAVE64 || c->kernel == TGS_KERNEL_STREAM) accountTicks(c, in, in.ctlPin.host<uint32_t>());
  }
  return TGS_OK;
}

// ---------------------------------------------------------------------------
// Instance state in either layout, through host copies (layout conversion,
// instance count changes).

int initialHost(tgs_ctx *c, uint32_t n, std::vector<HostInst> &h) {
  const auto &d = c->design;
  h.assign(n, HostInst());
  for (uint32_t i = 0; i < n; ++i) {
    h[i].reg.assign(d.regPoolSize, 0);
    h[i].mem.assign(d.memPoolSize, 0);
    h[i].exg.assign(d.exchangePoolSize, 0);
    std::copy(c->initReg.begin(), c->initReg.end(), h[i].reg.begin());
    std::copy(c->initMem.begin(), c->initMem.end(), h[i].mem.begin());
    h[i].xs = seedFor(c->stimSeed, i);
    h[i].probe.assign(c->nProbe(), 0);
  }
  return TGS_OK;
}

int instUpload(tgs_ctx *c, const std::vector<HostInst> &h) {
  c->inst.clear();
  for (uint32_t i = 0; i < h.size(); ++i) {
    c->inst.emplace_back(new Instance());
    Instance &in = *c->inst.back();
    int rc = resetInstance(c, in, i, c->stimSeed);
    if (rc) return rc;
    const auto &x = h[i];
    if (!x.reg.empty()) HIPCHK(c, hipMemcpy(in.reg.p, x.reg.data(), x.reg.size(), hipMemcpyHostToDevice));
    if (!x.mem.empty()) HIPCHK(c, hipMemcpy(in.mem.p, x.mem.data(), x.mem.size(), hipMemcpyHostToDevice));
    if (!x.exg.empty()) HIPCHK(c, hipMemcpy(in.exg.p, x.exg.data(), x.exg.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(in.xs.p, &x.xs, 8, hipMemcpyHostToDevice));
    if (x.probe.size() == c->nProbe() && !x.probe.empty())
      HIPCHK(c, hipMemcpy(in.probe.p, x.probe.data(), x.probe.size(), hipMemcpyHostToDevice));
    uint32_t ctl[C_WORDS] = {0};
    ctl[C_STOP] = x.stop;
    std::memcpy(&ctl[C_STOPCYC], &x.stopCycle, 8);
    HIPCHK(c, hipMemcpy(in.ctl.p, ctl, sizeof ctl, hipMemcpyHostToDevice));
    in.ctlRewritten();
    in.cycles = x.cycles;
  }
  return TGS_OK;
}

int instDownload(tgs_ctx *c, std::vector<HostInst> &h) {
  const auto &d = c->design;
  h.assign(c->inst.size(), HostInst());
  for (uint32_t i = 0; i < c->inst.size(); ++i) {
    Instance &in = *c->inst[i];
    auto &x = h[i];
    x.reg.resize(d.regPoolSize);
    x.mem.resize(d.memPoolSize);
    x.exg.resize(d.exchangePoolSize);
    if (!x.reg.empty()) HIPCHK(c, hipMemcpy(x.reg.data(), in.reg.p, x.reg.size(), hipMemcpyDeviceToHost));
    if (!x.mem.empty()) HIPCHK(c, hipMemcpy(x.mem.data(), in.mem.p, x.mem.size(), hipMemcpyDeviceToHost));
    if (!x.exg.empty()) HIPCHK(c, hipMemcpy(x.exg.data(), in.exg.p, x.exg.size(), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(&x.xs, in.xs.p, 8, hipMemcpyDeviceToHost));
    x.probe.resize(c->nProbe());
    if (!x.probe.empty()) HIPCHK(c, hipMemcpy(x.probe.data(), in.probe.p, x.probe.size(), hipMemcpyDeviceToHost));
    uint32_t ctl[C_WORDS];
    HIPCHK(c, hipMemcpy(ctl, in.ctl.p, sizeof ctl, hipMemcpyDeviceToHost));
    x.stop = ctl[C_STOP];
    std::memcpy(&x.stopCycle, &ctl[C_STOPCYC], 8);
    x.cycles = in.cycles;
  }
  return TGS_OK;
}

uint64_t regRows(const SimDesignInfo &d) { return up4(d.regPoolSize) + 4; }
uint64_t memRows(const SimDesignInfo &d) { return uint64_t(d.memPoolSize) + 1; }
uint64_t exgRows(const SimDesignInfo &d) { return up4(d.exchangePoolSize) + 4; }

int batchUpload(tgs_ctx *c, const std::vector<HostInst> &h) {
  auto &b = c->bs;
  const auto &d = c->design;
  const uint32_t n = h.size();
  b.n = n;
  b.instPad = (n + 63) / 64 * 64;
  const size_t P = b.instPad;
  HIPCHK(c, b.reg.alloc(regRows(d) * P));
  HIPCHK(c, b.mem.alloc(memRows(d) * P));
  HIPCHK(c, b.exg.alloc(exgRows(d) * P));
  HIPCHK(c, b.xs.alloc(8 * P));
  HIPCHK(c, b.draws.alloc(8 * P * 2 * kMaxDraws));
  HIPCHK(c, b.stop.alloc(8 * P));
  HIPCHK(c, b.stopCycle.alloc(8 * P));
  b.logCap = std::max<uint64_t>(kLogCap, std::min<uint64_t>(kBatchLogMax, c->stats.n_print * P));
  HIPCHK(c, b.logCount.alloc(8));
  HIPCHK(c, b.logCycle.alloc(b.logCap * 8));
  HIPCHK(c, b.logMsg.alloc(b.logCap * 8));
  HIPCHK(c, b.barrier.alloc(kBarTreeWords * 4));
  HIPCHK(c, b.runCycles.alloc(4 * P));
  HIPCHK(c, b.cycleBase.alloc(8 * P));
  HIPCHK(c, b.instOfD.alloc(4 * P));
  HIPCHK(c, b.perm.alloc(4 * P));
  HIPCHK(c, b.probe.alloc(std::max<uint64_t>(c->nProbe(), 1) * P));
  HIPCHK(c, b.scratch.alloc((std::max<uint64_t>({regRows(d), memRows(d), exgRows(d), uint64_t(c->nProbe())}) + 8) * P));
  b.slotOf.resize(n);
  b.instOf.assign(P, 0xFFFFFFFFu);
  b.cycles.assign(n, 0);
  for (uint32_t i = 0; i < n; ++i) {
    b.slotOf[i] = i;
    b.instOf[i] = i;
    b.cycles[i] = h[i].cycles;
  }
  auto put = [&](DBuf &buf, uint64_t rows, auto field) -> int {
    std::vector<uint8_t> a(rows * P, 0);
    for (uint32_t i = 0; i < n; ++i) {
      const auto &v = field(h[i]);
      for (size_t r = 0; r < v.size(); ++r) a[r * P + i] = v[r];
    }
    HIPCHK(c, hipMemcpy(buf.p, a.data(), a.size(), hipMemcpyHostToDevice));
    return TGS_OK;
  };
  int rc;
  if ((rc = put(b.reg, regRows(d), [](const HostInst &x) -> const std::vector<uint8_t> & { return x.reg; }))) return rc;
  if ((rc = put(b.mem, memRows(d), [](const HostInst &x) -> const std::vector<uint8_t> & { return x.mem; }))) return rc;
  if ((rc = put(b.exg, exgRows(d), [](const HostInst &x) -> const std::vector<uint8_t> & { return x.exg; }))) return rc;
  if (c->nProbe() &&
      (rc = put(b.probe, c->nProbe(), [](const HostInst &x) -> const std::vector<uint8_t> & { return x.probe; })))
    return rc;
  std::vector<uint64_t> xs(P, 0), st(P, 0), sc(P, UINT64_MAX);
  for (uint32_t i = 0; i < n; ++i) {
    xs[i] = h[i].xs;
    st[i] = h[i].stop;
    sc[i] = h[i].stopCycle;
  }
  HIPCHK(c, hipMemcpy(b.xs.p, xs.data(), 8 * P, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(b.stop.p, st.data(), 8 * P, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(b.stopCycle.p, sc.data(), 8 * P, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemset(b.logCount.p, 0, 8));
  HIPCHK(c, hipMemcpy(b.instOfD.p, b.instOf.data(), 4 * P, hipMemcpyHostToDevice));
  return TGS_OK;
}

int batchDownload(tgs_ctx *c, std::vector<HostInst> &h) {
  auto &b = c->bs;
  const auto &d = c->design;
  const size_t P = b.instPad;
  h.assign(b.n, HostInst());
  auto get = [&](DBuf &buf, uint64_t rows, uint32_t keep, auto field) -> int {
    std::vector<uint8_t> a(rows * P);
    HIPCHK(c, hipMemcpy(a.data(), buf.p, a.size(), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < b.n; ++i) {
      auto &v = field(h[i]);
      v.resize(keep);
      for (uint32_t r = 0; r < keep; ++r) v[r] = a[r * P + b.slotOf[i]];
    }
    return TGS_OK;
  };
  int rc;
  if ((rc = get(b.reg, regRows(d), d.regPoolSize, [](HostInst &x) -> std::vector<uint8_t> & { return x.reg; }))) return rc;
  if ((rc = get(b.mem, memRows(d), d.memPoolSize, [](HostInst &x) -> std::vector<uint8_t> & { return x.mem; }))) return rc;
  if ((rc = get(b.exg, exgRows(d), d.exchangePoolSize, [](HostInst &x) -> std::vector<uint8_t> & { return x.exg; }))) return rc;
  if (c->nProbe() &&
      (rc = get(b.probe, c->nProbe(), c->nProbe(), [](HostInst &x) -> std::vector<uint8_t> & { return x.probe; })))
    return rc;
  std::vector<uint64_t> xs(P), st(P), sc(P);
  HIPCHK(c, hipMemcpy(xs.data(), b.xs.p, 8 * P, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(st.data(), b.stop.p, 8 * P, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(sc.data(), b.stopCycle.p, 8 * P, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < b.n; ++i) {
    h[i].xs = xs[b.slotOf[i]];
    h[i].stop = uint32_t(st[b.slotOf[i]]);
    h[i].stopCycle = sc[b.slotOf[i]];
    h[i].cycles = b.cycles[i];
  }
  return TGS_OK;
}

int downloadAll(tgs_ctx *c, std::vector<HostInst> &h) { return c->batched ? batchDownload(c, h) : instDownload(c, h); }
int uploadAll(tgs_ctx *c, const std::vector<HostInst> &h) { return c->batched ? batchUpload(c, h) : instUpload(c, h); }
uint32_t instanceCount(const tgs_ctx *c) { return c->batched ? c->bs.n : uint32_t(c->inst.size()); }

// Build and upload the batched kernel's program image and choose its launch
// shape: instances per wave I (a wave evaluates 64 / I ops per instruction
// round for I instances) and waves per workgroup. Each wave holds its group's
// value pool and level buffers in LDS ([slot][I], [2][2][32][I]); the LUT and
// the staged partition program are shared by the workgroup. The kernel is
// bound by the latency of each wave's dependent LDS chain, so its throughput
// follows the instances in flight per CU: co-resident workgroups (their
// waves at most 16 per CU) times the instances one item carries on average
// (an item is one partition for `waves` consecutive groups of I; this
// context's instances fill ceil(groups / waves) items), weighted for the
// longer LUT rounds of wide groups (a T_LUTFAST round is 256 / I ops: 0.8
// at I = 32, 0.5 at I = 64, as measured against I = 8 and 16). The shape
// with the most of them runs; ties go to more waves per CU, then to the
// smaller I. The LUT and the program buffers are paid per workgroup, so one
// large workgroup per CU can carry more instances than two small ones
// (1,024 instances of the 200K-LUT4 design: 12 waves of I = 16 against two
// workgroups of 8 waves of I = 8, 310 against 377 ms on one box).
int ensureBatchImage(tgs_ctx *c) {
  if (c->batchReady) return TGS_OK;
  const uint32_t imageWaves = c->batchWaves ? c->batchWaves : 4;
  uint32_t maxVp = 16;
  for (const auto &p : c->design.parts) maxVp = std::max(maxVp, p.valuePoolSize);
  const uint32_t vpRows = up16(maxVp);
  if (vpRows + kLevelRows > 8192)
    return c->fail(TGS_E_LIMIT, "batched execution addresses value pools of at most %u slots (%u here)",
                   8192 - kLevelRows, maxVp);
  c->w64 = buildWave64Image(c->design, c->stim, imageWaves, vpRows);
  c->w64Ready = false;  // the wave64 schedule must be rebuilt for this image
  c->batchFast = false;
  const uint32_t vpBytes = up16(std::max<uint32_t>(c->w64.maxVp, 16));
  auto maxWords = [](const Wave64Image &im) -> uint32_t {
    uint32_t m = 0;
    for (auto w : im.partWords) m = std::max(m, (w + 3) & ~3u);
    return m;
  };
  uint32_t maxPart = maxWords(c->w64);
  // T_LUTFAST image, when every value is a nibble: its tasks' word counts do
  // not depend on I (levels are unpadded), so one build sizes the program
  // buffers, and every shape below is sized for the larger of the two images
  const char *envFast = std::getenv("TGS_BATCH_FAST");
  const bool fastAllowed = !(envFast && envFast[0] == '0');
  Wave64Image fastImg;
  if (fastAllowed && c->nibbleState && (vpRows + kLevelRows) * 8u <= 65536u && nibbleProgram(c->design)) {
    fastImg = buildWave64Image(c->design, c->stim, imageWaves, vpRows, 8);
    if (fastImg.fastTasks) maxPart = std::max(maxPart, maxWords(fastImg));
  }
  const uint32_t ldsMax = uint32_t(c->prop.sharedMemPerBlock) - kLdsGranuleMargin;
  HIPCHK(c, batchCheckNoStaticLds());
  const char *envInst = std::getenv("TGS_BATCH_INST");
  const uint32_t forcedInst = c->batchInst ? c->batchInst : envInst ? uint32_t(std::atoi(envInst)) : 0u;
  std::vector<uint32_t> insts, waveList;
  for (uint32_t i : {8u, 16u, 32u, 64u})
    if (!forcedInst || forcedInst == i) insts.push_back(i);
  if (insts.empty()) return c->fail(TGS_E_ARG, "instances per wave must be 8, 16, 32 or 64 (%u given)", forcedInst);
  if (c->batchWaves) waveList.push_back(c->batchWaves);
  else waveList = {2u, 4u, 8u, 12u, 16u};
  // nibble rows (a nibble per instance) when every LUT task can be T_LUTFAST
  const char *envNib = std::getenv("TGS_BATCH_NIB");
  const bool nibOk = fastImg.fastTasks && fastImg.rowTasks == 0 && !(envNib && envNib[0] == '0');
  const uint64_t nInst = std::max<uint32_t>(c->bs.n, 1);
  c->batchShapeN = uint32_t(nInst);
  bool have = false, bestNib = false;
  uint32_t bestScore = 0, bestInst = 0, bestWaves = 0, bestLds = 0;  // bestScore: waves per CU
  double bestCols = 0;
  int bestPerCU = 0;
  for (uint32_t inst : insts)
    for (bool nib : {false, true}) {
      if (nib && !nibOk) continue;
      const uint32_t rowBytes = nib ? inst / 2 : inst;
      if (nib && (vpRows + kLevelRows) * rowBytes > 65536u) continue;
      for (uint32_t waves : waveList) {
        const uint32_t lds = batchLdsBytes(inst, waves, vpBytes, maxPart, nib);
        if (lds > ldsMax) continue;
        HIPCHK(c, batchSetLds(lds));
        int perCU = 0;
        HIPCHK(c, batchMaxBlocksPerCU(&perCU, inst, nib, waves * 64, lds));
        perCU = residentPerCU(c, perCU, lds);
        if (perCU < 1) continue;
        perCU = std::min<int>(perCU, std::max<uint32_t>(1, 16 / waves));
        const uint32_t score = uint32_t(perCU) * waves;
        const uint64_t groups = (nInst + inst - 1) / inst, items = (groups + waves - 1) / waves;
        // a T_LUTFAST round is 256 / rowBytes ops
        const double roundWeight = rowBytes >= 64 ? 0.5 : rowBytes >= 32 ? 0.8 : rowBytes <= 4 ? 0.7 : 1.0;
        const double cols = double(perCU) * double(std::min<uint64_t>(nInst, items * waves * inst)) / double(items) *
                            roundWeight;
        const bool same = std::abs(cols - bestCols) <= 1e-6 * cols;
        if (!have || (!same && cols > bestCols) ||
            (same && (score > bestScore || (score == bestScore && inst < bestInst)))) {
          have = true;
          bestCols = cols;
          bestScore = score;
          bestInst = inst;
          bestNib = nib;
          bestWaves = waves;
          bestLds = lds;
          bestPerCU = perCU;
        }
      }
    }
  if (!have)
    return c->fail(TGS_E_LIMIT,
                   "batched execution needs %u B of LDS (value pool %u slots x %u instances); device allows %u B",
                   batchLdsBytes(insts[0], 1, vpBytes, maxPart, false), c->w64.maxVp, insts[0], ldsMax);
  // T_LUTFAST tasks for the chosen rows (offsets row * bytes per row) when
  // those offsets fit 16 bits (the shape was sized for this image's program)
  const uint32_t bestRowBytes = bestNib ? bestInst / 2 : bestInst;
  c->batchNib = false;
  if (fastImg.fastTasks && (vpRows + kLevelRows) * bestRowBytes <= 65536u) {
    Wave64Image f = bestRowBytes == 8 ? std::move(fastImg)
                                      : buildWave64Image(c->design, c->stim, imageWaves, vpRows, bestRowBytes);
    if (f.fastTasks && maxWords(f) <= maxPart && (!bestNib || f.rowTasks == 0)) {
      c->w64 = std::move(f);
      c->batchFast = true;
      c->batchNib = bestNib;
    }
  }
  if (bestNib && !c->batchNib)
    return c->fail(TGS_E_INVALID, "nibble rows chosen for an image that is not all T_LUTFAST");
  {
    auto &im = c->w64;
    int rc;
    if ((rc = upload(c, c->wWords, im.words.data(), im.words.size() * 4))) return rc;
    if ((rc = upload(c, c->wPartOff, im.partOff.data(), im.partOff.size() * 4))) return rc;
    if ((rc = upload(c, c->wPartWords, im.partWords.data(), im.partWords.size() * 4))) return rc;
    if ((rc = upload(c, c->wConsts, im.consts.data(), im.consts.size()))) return rc;
  }
  c->batchInstEff = bestInst;
  c->batchProgWords = maxPart;
  c->batchBlockEff = bestWaves * 64;
  c->batchLds = bestLds;
  HIPCHK(c, batchSetLds(c->batchLds));
  // at most 16 waves per CU: more only lengthens the grid barrier
  const int perCU = std::min<int>(bestPerCU, std::max<uint32_t>(1, 16 / bestWaves));
  c->batchGrid = uint32_t(perCU) * c->prop.multiProcessorCount;
  if (c->batchGridOverride) c->batchGrid = std::min(c->batchGrid, c->batchGridOverride);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->batchReady = true;
  return TGS_OK;
}

// Reorder slots so that slot order is decreasing remaining cycle count.
int batchOrder(tgs_ctx *c, const uint32_t *cycles) {
  auto &b = c->bs;
  const auto &d = c->design;
  const uint32_t n = b.n, P = b.instPad;
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return cycles[x] > cycles[y]; });
  std::vector<uint32_t> newSlot(n);
  bool same = true;
  for (uint32_t k = 0; k < n; ++k) {
    newSlot[order[k]] = k;
    same = same && b.slotOf[order[k]] == k;
  }
  if (same) return TGS_OK;
  std::vector<uint32_t> src(P);
  for (uint32_t k = n; k < P; ++k) src[k] = k;
  for (uint32_t i = 0; i < n; ++i) src[newSlot[i]] = b.slotOf[i];
  HIPCHK(c, hipMemcpyAsync(b.perm.p, src.data(), 4 * P, hipMemcpyHostToDevice, c->stream));
  const uint32_t *sp = b.perm.as<uint32_t>();
  auto perm8 = [&](DBuf &buf, uint64_t rows) -> int {
    HIPCHK(c, launchPermute(buf.as<uint8_t>(), b.scratch.as<uint8_t>(), rows, P, sp, c->stream));
    HIPCHK(c, hipMemcpyAsync(buf.p, b.scratch.p, rows * P, hipMemcpyDeviceToDevice, c->stream));
    return TGS_OK;
  };
  auto perm64 = [&](DBuf &buf) -> int {
    HIPCHK(c, launchPermute64(buf.as<uint64_t>(), b.scratch.as<uint64_t>(), 1, P, sp, c->stream));
    HIPCHK(c, hipMemcpyAsync(buf.p, b.scratch.p, 8 * P, hipMemcpyDeviceToDevice, c->stream));
    return TGS_OK;
  };
  int rc;
  if ((rc = perm8(b.reg, regRows(d))) || (rc = perm8(b.mem, memRows(d))) || (rc = perm8(b.exg, exgRows(d))) ||
      (rc = perm64(b.xs)) || (rc = perm64(b.stop)) || (rc = perm64(b.stopCycle)) ||
      (c->nProbe() && (rc = perm8(b.probe, c->nProbe()))))
    return rc;
  for (uint32_t i = 0; i < n; ++i) b.slotOf[i] = newSlot[i];
  std::fill(b.instOf.begin(), b.instOf.end(), 0xFFFFFFFFu);
  for (uint32_t i = 0; i < n; ++i) b.instOf[newSlot[i]] = i;
  HIPCHK(c, hipMemcpyAsync(b.instOfD.p, b.instOf.data(), 4 * P, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TGS_OK;
}

// Device memory a batched run needs beyond the instance state: per-slot
// cycle counts and bases of the launch, the per-launch-cycle active-group
// table and the grid-barrier counter. `fixedCycles` = 0: the context's own
// buffers, the table allocated per launch; otherwise preallocated memory
// whose table holds at most fixedCycles entries (launches are cut to fit).
struct BatchScratch {
  uint32_t *runCycles = nullptr;
  uint64_t *cycleBase = nullptr;
  uint32_t *groupsActive = nullptr;
  uint32_t *barrier = nullptr;
  BatchArgs *args = nullptr;  // the launch's BatchArgs (one context)
  uint64_t fixedCycles = 0;
};

// BatchArgs of context c for one launch: its image and pools, and the
// launch's per-slot cycle counts and bases and active-group table.
BatchArgs batchArgsOf(tgs_ctx *c, const BatchScratch &sc, uint32_t nCycles) {
  auto &b = c->bs;
  BatchArgs a{};
  a.lut = c->lut.as<uint8_t>();
  a.consts = c->wConsts.as<uint8_t>();
  a.reg = b.reg.as<uint8_t>();
  a.mem = b.mem.as<uint8_t>();
  a.exg = b.exg.as<uint8_t>();
  a.regSize = c->design.regPoolSize;
  a.memSize = c->design.memPoolSize;
  a.exgSize = c->design.exchangePoolSize;
  a.instPad = b.instPad;
  a.words = c->wWords.as<uint32_t>();
  a.partOff = c->wPartOff.as<uint32_t>();
  a.partWords = c->wPartWords.as<uint32_t>();
  a.n0 = c->w64.n0;
  a.nTotal = c->w64.nTotal;
  a.vpBytes = up16(std::max<uint32_t>(c->w64.maxVp, 16));
  a.runCycles = sc.runCycles;
  a.cycleBase = sc.cycleBase;
  a.groupsActive = sc.groupsActive;
  a.nCycles = nCycles;
  a.xs = b.xs.as<uint64_t>();
  a.draws = b.draws.as<uint64_t>();
  a.inIds = c->inIds.as<uint32_t>();
  a.nIn = c->stim.ids.size();
  a.stop = b.stop.as<uint64_t>();
  a.stopCycle = b.stopCycle.as<uint64_t>();
  a.logCount = b.logCount.as<uint64_t>();
  a.logCycle = b.logCycle.as<uint64_t>();
  a.logMsg = b.logMsg.as<uint64_t>();
  a.instOf = b.instOfD.as<uint32_t>();
  a.logCap = uint32_t(b.logCap);
  a.printEnable = c->printEnable;
  if (c->nProbe()) {
    a.probeOff = c->prOff.as<uint32_t>();
    a.probeSlot = c->prSlot.as<uint32_t>();
    a.probeVal = b.probe.as<uint8_t>();
  }
  return a;
}

// Launch shape of a batched launch over several contexts: their common
// instances per wave, the first context's waves per workgroup (fewer when
// the largest value pool and program do not fit), LDS sized for the largest.
struct BatchShape {
  uint32_t inst = 0, block = 0, grid = 0, lds = 0, vpRows = 0, progWords = 0;
  bool nib = false;
};

int batchShapeOf(tgs_ctx *const *cs, uint32_t nCtx, BatchShape *out) {
  tgs_ctx *c0 = cs[0];
  BatchShape sh;
  sh.inst = c0->batchInstEff;
  sh.nib = c0->batchNib;
  for (uint32_t k = 0; k < nCtx; ++k) {
    sh.vpRows = std::max(sh.vpRows, up16(std::max<uint32_t>(cs[k]->w64.maxVp, 16)));
    sh.progWords = std::max(sh.progWords, cs[k]->batchProgWords);
  }
  if (nCtx == 1) {
    sh.block = c0->batchBlockEff;
    sh.grid = c0->batchGrid;
    sh.lds = c0->batchLds;
    *out = sh;
    return TGS_OK;
  }
  const uint32_t ldsMax = uint32_t(c0->prop.sharedMemPerBlock) - kLdsGranuleMargin;
  for (uint32_t waves = c0->batchBlockEff / 64; waves >= 1; waves /= 2) {
    const uint32_t lds = batchLdsBytes(sh.inst, waves, sh.vpRows, sh.progWords, sh.nib);
    if (lds > ldsMax) continue;
    HIPCHK(c0, batchSetLds(lds));
    int perCU = 0;
    HIPCHK(c0, batchMaxBlocksPerCU(&perCU, sh.inst, sh.nib, waves * 64, lds));
    perCU = residentPerCU(c0, perCU, lds);
    if (perCU < 1) continue;
    perCU = std::min<int>(perCU, std::max<uint32_t>(1, 16 / waves));
    sh.block = waves * 64;
    sh.lds = lds;
    sh.grid = uint32_t(perCU) * c0->prop.multiProcessorCount;
    if (c0->batchGridOverride) sh.grid = std::min(sh.grid, c0->batchGridOverride);
    *out = sh;
    return TGS_OK;
  }
  return c0->fail(TGS_E_LIMIT, "the contexts' value pools and programs do not fit one workgroup's LDS together");
}

// Run the instances of nCtx batched contexts of one device in persistent
// launches that serve all of them: each launch cycle runs region 0 of every
// context, then region 1, so small batches of different designs share the
// GPU. cycles[k] holds one count per instance of context k. Launches go on
// the first context's stream after every context's stream has drained.
// `fixed`: preallocated scratch (one context only, see BatchScratch).
int runBatchedMulti(tgs_ctx *const *cs, const uint32_t *const *cycles, uint32_t nCtx,
                    const BatchScratch *fixed = nullptr) {
  int rc;
  tgs_ctx *c0 = cs[0];
  for (uint32_t k = 0; k < nCtx; ++k) {
    tgs_ctx *c = cs[k];
    if ((rc = ensureBatchImage(c))) return rc;
    if ((rc = batchOrder(c, cycles[k]))) return rc;
    if (c != c0) HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  BatchShape sh;
  if ((rc = batchShapeOf(cs, nCtx, &sh))) return rc;
  if (nCtx > 1) HIPCHK(c0, batchSetLds(sh.lds));
  // host bookkeeping kept in the contexts: no host allocation once sized
  uint64_t maxAll = 0, perLaunch = std::min<uint64_t>({1u << 20, 0xE0000000ull / (2ull * sh.grid + 2),
                                                        fixed ? fixed->fixedCycles : uint64_t(UINT32_MAX)});
  for (uint32_t k = 0; k < nCtx; ++k) {
    tgs_ctx *c = cs[k];
    auto &b = c->bs;
    const uint32_t P = b.instPad;
    auto &remaining = c->bhRemaining, &base = c->bhBase;
    remaining.assign(P, 0);
    base.assign(P, 0);
    for (uint32_t s = 0; s < P; ++s)
      if (b.instOf[s] != 0xFFFFFFFFu) {
        remaining[s] = cycles[k][b.instOf[s]];
        base[s] = b.cycles[b.instOf[s]];
        maxAll = std::max(maxAll, remaining[s]);
      }
    perLaunch = std::min<uint64_t>(perLaunch, printCycleCap(c, b.logCap, P));
    c->bhRun.assign(P, 0);
  }
  std::vector<BatchArgs> &args = c0->bhArgs;
  args.resize(nCtx);
  while (maxAll > 0) {
    const uint32_t L = uint32_t(std::min<uint64_t>(maxAll, perLaunch));
    const uint32_t I = sh.inst;
    uint32_t *barrier = nullptr;
    for (uint32_t k = 0; k < nCtx; ++k) {
      tgs_ctx *c = cs[k];
      auto &b = c->bs;
      const uint32_t P = b.instPad;
      if ((rc = drainLogs(c))) return rc;  // every launch starts with an empty device log
      auto &run = c->bhRun, &active = c->bhActive;
      uint32_t Lk = 0;
      for (uint32_t s = 0; s < P; ++s) {
        run[s] = uint32_t(std::min<uint64_t>(c->bhRemaining[s], L));
        Lk = std::max(Lk, run[s]);
      }
      active.assign(std::max<uint32_t>(Lk, 1), 0);
      for (uint32_t cy = 0; cy < Lk; ++cy) {
        uint32_t g = 0;
        while (g < P / I && run[g * I] > cy) ++g;  // slots sorted by decreasing cycles
        active[cy] = g;
      }
      BatchScratch sc;
      if (fixed) {
        sc = *fixed;
      } else {
        HIPCHK(c, b.groupsActive.alloc(4ull * active.size()));
        sc.runCycles = b.runCycles.as<uint32_t>();
        sc.cycleBase = b.cycleBase.as<uint64_t>();
        sc.groupsActive = b.groupsActive.as<uint32_t>();
        sc.barrier = b.barrier.as<uint32_t>();
      }
      if (k == 0) barrier = sc.barrier;
      HIPCHK(c, hipMemcpyAsync(sc.runCycles, run.data(), 4ull * P, hipMemcpyHostToDevice, c0->stream));
      HIPCHK(c, hipMemcpyAsync(sc.cycleBase, c->bhBase.data(), 8ull * P, hipMemcpyHostToDevice, c0->stream));
      HIPCHK(c, hipMemcpyAsync(sc.groupsActive, active.data(), 4ull * active.size(), hipMemcpyHostToDevice,
                               c0->stream));
      args[k] = batchArgsOf(c, sc, Lk);
    }
    BatchArgs *argsDev;
    if (fixed) {
      argsDev = fixed->args;
    } else {
      HIPCHK(c0, c0->bs.argsDev.alloc(sizeof(BatchArgs) * nCtx));
      argsDev = c0->bs.argsDev.as<BatchArgs>();
    }
    HIPCHK(c0, hipMemcpyAsync(argsDev, args.data(), sizeof(BatchArgs) * nCtx, hipMemcpyHostToDevice, c0->stream));
    HIPCHK(c0, launchFill32(barrier, kBarTreeWords, 0, c0->stream));  // a kernel, not a memset node
    BatchLaunch bl{argsDev, nCtx, L, barrier, sh.vpRows, sh.progWords};
    HIPCHK(c0, hipEventRecord(c0->ev0, c0->stream));
    HIPCHK(c0, launchBatch(bl, sh.inst, sh.nib, sh.grid, sh.block, sh.lds, c0->stream));
    HIPCHK(c0, hipEventRecord(c0->ev1, c0->stream));
    HIPCHK(c0, hipEventSynchronize(c0->ev1));
    float ms = 0;
    HIPCHK(c0, hipEventElapsedTime(&ms, c0->ev0, c0->ev1));
    maxAll = 0;
    for (uint32_t k = 0; k < nCtx; ++k) {
      tgs_ctx *c = cs[k];
      const uint32_t P = c->bs.instPad;
      c->prof.kernel_ms += ms;  // the shared launch, counted for every context it served
      c->prof.launches += 1;
      c->prof.kernel = TGS_KERNEL_BATCH;
      if (c->printEnable && c->stats.n_print) c->logDirty = true;
      c->prof.grid = sh.grid;
      c->prof.block = sh.block;
      c->prof.batch_group = sh.inst;
      c->prof.batch_fast_tasks = c->batchFast ? c->w64.fastTasks : 0;
      c->prof.batch_nibble_rows = c->batchNib ? 1 : 0;
      c->prof.lds_bytes = sh.lds;
      c->prof.prog_bytes = c->w64.words.size() * 4;
      for (uint32_t s = 0; s < P; ++s) {
        c->prof.cycles += c->bhRun[s];
        c->bhRemaining[s] -= c->bhRun[s];
        c->bhBase[s] += c->bhRun[s];
        maxAll = std::max(maxAll, c->bhRemaining[s]);
      }
    }
  }
  for (uint32_t k = 0; k < nCtx; ++k)
    for (uint32_t i = 0; i < cs[k]->bs.n; ++i) cs[k]->bs.cycles[i] += cycles[k][i];
  return TGS_OK;
}

int runBatched(tgs_ctx *c, const uint32_t *cycles, const BatchScratch *fixed = nullptr) {
  return runBatchedMulti(&c, &cycles, 1, fixed);
}

int switchLayout(tgs_ctx *c, bool batched) {
  if (c->batched == batched) return TGS_OK;
  std::vector<HostInst> h;
  int rc = drainLogs(c);
  if (rc) return rc;
  rc = downloadAll(c, h);
  if (rc) return rc;
  c->batched = batched;
  if (batched) c->inst.clear();
  return uploadAll(c, h);
}

int checkLoaded(tgs_ctx *c) {
  if (!c) return TGS_E_ARG;
  if (!c->loaded) return c->fail(TGS_E_STATE, "no design loaded");
  return setDevice(c);
}

int readStop(tgs_ctx *c, Instance &in, int *stopped) {
  uint32_t s = 0;
  HIPCHK(c, hipMemcpy(&s, in.ctl.as<uint32_t>() + C_STOP, 4, hipMemcpyDeviceToHost));
  *stopped = s != 0;
  return TGS_OK;
}


}  // namespace

extern "C" {

int tgs_program_stats_from_bytes(const uint8_t *design, size_t size, tgs_program_stats *out, char *err,
                                 size_t errlen) {
  if (!design || !out) return TGS_E_ARG;
  SimDesignInfo d;
  std::string e;
  if (!toucanGPUSim::parseSimDesignInfo(design, size, d, &e)) {
    if (err && errlen) snprintf(err, errlen, "%s", e.c_str());
    return TGS_E_FORMAT;
  }
  int rc = validateProgram(d, &e);
  if (rc != TGS_OK) {
    if (err && errlen) snprintf(err, errlen, "%s", e.c_str());
    return rc;
  }
  programStats(d, 0, out);
  return TGS_OK;
}

int tgs_initial_state(const uint8_t *design, size_t designSize, const uint8_t *symbols, size_t symbolsSize,
                      int randomize, uint32_t seed, uint8_t *reg, size_t regSize, uint8_t *mem, size_t memSize) {
  if (!design) return TGS_E_ARG;
  SimDesignInfo d;
  SimDebugInfo s;
  if (!toucanGPUSim::parseSimDesignInfo(design, designSize, d, nullptr)) return TGS_E_FORMAT;
  if (symbols && !toucanGPUSim::parseSimDebugInfo(symbols, symbolsSize, s, nullptr)) return TGS_E_FORMAT;
  if (randomize && !symbols) return TGS_E_ARG;
  d.Init();
  if (randomize) d.Randomize(seed, s);
  if (regSize != d.regPool.size() || memSize != d.memPool.size()) return TGS_E_ARG;
  if (regSize) std::memcpy(reg, d.regPool.data(), regSize);
  if (memSize) std::memcpy(mem, d.memPool.data(), memSize);
  return TGS_OK;
}

// This device's address of `scratch`, resolved once per device with the
// device current (the array exists on every GPU the library runs on).
static hipError_t scratchFor(int device, void **out) {
  static std::mutex mu;
  static std::vector<void *> addr;
  std::lock_guard<std::mutex> lock(mu);
  if (size_t(device) >= addr.size()) addr.resize(size_t(device) + 1, nullptr);
  if (!addr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *out = addr[device];
  return hipSuccess;
}

static int createCtx(int device, hipStream_t userStream, tgs_ctx **out) {
  if (!out) return TGS_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return TGS_E_NODEV;
  }
  if (device < 0 || device >= n) return TGS_E_NODEV;
  auto c = std::make_unique<tgs_ctx>();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) return TGS_E_NODEV;
  if (hipGetDeviceProperties(&c->prop, device) != hipSuccess) return TGS_E_NODEV;
  // setupGPU (ToucanSim.cu:40-43): the cycle kernels need cooperative launch.
  if (!c->prop.cooperativeLaunch) return TGS_E_NODEV;
  if (userStream) {
    c->stream = userStream;
    c->ownStream = false;
  } else if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    c->stream = nullptr;
    return TGS_E_HIP;
  }
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) return TGS_E_HIP;
  if (scratchFor(device, &c->scratchDev) != hipSuccess) {
    (void)hipGetLastError();
    return TGS_E_HIP;
  }
  const char *blk = std::getenv("TGS_W64_BLOCK");
  if (blk) c->w64Block = std::max(64u, std::min(1024u, uint32_t(std::atoi(blk)) / 64 * 64));
  const char *grp = std::getenv("TGS_W64_GROUPS");
  if (grp) c->w64Groups = uint32_t(std::atoi(grp));
  const char *mode = std::getenv("TGS_W64_MODE");
  if (mode) c->w64ModeForce = std::atoi(mode);
  *out = c.release();
  return TGS_OK;
}

int tgs_create(int device, tgs_ctx **out) { return createCtx(device, nullptr, out); }

int tgs_create_on_stream(int device, void *stream, tgs_ctx **out) {
  if (!stream) return TGS_E_ARG;
  return createCtx(device, static_cast<hipStream_t>(stream), out);
}

void tgs_destroy(tgs_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->snap.release();
  c->box.reset();
  c->inst.clear();
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream && c->ownStream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int tgs_last_error(const tgs_ctx *c, char *buf, size_t len) {
  if (!c || !buf || !len) return TGS_E_ARG;
  snprintf(buf, len, "%s", c->err.c_str());
  return TGS_OK;
}

int tgs_load(tgs_ctx *c, const uint8_t *design, size_t designSize, const uint8_t *symbols, size_t symbolsSize,
             int randomize, int enablePrint) {
  if (!c || !design) return TGS_E_ARG;
  int rc = setDevice(c);
  if (rc) return rc;
  c->loaded = false;
  c->plainReady = c->w64Ready = c->streamReady = false;
  c->probeCsrOf.clear();
  c->inst.clear();
  c->design = SimDesignInfo();
  c->symbols = SimDebugInfo();
  std::string e;
  if (!toucanGPUSim::parseSimDesignInfo(design, designSize, c->design, &e))
    return c->fail(TGS_E_FORMAT, "design: %s", e.c_str());
  if (symbols && !toucanGPUSim::parseSimDebugInfo(symbols, symbolsSize, c->symbols, &e))
    return c->fail(TGS_E_FORMAT, "symbols: %s", e.c_str());
  if (randomize && !symbols) return c->fail(TGS_E_ARG, "Randomize needs the symbol file");
  rc = validateProgram(c->design, &e);
  if (rc != TGS_OK) return c->fail(rc, "%s", e.c_str());
  // init (ToucanSim.cu:128-134): Init, then Randomize(0) when asked.
  c->design.Init();
  if (randomize) c->design.Randomize(0, c->symbols);
  c->initReg = c->design.regPool;
  c->initMem = c->design.memPool;
  c->nibbleState = true;
  c->batchFast = false;
  c->noteWrite(c->initReg.data(), c->initReg.size());
  c->noteWrite(c->initMem.data(), c->initMem.size());
  c->designHash = 0xCBF29CE484222325ull;
  for (size_t i = 0; i < designSize; ++i) c->designHash = (c->designHash ^ design[i]) * 0x100000001B3ull;
  c->stim = StimMap();
  c->printEnable = enablePrint ? 1 : 0;
  programStats(c->design, 0, &c->stats);
  std::vector<uint8_t> lut(wave64LutBytes(), 0);
  std::copy(c->design.lut.begin(), c->design.lut.end(), lut.begin());
  if ((rc = upload(c, c->lut, lut.data(), lut.size()))) return rc;
  if ((rc = upload(c, c->inIds, nullptr, 0))) return rc;
  c->batched = false;
  c->batchReady = false;
  if (!c->kernelChosen) c->kernel = TGS_KERNEL_STREAM;
  c->kernel = c->kernel == TGS_KERNEL_BATCH ? TGS_KERNEL_STREAM : c->kernel;
  c->inst.emplace_back(new Instance());
  if ((rc = resetInstance(c, *c->inst[0], 0, c->stimSeed))) return rc;
  c->logs.assign(1, HostLog());
  c->logDirty = false;
  c->partProf = false;
  c->profCollected = 0;
  c->loaded = true;
  c->prof = tgs_profile{};
  return TGS_OK;
}

int tgs_load_files(tgs_ctx *c, const char *designPath, const char *symbolPath, int randomize, int enablePrint) {
  if (!c || !designPath) return TGS_E_ARG;
  auto slurp = [](const char *path, std::vector<uint8_t> &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
  };
  std::vector<uint8_t> d, s;
  if (!slurp(designPath, d)) return c->fail(TGS_E_IO, "Error opening file: %s", designPath);
  if (symbolPath && !slurp(symbolPath, s)) return c->fail(TGS_E_IO, "Error opening file: %s", symbolPath);
  return tgs_load(c, d.data(), d.size(), symbolPath ? s.data() : nullptr, s.size(), randomize, enablePrint);
}

int tgs_set_kernel(tgs_ctx *c, int kernel) {
  if (!c) return TGS_E_ARG;
  if (kernel != TGS_KERNEL_PLAIN && kernel != TGS_KERNEL_WAVE64 && kernel != TGS_KERNEL_BATCH &&
      kernel != TGS_KERNEL_STREAM)
    return c->fail(TGS_E_ARG, "unknown kernel %d", kernel);
  if (c->loaded) {
    int rc = setDevice(c);
    if (rc) return rc;
    if ((rc = switchLayout(c, kernel == TGS_KERNEL_BATCH))) return rc;
    if (kernel == TGS_KERNEL_BATCH && (rc = ensureBatchImage(c))) return rc;
  }
  c->kernel = kernel;
  c->kernelChosen = true;
  return TGS_OK;
}

int tgs_set_launch(tgs_ctx *c, int kernel, uint32_t block, uint32_t groups) {
  if (!c) return TGS_E_ARG;
  if (block % 64 || block > 1024) return c->fail(TGS_E_ARG, "block must be a multiple of 64 up to 1024");
  switch (kernel) {
    case TGS_KERNEL_PLAIN:
      if (block) c->plainBlock = block;
      c->plainReady = false;
      break;
    case TGS_KERNEL_WAVE64:
      if (block) c->w64Block = block;
      c->w64Groups = groups;
      c->w64Ready = false;
      break;
    case TGS_KERNEL_BATCH:
      if (block > 1024) return c->fail(TGS_E_ARG, "batched kernel block must be at most 1024 (sixteen waves)");
      c->batchWaves = block / 64;
      c->batchGridOverride = groups;
      c->batchReady = false;
      if (c->loaded && c->batched) {  // the batched image is rebuilt now, not at the next run
        int rc = setDevice(c);
        if (rc || (rc = ensureBatchImage(c))) return rc;
      }
      break;
    case TGS_KERNEL_STREAM:
      c->streamNC = 0;  // automatic
      for (uint32_t nc : {2u, 4u, 8u})
        if (block && streamBlock(nc) == block) c->streamNC = nc;
      if (block && !c->streamNC)
        return c->fail(TGS_E_ARG, "streamed kernel block must be %u, %u or %u (2, 4 or 8 compute waves + loaders)",
                       streamBlock(2), streamBlock(4), streamBlock(8));
      c->streamGroups = groups;
      c->streamReady = false;
      break;
    default:
      return c->fail(TGS_E_ARG, "unknown kernel %d", kernel);
  }
  return TGS_OK;
}

int tgs_set_print(tgs_ctx *c, int enable) {
  if (!c) return TGS_E_ARG;
  c->printEnable = enable ? 1 : 0;
  return TGS_OK;
}

int tgs_set_instances(tgs_ctx *c, uint32_t n) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (n == 0) return c->fail(TGS_E_ARG, "at least one instance");
  std::vector<HostInst> h;
  if ((rc = initialHost(c, n, h))) return rc;
  if ((rc = uploadAll(c, h))) return rc;
  c->logs.assign(n, HostLog());
  c->logDirty = false;
  if (c->batched && c->batchShapeN != n) c->batchReady = false;  // the launch shape depends on the count
  return c->batched ? ensureBatchImage(c) : TGS_OK;
}

int tgs_set_stimulus(tgs_ctx *c, const uint32_t *ids, uint32_t n, uint64_t seed) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (n && !ids) return TGS_E_ARG;
  if (n > kMaxDraws * 64) return c->fail(TGS_E_LIMIT, "at most %u stimulus inputs", kMaxDraws * 64);
  std::string e;
  StimMap m;
  if ((rc = buildStimMap(c->design, ids, n, &m, &e))) return c->fail(rc, "%s", e.c_str());
  c->stim = std::move(m);
  c->stimSeed = seed;
  c->plainReady = c->w64Ready = c->batchReady = c->streamReady = false;
  if ((rc = upload(c, c->inIds, c->stim.ids.data(), c->stim.ids.size() * 4))) return rc;
  programStats(c->design, n, &c->stats);
  // every instance restarts its generator from splitmix64(seed + k)
  if (c->batched) {
    std::vector<uint64_t> xs(c->bs.instPad, 0);
    for (uint32_t i = 0; i < c->bs.n; ++i) xs[c->bs.slotOf[i]] = seedFor(seed, i);
    HIPCHK(c, hipMemcpy(c->bs.xs.p, xs.data(), 8 * xs.size(), hipMemcpyHostToDevice));
  } else {
    for (uint32_t i = 0; i < c->inst.size(); ++i) {
      uint64_t z = seedFor(seed, i);
      HIPCHK(c, hipMemcpy(c->inst[i]->xs.p, &z, 8, hipMemcpyHostToDevice));
    }
  }
  return c->batched ? ensureBatchImage(c) : TGS_OK;
}

int tgs_run_batch(tgs_ctx *c, const uint32_t *cycles, uint32_t n, int *anyStopped) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (!cycles || n != instanceCount(c)) return c->fail(TGS_E_ARG, "need one cycle count per instance");
  int any = 0;
  if (c->batched) {
    if ((rc = runBatched(c, cycles))) return rc;
    std::vector<uint64_t> st(c->bs.instPad);
    HIPCHK(c, hipMemcpy(st.data(), c->bs.stop.p, 8 * st.size(), hipMemcpyDeviceToHost));
    for (auto v : st) any |= v != 0;
  } else {
    for (uint32_t i = 0; i < n; ++i) {
      if ((rc = runInstance(c, *c->inst[i], cycles[i]))) return rc;
      int s = 0;
      if ((rc = readStop(c, *c->inst[i], &s))) return rc;
      any |= s;
    }
  }
  if (anyStopped) *anyStopped = any;
  return TGS_OK;
}

// Device memory the context holds: every buffer of its instances, images,
// batched state, probes, profile and testbench mailbox, and the waveform
// staging ring.
uint64_t deviceBytes(const tgs_ctx *c) {
  uint64_t t = 0;
  for (const DBuf *b : {&c->lut, &c->inIds, &c->pNet, &c->pParts, &c->pLevels, &c->pMparts, &c->pConsts, &c->wWords,
                        &c->wPartOff, &c->wPartWords, &c->wSchedOff, &c->wSchedPart, &c->wSchedLds, &c->wConsts,
                        &c->sWords, &c->sSteps, &c->sStepBase, &c->sStepCount, &c->sConsts, &c->profUseful,
                        &c->profWg, &c->profOwner, &c->boxDev, &c->prOff, &c->prSlot})
    t += b->p ? b->n : 0;
  const auto &b = c->bs;
  for (const DBuf *d : {&b.reg, &b.mem, &b.exg, &b.xs, &b.draws, &b.stop, &b.stopCycle, &b.logCount, &b.logCycle,
                        &b.logMsg, &b.barrier, &b.probe, &b.runCycles, &b.cycleBase, &b.groupsActive, &b.instOfD,
                        &b.perm, &b.scratch, &b.argsDev})
    t += d->p ? d->n : 0;
  for (const auto &in : c->inst)
    for (const DBuf *d : {&in->reg, &in->mem, &in->exg, &in->xs, &in->ctl, &in->logCycle, &in->logMsg, &in->probe})
      t += d->p ? d->n : 0;
  if (c->snap.staging) t += uint64_t(c->snap.slots) * c->snap.bytes;
  return t;
}

int tgs_device_bytes(tgs_ctx *c, uint64_t *bytes) {
  if (!c || !bytes) return TGS_E_ARG;
  *bytes = deviceBytes(c);
  return TGS_OK;
}

int tgs_run_batch_multi(tgs_ctx *const *ctxs, uint32_t n, const uint32_t *const *cycles, int *anyStopped) {
  if (!ctxs || !cycles || n == 0) return TGS_E_ARG;
  for (uint32_t k = 0; k < n; ++k) {
    tgs_ctx *c = ctxs[k];
    if (!c) return TGS_E_ARG;
    int rc = checkLoaded(c);
    if (rc) return rc;
    if (!c->batched) return c->fail(TGS_E_STATE, "tgs_run_batch_multi runs batched contexts (tgs_set_kernel batched)");
    if (!cycles[k]) return c->fail(TGS_E_ARG, "need one cycle count per instance");
    if (c->device != ctxs[0]->device) return c->fail(TGS_E_ARG, "the contexts of one call share one device");
    for (uint32_t j = 0; j < k; ++j)
      if (ctxs[j] == c) return c->fail(TGS_E_ARG, "a context appears twice in one call");
  }
  if (n > kMaxBatchContexts)
    return ctxs[0]->fail(TGS_E_LIMIT, "at most %u contexts per call (%u given)", kMaxBatchContexts, n);
  int rc;
  for (uint32_t k = 0; k < n; ++k)
    if ((rc = setDevice(ctxs[k])) || (rc = ensureBatchImage(ctxs[k]))) return rc;
  // one launch per instance count per wave and row format (the kernel is
  // specialised on both)
  std::vector<tgs_ctx *> grp;
  std::vector<const uint32_t *> cyc;
  for (uint32_t inst : {8u, 16u, 32u, 64u})
    for (bool nib : {false, true}) {
      grp.clear();
      cyc.clear();
      for (uint32_t k = 0; k < n; ++k)
        if (ctxs[k]->batchInstEff == inst && ctxs[k]->batchNib == nib) {
          grp.push_back(ctxs[k]);
          cyc.push_back(cycles[k]);
        }
      if (!grp.empty() && (rc = runBatchedMulti(grp.data(), cyc.data(), uint32_t(grp.size())))) return rc;
    }
  for (uint32_t k = 0; k < n; ++k) {
    tgs_ctx *c = ctxs[k];
    auto &st = c->bhStop;
    st.resize(c->bs.instPad);
    HIPCHK(c, hipMemcpy(st.data(), c->bs.stop.p, 8 * st.size(), hipMemcpyDeviceToHost));
    int any = 0;
    for (auto v : st) any |= v != 0;
    if (anyStopped) anyStopped[k] = any;
  }
  return TGS_OK;
}

int tgs_run_batch_noalloc(tgs_ctx *c, const uint32_t *cycles, uint32_t n, int *anyStopped) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (!c->batched) return c->fail(TGS_E_STATE, "tgs_run_batch_noalloc runs batched contexts (tgs_set_kernel batched)");
  if (!cycles || n != instanceCount(c)) return c->fail(TGS_E_ARG, "need one cycle count per instance");
  if (c->bs.instPad > kScratchSlots)
    return c->fail(TGS_E_LIMIT, "tgs_run_batch_noalloc holds at most %u instances (%u given)", kScratchSlots, n);
  if (!c->batchReady)
    return c->fail(TGS_E_STATE, "the batched program image is not built (tgs_set_instances or tgs_set_stimulus "
                                "after the last tgs_set_launch builds it)");
  uint8_t *base = static_cast<uint8_t *>(c->scratchDev);
  BatchScratch sc;
  sc.runCycles = reinterpret_cast<uint32_t *>(base + kScratchRunOff);
  sc.cycleBase = reinterpret_cast<uint64_t *>(base + kScratchBaseOff);
  sc.barrier = reinterpret_cast<uint32_t *>(base + kScratchBarrierOff);
  sc.groupsActive = reinterpret_cast<uint32_t *>(base + kScratchActiveOff);
  sc.args = reinterpret_cast<BatchArgs *>(base + kScratchArgsOff);
  sc.fixedCycles = kScratchCycles;
  if ((rc = runBatched(c, cycles, &sc))) return rc;
  auto &st = c->bhStop;
  st.resize(c->bs.instPad);
  HIPCHK(c, hipMemcpy(st.data(), c->bs.stop.p, 8 * st.size(), hipMemcpyDeviceToHost));
  int any = 0;
  for (auto v : st) any |= v != 0;
  if (anyStopped) *anyStopped = any;
  return TGS_OK;
}

int tgs_get_stop(tgs_ctx *c, uint32_t instance, int *stopped);

int tgs_eval_free_running(tgs_ctx *c, uint32_t cycles, int *stopped) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  std::vector<uint32_t> all(instanceCount(c), cycles);
  if ((rc = tgs_run_batch(c, all.data(), uint32_t(all.size()), nullptr))) return rc;
  if (stopped) return tgs_get_stop(c, 0, stopped);
  return TGS_OK;
}

int tgs_eval(tgs_ctx *c, int *stopped) { return tgs_eval_free_running(c, 1, stopped); }

int tgs_pool_size(const tgs_ctx *c, int pool, uint64_t *size) {
  if (!c || !size || !c->loaded) return TGS_E_ARG;
  switch (pool) {
    case TGS_POOL_REG: *size = c->design.regPoolSize; return TGS_OK;
    case TGS_POOL_MEM: *size = c->design.memPoolSize; return TGS_OK;
    case TGS_POOL_EXCHANGE: *size = c->design.exchangePoolSize; return TGS_OK;
    default: return TGS_E_ARG;
  }
}

// Copy [off, off + n) of one instance's pool to or from host memory, in either layout.
static int poolIO(tgs_ctx *c, uint32_t instance, int pool, uint64_t off, uint64_t n, uint8_t *host, bool write) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (instance >= instanceCount(c)) return c->fail(TGS_E_ARG, "instance %u out of range", instance);
  uint64_t size = 0;
  if (tgs_pool_size(c, pool, &size)) return c->fail(TGS_E_ARG, "unknown pool %d", pool);
  if (off > size || n > size - off) return c->fail(TGS_E_ARG, "pool range out of bounds");
  if (n == 0) return TGS_OK;
  if (!host) return TGS_E_ARG;
  if (write) c->noteWrite(host, n);
  if (!c->batched) {
    Instance *in = c->inst[instance].get();
    DBuf &b = pool == TGS_POOL_REG ? in->reg : pool == TGS_POOL_MEM ? in->mem : in->exg;
    uint8_t *p = b.as<uint8_t>() + off;
    HIPCHK(c, write ? hipMemcpy(p, host, n, hipMemcpyHostToDevice) : hipMemcpy(host, p, n, hipMemcpyDeviceToHost));
    return TGS_OK;
  }
  auto &bs = c->bs;
  DBuf &b = pool == TGS_POOL_REG ? bs.reg : pool == TGS_POOL_MEM ? bs.mem : bs.exg;
  const size_t P = bs.instPad;
  // gather / scatter the instance's column through the dense staging buffer
  uint8_t *p = b.as<uint8_t>() + off * P;
  uint8_t *dense = bs.scratch.as<uint8_t>();
  if (write) {
    HIPCHK(c, hipMemcpyAsync(dense, host, n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launchSlotColumn(p, uint32_t(P), bs.slotOf[instance], n, dense, true, c->stream));
  } else {
    HIPCHK(c, launchSlotColumn(p, uint32_t(P), bs.slotOf[instance], n, dense, false, c->stream));
    HIPCHK(c, hipMemcpyAsync(host, dense, n, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TGS_OK;
}

int tgs_read_pool(tgs_ctx *c, uint32_t instance, int pool, uint64_t off, uint64_t n, uint8_t *out) {
  return poolIO(c, instance, pool, off, n, out, false);
}

int tgs_write_pool(tgs_ctx *c, uint32_t instance, int pool, uint64_t off, uint64_t n, const uint8_t *src) {
  return poolIO(c, instance, pool, off, n, const_cast<uint8_t *>(src), true);
}

int tgs_read_regs(tgs_ctx *c, uint32_t instance, const uint32_t *ids, uint32_t n, uint8_t *out) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (n && (!ids || !out)) return TGS_E_ARG;
  for (uint32_t i = 0; i < n; ++i)
    if (ids[i] >= c->design.regPoolSize) return c->fail(TGS_E_ARG, "register id %u out of range", ids[i]);
  std::vector<uint8_t> pool(c->design.regPoolSize);
  if ((rc = poolIO(c, instance, TGS_POOL_REG, 0, pool.size(), pool.data(), false))) return rc;
  for (uint32_t i = 0; i < n; ++i) out[i] = pool[ids[i]];
  return TGS_OK;
}

int tgs_write_regs(tgs_ctx *c, uint32_t instance, const uint32_t *ids, uint32_t n, const uint8_t *vals) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (n && (!ids || !vals)) return TGS_E_ARG;
  for (uint32_t i = 0; i < n; ++i)
    if (ids[i] >= c->design.regPoolSize) return c->fail(TGS_E_ARG, "register id %u out of range", ids[i]);
  std::vector<uint8_t> pool(c->design.regPoolSize);
  if ((rc = poolIO(c, instance, TGS_POOL_REG, 0, pool.size(), pool.data(), false))) return rc;
  for (uint32_t i = 0; i < n; ++i) pool[ids[i]] = vals[i];
  return poolIO(c, instance, TGS_POOL_REG, 0, pool.size(), pool.data(), true);
}

int tgs_step(tgs_ctx *c, const uint32_t *writeIds, uint32_t nWrite, const uint8_t *writeVals, uint32_t cycles,
             const uint32_t *readIds, uint32_t nRead, uint8_t *readVals, int *anyStopped) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if ((nWrite && (!writeIds || !writeVals)) || (nRead && (!readIds || !readVals))) return TGS_E_ARG;
  const uint32_t regs = c->design.regPoolSize, ni = instanceCount(c);
  for (uint32_t k = 0; k < nWrite; ++k)
    if (writeIds[k] >= regs) return c->fail(TGS_E_ARG, "register id %u out of range", writeIds[k]);
  for (uint32_t k = 0; k < nRead; ++k)
    if (readIds[k] >= regs) return c->fail(TGS_E_ARG, "register id %u out of range", readIds[k]);
  {
    std::vector<uint32_t> w(writeIds, writeIds + nWrite);
    std::sort(w.begin(), w.end());
    if (std::adjacent_find(w.begin(), w.end()) != w.end())
      return c->fail(TGS_E_ARG, "a register id is written twice in one step");
  }
  if ((rc = setDevice(c))) return rc;
  // A step that failed may have left copies or kernels queued that still
  // read the mailbox: wait for them before the mailbox is rewritten.
  if (c->stepFailed) HIPCHK(c, hipStreamSynchronize(c->stream));
  c->stepFailed = true;  // until this step returns successfully
  // mailbox: [write ids][read ids][write values, per instance][read values,
  // per instance][control words 0..kCtlBack-1 per instance: stop flag,
  // barrier ticks of the last launch]
  constexpr uint32_t kCtlBack = C_TOTTICKS + 2;
  const size_t oRead = 4ull * nWrite, oWv = oRead + 4ull * nRead, oRv = oWv + size_t(ni) * nWrite;
  const size_t oCtl = (oRv + size_t(ni) * nRead + 7) & ~size_t(7);
  const size_t bytes = oCtl + 4ull * kCtlBack * ni;
  if (bytes > c->box.n) {
    HIPCHK(c, hipStreamSynchronize(c->stream));  // no copy of a failed earlier step may still use the old box
    HIPCHK(c, c->box.alloc(bytes));
    HIPCHK(c, c->boxDev.alloc(bytes));
  }
  c->noteWrite(writeVals, size_t(ni) * nWrite);
  uint8_t *h = c->box.host<uint8_t>(), *d = c->boxDev.as<uint8_t>();
  std::memcpy(h, writeIds, 4ull * nWrite);
  std::memcpy(h + oRead, readIds, 4ull * nRead);
  std::memcpy(h + oWv, writeVals, size_t(ni) * nWrite);
  // One instance on the streamed or wave64 kernel: the step is one launch
  // (the kernel applies the writes before its first cycle and takes the
  // reads after its last, through the pinned mailbox) and one
  // synchronization; no copies, no mailbox kernels.
  if (!c->batched && ni == 1 && cycles > 0 && c->kernel != TGS_KERNEL_PLAIN) {
    Instance &in = *c->inst[0];
    const uint8_t *hd = c->box.dev<uint8_t>();
    const StepBox mb{reinterpret_cast<const uint32_t *>(hd), hd + oWv, nWrite,
                     reinterpret_cast<const uint32_t *>(hd + oRead), const_cast<uint8_t *>(hd) + oRv, nRead};
    if ((rc = runInstance(c, in, cycles, false, &mb))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->timingPending) {
      c->timingPending = false;
      float ms = 0;
      HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
      c->prof.kernel_ms += ms;
    }
    const uint32_t *ctl = in.ctlPin.host<uint32_t>();
    accountTicks(c, in, ctl);
    if (nRead) std::memcpy(readVals, h + oRv, nRead);
    if (anyStopped) *anyStopped = ctl[C_STOP] != 0;
    c->stepFailed = false;
    return TGS_OK;
  }
  auto box = [&](const uint32_t *ids, uint32_t n, uint8_t *vals, bool toPool) -> int {
    if (c->batched) {
      HIPCHK(c, launchRegMailbox(c->bs.reg.as<uint8_t>(), c->bs.instPad, c->bs.instOfD.as<uint32_t>(), c->bs.n, ids,
                                 n, vals, toPool, c->stream));
    } else {
      for (uint32_t i = 0; i < ni; ++i)
        HIPCHK(c, launchRegMailbox(c->inst[i]->reg.as<uint8_t>(), 1, nullptr, 1, ids, n, vals + size_t(i) * n, toPool,
                                   c->stream));
    }
    return TGS_OK;
  };
  if (nWrite) {
    HIPCHK(c, hipMemcpyAsync(d, h, oRv, hipMemcpyHostToDevice, c->stream));
    if ((rc = box(reinterpret_cast<const uint32_t *>(d), nWrite, d + oWv, true))) return rc;
  } else if (nRead) {
    HIPCHK(c, hipMemcpyAsync(d + oRead, h + oRead, 4ull * nRead, hipMemcpyHostToDevice, c->stream));
  }
  int any = 0;
  if (c->batched) {
    if (cycles) {
      std::vector<uint32_t> all(ni, cycles);
      if ((rc = tgs_run_batch(c, all.data(), ni, &any))) return rc;
    } else {  // every slot's sticky stop flag in one copy
      std::vector<uint64_t> st(c->bs.instPad);
      HIPCHK(c, hipMemcpyAsync(st.data(), c->bs.stop.p, 8 * st.size(), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (auto v : st) any |= v != 0;
    }
  } else {
    // the last launch is not waited for: its timing, the stop flags and its
    // barrier ticks come back with the read values (the gather copies each
    // instance's control words)
    for (uint32_t i = 0; i < ni; ++i)
      if (cycles && (rc = runInstance(c, *c->inst[i], cycles, i + 1 < ni))) return rc;
  }
  if (c->batched) {
    if (nRead && (rc = box(reinterpret_cast<const uint32_t *>(d + oRead), nRead, d + oRv, false))) return rc;
  } else {
    for (uint32_t i = 0; i < ni; ++i)
      HIPCHK(c, launchRegMailbox(c->inst[i]->reg.as<uint8_t>(), 1, nullptr, 1,
                                 reinterpret_cast<const uint32_t *>(d + oRead), nRead, d + oRv + size_t(i) * nRead,
                                 false, c->stream, c->inst[i]->ctl.as<uint32_t>(),
                                 reinterpret_cast<uint32_t *>(d + oCtl) + size_t(kCtlBack) * i, kCtlBack));
  }
  // one copy back: read values, stop words and ticks
  const size_t back0 = oRv, back1 = c->batched ? oRv + size_t(ni) * nRead : bytes;
  if (back1 > back0) HIPCHK(c, hipMemcpyAsync(h + back0, d + back0, back1 - back0, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->timingPending) {
    c->timingPending = false;
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->prof.kernel_ms += ms;
    if (c->kernel == TGS_KERNEL_WAVE64 || c->kernel == TGS_KERNEL_STREAM)
      accountTicks(c, *c->inst[ni - 1], reinterpret_cast<const uint32_t *>(h + oCtl) + size_t(kCtlBack) * (ni - 1));
  }
  if (!c->batched)
    for (uint32_t i = 0; i < ni; ++i) {
      uint32_t w = 0;
      std::memcpy(&w, h + oCtl + 4ull * (size_t(kCtlBack) * i + C_STOP), 4);
      any |= w != 0;
    }
  if (nRead) std::memcpy(readVals, h + oRv, size_t(ni) * nRead);
  if (anyStopped) *anyStopped = any;
  c->stepFailed = false;
  return TGS_OK;
}

int tgs_get_stop(tgs_ctx *c, uint32_t instance, int *stopped) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (!stopped || instance >= instanceCount(c)) return c->fail(TGS_E_ARG, "bad instance or pointer");
  if (!c->batched) return readStop(c, *c->inst[instance], stopped);
  uint64_t v = 0;
  HIPCHK(c, hipMemcpy(&v, c->bs.stop.as<uint64_t>() + c->bs.slotOf[instance], 8, hipMemcpyDeviceToHost));
  *stopped = v != 0;
  return TGS_OK;
}

int tgs_reset_instance(tgs_ctx *c, uint32_t instance, uint64_t seed, uint32_t index) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (instance >= instanceCount(c)) return c->fail(TGS_E_ARG, "instance %u out of range", instance);
  if ((rc = setDevice(c))) return rc;
  if ((rc = drainLogs(c))) return rc;
  c->logs[instance].clear();  // the new run's log starts empty
  if (!c->batched) return resetInstance(c, *c->inst[instance], index, seed);
  auto &b = c->bs;
  const auto &d = c->design;
  std::vector<uint8_t> reg(d.regPoolSize, 0), mem(d.memPoolSize, 0), exg(d.exchangePoolSize, 0);
  std::copy(c->initReg.begin(), c->initReg.end(), reg.begin());
  std::copy(c->initMem.begin(), c->initMem.end(), mem.begin());
  if ((rc = tgs_write_pool(c, instance, TGS_POOL_REG, 0, reg.size(), reg.data()))) return rc;
  if ((rc = tgs_write_pool(c, instance, TGS_POOL_MEM, 0, mem.size(), mem.data()))) return rc;
  if ((rc = tgs_write_pool(c, instance, TGS_POOL_EXCHANGE, 0, exg.size(), exg.data()))) return rc;
  const uint32_t s = b.slotOf[instance];
  const uint64_t xs = seedFor(seed, index), zero = 0, never = UINT64_MAX;
  HIPCHK(c, hipMemcpy(b.xs.as<uint64_t>() + s, &xs, 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(b.stop.as<uint64_t>() + s, &zero, 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(b.stopCycle.as<uint64_t>() + s, &never, 8, hipMemcpyHostToDevice));
  b.cycles[instance] = 0;
  return TGS_OK;
}

int tgs_snapshot_begin(tgs_ctx *c, uint32_t instance, uint64_t *ticket) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (!ticket || instance >= instanceCount(c)) return c->fail(TGS_E_ARG, "bad instance or pointer");
  if ((rc = setDevice(c))) return rc;
  auto &s = c->snap;
  const uint64_t nReg = c->design.regPoolSize, nPr = c->nProbe(), n = nReg + nPr;
  if (s.slots == 0 || s.bytes < n) {
    s.release();
    s.slots = 4;
    s.bytes = (n + 255) & ~uint64_t(255);
    HIPCHK(c, hipStreamCreateWithFlags(&s.side, hipStreamNonBlocking));
    HIPCHK(c, hipMalloc(&s.staging, s.slots * std::max<uint64_t>(s.bytes, 256)));
    s.host.assign(s.slots, nullptr);
    s.staged.assign(s.slots, nullptr);
    s.done.assign(s.slots, nullptr);
    s.ticketOf.assign(s.slots, UINT64_MAX);
    s.cycleOf.assign(s.slots, 0);
    s.probes.assign(s.slots, 0);
    s.ordered.assign(s.slots, false);
    for (uint32_t k = 0; k < s.slots; ++k) {
      HIPCHK(c, hipHostMalloc(&s.host[k], std::max<uint64_t>(s.bytes, 256), hipHostMallocDefault));
      HIPCHK(c, hipEventCreateWithFlags(&s.staged[k], hipEventDisableTiming));
      HIPCHK(c, hipEventCreateWithFlags(&s.done[k], hipEventDisableTiming));
    }
  }
  const uint32_t k = uint32_t(s.next % s.slots);
  if (s.ticketOf[k] != UINT64_MAX) HIPCHK(c, hipEventSynchronize(s.done[k]));  // slot still in flight
  uint8_t *stage = static_cast<uint8_t *>(s.staging) + k * s.bytes;
  if (c->batched) {
    const auto &b = c->bs;
    if (nReg) HIPCHK(c, launchSlotColumn(b.reg.as<uint8_t>(), b.instPad, b.slotOf[instance], nReg, stage, false, c->stream));
    if (nPr)
      HIPCHK(c, launchSlotColumn(b.probe.as<uint8_t>(), b.instPad, b.slotOf[instance], nPr, stage + nReg, false,
                                 c->stream));
  } else {
    if (nReg) HIPCHK(c, hipMemcpyAsync(stage, c->inst[instance]->reg.p, nReg, hipMemcpyDeviceToDevice, c->stream));
    if (nPr)
      HIPCHK(c, hipMemcpyAsync(stage + nReg, c->inst[instance]->probe.p, nPr, hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(c, hipEventRecord(s.staged[k], c->stream));
  HIPCHK(c, hipStreamWaitEvent(s.side, s.staged[k], 0));
  if (n) HIPCHK(c, hipMemcpyAsync(s.host[k], stage, n, hipMemcpyDeviceToHost, s.side));
  HIPCHK(c, hipEventRecord(s.done[k], s.side));
  s.ticketOf[k] = s.next;
  s.cycleOf[k] = c->batched ? c->bs.cycles[instance] : c->inst[instance]->cycles;
  s.probes[k] = uint32_t(nPr);
  s.ordered[k] = false;
  *ticket = s.next++;
  return TGS_OK;
}

int tgs_snapshot_wait(tgs_ctx *c, uint64_t ticket, const uint8_t **data, uint64_t *cycle) {
  if (!c || !data) return TGS_E_ARG;
  auto &s = c->snap;
  if (s.slots == 0) return c->fail(TGS_E_STATE, "no snapshot was started");
  const uint32_t k = uint32_t(ticket % s.slots);
  if (s.ticketOf[k] != ticket) return c->fail(TGS_E_STATE, "snapshot %llu was overwritten or never started",
                                               (unsigned long long)ticket);
  int rc = setDevice(c);
  if (rc) return rc;
  HIPCHK(c, hipEventSynchronize(s.done[k]));
  // probe values arrive grouped by partition: put them in probe-list order once
  if (!s.ordered[k] && s.probes[k] && s.probes[k] == c->nProbe()) {
    uint8_t *pv = static_cast<uint8_t *>(s.host[k]) + c->design.regPoolSize;
    std::vector<uint8_t> csr(pv, pv + s.probes[k]);
    for (uint32_t i = 0; i < s.probes[k]; ++i) pv[i] = csr[c->probeCsrOf[i]];
  }
  s.ordered[k] = true;
  *data = static_cast<const uint8_t *>(s.host[k]);
  if (cycle) *cycle = s.cycleOf[k];
  return TGS_OK;
}

int tgs_set_probes(tgs_ctx *c, const uint32_t *parts, const uint32_t *slots, uint32_t n) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (n && (!parts || !slots)) return TGS_E_ARG;
  const auto &d = c->design;
  const uint32_t nParts = d.parts.size();
  for (uint32_t i = 0; i < n; ++i) {
    if (parts[i] >= nParts) return c->fail(TGS_E_ARG, "probe %u: partition %u out of range", i, parts[i]);
    if (slots[i] >= d.parts[parts[i]].valuePoolSize)
      return c->fail(TGS_E_ARG, "probe %u: slot %u outside partition %u's value pool", i, slots[i], parts[i]);
  }
  if ((rc = setDevice(c))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));  // no launch still reads the old list
  std::vector<uint32_t> off(nParts + 1, 0), fill(nParts, 0), slot(n), csrOf(n);
  for (uint32_t i = 0; i < n; ++i) off[parts[i] + 1]++;
  for (uint32_t q = 0; q < nParts; ++q) off[q + 1] += off[q];
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t at = off[parts[i]] + fill[parts[i]]++;
    slot[at] = slots[i];
    csrOf[i] = at;
  }
  c->probeCsrOf = csrOf;
  if ((rc = upload(c, c->prOff, off.data(), 4ull * off.size()))) return rc;
  if ((rc = upload(c, c->prSlot, slot.data(), 4ull * n))) return rc;
  // new, zeroed probe buffers; the rest of every instance's state (pools,
  // control words, print log) stays where it is
  if (c->batched) {
    auto &b = c->bs;
    const auto &d = c->design;
    HIPCHK(c, b.probe.alloc(std::max<uint64_t>(n, 1) * b.instPad));
    HIPCHK(c, hipMemsetAsync(b.probe.p, 0, b.probe.n, c->stream));
    const uint64_t rows = std::max<uint64_t>({regRows(d), memRows(d), exgRows(d), uint64_t(n)}) + 8;
    if (b.scratch.n < rows * b.instPad) HIPCHK(c, b.scratch.alloc(rows * b.instPad));
  } else {
    for (auto &in : c->inst) {
      HIPCHK(c, in->probe.alloc(std::max<uint32_t>(n, 1)));
      HIPCHK(c, hipMemsetAsync(in->probe.p, 0, in->probe.n, c->stream));
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TGS_OK;
}

int tgs_read_probes(tgs_ctx *c, uint32_t instance, uint8_t *out, uint32_t n) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (instance >= instanceCount(c)) return c->fail(TGS_E_ARG, "instance %u out of range", instance);
  if (n != c->nProbe()) return c->fail(TGS_E_ARG, "%u probes are set, not %u", c->nProbe(), n);
  if (n == 0) return TGS_OK;
  if (!out) return TGS_E_ARG;
  std::vector<uint8_t> v(n);
  if (c->batched) {
    auto &b = c->bs;
    uint8_t *dense = b.scratch.as<uint8_t>();
    HIPCHK(c, launchSlotColumn(b.probe.as<uint8_t>(), b.instPad, b.slotOf[instance], n, dense, false, c->stream));
    HIPCHK(c, hipMemcpyAsync(v.data(), dense, n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else {
    HIPCHK(c, hipMemcpy(v.data(), c->inst[instance]->probe.p, n, hipMemcpyDeviceToHost));
  }
  for (uint32_t i = 0; i < n; ++i) out[i] = v[c->probeCsrOf[i]];
  return TGS_OK;
}

// Checkpoint image: a fixed header, then the register, memory and exchange
// pools of one instance (nibble per byte, instance-major whatever the
// context's layout).
namespace {
struct CkptHeader {
  char magic[8];
  uint64_t designHash;
  uint32_t regBytes, memBytes, exgBytes, stop;
  uint64_t cycles, stopCycle, xs, reserved;
};
static_assert(sizeof(CkptHeader) == 64, "checkpoint header is 64 bytes");
constexpr char kCkptMagic[8] = {'T', 'G', 'S', 'C', 'K', 'P', 'T', '1'};

uint64_t ckptBytes(const SimDesignInfo &d) {
  return sizeof(CkptHeader) + uint64_t(d.regPoolSize) + d.memPoolSize + d.exchangePoolSize;
}
}  // namespace

int tgs_checkpoint_size(tgs_ctx *c, uint64_t *bytes) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (!bytes) return TGS_E_ARG;
  *bytes = ckptBytes(c->design);
  return TGS_OK;
}

int tgs_checkpoint_save(tgs_ctx *c, uint32_t instance, uint8_t *buf, uint64_t bytes) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  const auto &d = c->design;
  if (instance >= instanceCount(c)) return c->fail(TGS_E_ARG, "instance %u out of range", instance);
  if (!buf || bytes < ckptBytes(d)) return c->fail(TGS_E_ARG, "checkpoint buffer smaller than %llu bytes",
                                                   (unsigned long long)ckptBytes(d));
  if ((rc = setDevice(c))) return rc;
  CkptHeader h{};
  std::memcpy(h.magic, kCkptMagic, 8);
  h.designHash = c->designHash;
  h.regBytes = d.regPoolSize;
  h.memBytes = d.memPoolSize;
  h.exgBytes = d.exchangePoolSize;
  uint8_t *p = buf + sizeof h;
  if ((rc = poolIO(c, instance, TGS_POOL_REG, 0, h.regBytes, p, false))) return rc;
  if ((rc = poolIO(c, instance, TGS_POOL_MEM, 0, h.memBytes, p + h.regBytes, false))) return rc;
  if ((rc = poolIO(c, instance, TGS_POOL_EXCHANGE, 0, h.exgBytes, p + h.regBytes + h.memBytes, false))) return rc;
  if (c->batched) {
    const auto &b = c->bs;
    const uint32_t s = b.slotOf[instance];
    uint64_t st = 0;
    HIPCHK(c, hipMemcpy(&h.xs, b.xs.as<uint64_t>() + s, 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(&st, b.stop.as<uint64_t>() + s, 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(&h.stopCycle, b.stopCycle.as<uint64_t>() + s, 8, hipMemcpyDeviceToHost));
    h.stop = st != 0;
    h.cycles = b.cycles[instance];
  } else {
    Instance &in = *c->inst[instance];
    uint32_t ctl[C_WORDS];
    HIPCHK(c, hipMemcpy(&h.xs, in.xs.p, 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(ctl, in.ctl.p, sizeof ctl, hipMemcpyDeviceToHost));
    h.stop = ctl[C_STOP] != 0;
    std::memcpy(&h.stopCycle, &ctl[C_STOPCYC], 8);
    h.cycles = in.cycles;
  }
  std::memcpy(buf, &h, sizeof h);
  return TGS_OK;
}

int tgs_checkpoint_restore(tgs_ctx *c, uint32_t instance, const uint8_t *buf, uint64_t bytes) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  const auto &d = c->design;
  if (instance >= instanceCount(c)) return c->fail(TGS_E_ARG, "instance %u out of range", instance);
  CkptHeader h;
  if (!buf || bytes < sizeof h) return c->fail(TGS_E_FORMAT, "checkpoint shorter than its header");
  std::memcpy(&h, buf, sizeof h);
  if (std::memcmp(h.magic, kCkptMagic, 8) != 0) return c->fail(TGS_E_FORMAT, "not a checkpoint image");
  if (h.designHash != c->designHash || h.regBytes != d.regPoolSize || h.memBytes != d.memPoolSize ||
      h.exgBytes != d.exchangePoolSize)
    return c->fail(TGS_E_ARG, "checkpoint was taken on another design");
  if (bytes < ckptBytes(d)) return c->fail(TGS_E_FORMAT, "checkpoint truncated");
  if ((rc = setDevice(c))) return rc;
  if ((rc = drainLogs(c))) return rc;
  c->logs[instance].clear();  // print events are output of the replaced run, not state
  uint8_t *p = const_cast<uint8_t *>(buf) + sizeof h;
  if ((rc = poolIO(c, instance, TGS_POOL_REG, 0, h.regBytes, p, true))) return rc;
  if ((rc = poolIO(c, instance, TGS_POOL_MEM, 0, h.memBytes, p + h.regBytes, true))) return rc;
  if ((rc = poolIO(c, instance, TGS_POOL_EXCHANGE, 0, h.exgBytes, p + h.regBytes + h.memBytes, true))) return rc;
  if (c->batched) {
    auto &b = c->bs;
    const uint32_t s = b.slotOf[instance];
    const uint64_t st = h.stop;
    HIPCHK(c, hipMemcpy(b.xs.as<uint64_t>() + s, &h.xs, 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b.stop.as<uint64_t>() + s, &st, 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b.stopCycle.as<uint64_t>() + s, &h.stopCycle, 8, hipMemcpyHostToDevice));
    b.cycles[instance] = h.cycles;
  } else {
    Instance &in = *c->inst[instance];
    uint32_t ctl[C_WORDS];
    HIPCHK(c, hipMemcpy(ctl, in.ctl.p, sizeof ctl, hipMemcpyDeviceToHost));
    ctl[C_STOP] = h.stop;
    std::memcpy(&ctl[C_STOPCYC], &h.stopCycle, 8);
    HIPCHK(c, hipMemcpy(in.ctl.p, ctl, sizeof ctl, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(in.xs.p, &h.xs, 8, hipMemcpyHostToDevice));
    in.cycles = h.cycles;
  }
  return TGS_OK;
}

int tgs_get_cycles(tgs_ctx *c, uint32_t instance, uint64_t *cycles) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (!cycles || instance >= instanceCount(c)) return c->fail(TGS_E_ARG, "bad instance or pointer");
  *cycles = c->batched ? c->bs.cycles[instance] : c->inst[instance]->cycles;
  return TGS_OK;
}

int tgs_get_stop_cycle(tgs_ctx *c, uint32_t instance, uint64_t *cycle) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (!cycle || instance >= instanceCount(c)) return c->fail(TGS_E_ARG, "bad instance or pointer");
  if (c->batched) {
    HIPCHK(c, hipMemcpy(cycle, c->bs.stopCycle.as<uint64_t>() + c->bs.slotOf[instance], 8, hipMemcpyDeviceToHost));
  } else {
    HIPCHK(c, hipMemcpy(cycle, c->inst[instance]->ctl.as<uint32_t>() + C_STOPCYC, 8, hipMemcpyDeviceToHost));
  }
  return TGS_OK;
}

int tgs_get_print_log(tgs_ctx *c, uint32_t instance, uint64_t *cycles, uint32_t *msgs, uint64_t cap,
                      uint64_t *count) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (!count || instance >= instanceCount(c)) return c->fail(TGS_E_ARG, "bad instance or pointer");
  if ((rc = drainLogs(c))) return rc;
  const HostLog &L = c->logs[instance];
  *count = L.cyc.size();
  const uint64_t k = std::min<uint64_t>(cap, *count);
  if (k && cycles) std::memcpy(cycles, L.cyc.data(), k * 8);
  if (k && msgs) std::memcpy(msgs, L.msg.data(), k * 4);
  return TGS_OK;
}

int tgs_clear_print_log(tgs_ctx *c, uint32_t instance, uint64_t *dropped) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (instance >= instanceCount(c)) return c->fail(TGS_E_ARG, "instance %u out of range", instance);
  if ((rc = drainLogs(c))) return rc;
  if (dropped) *dropped = c->logs[instance].dropped;
  c->logs[instance].clear();
  return TGS_OK;
}

int tgs_get_profile(tgs_ctx *c, tgs_profile *out) {
  if (!c || !out) return TGS_E_ARG;
  *out = c->prof;
  return TGS_OK;
}

int tgs_reset_profile(tgs_ctx *c) {
  if (!c) return TGS_E_ARG;
  c->prof = tgs_profile{};
  return TGS_OK;
}

int tgs_set_part_profile(tgs_ctx *c, int enable, uint32_t start, uint32_t count) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  c->partProf = false;
  c->profCollected = 0;
  if (!enable) return TGS_OK;
  if (count == 0) return c->fail(TGS_E_ARG, "the profile window needs at least one cycle");
  const size_t parts = c->design.parts.size();
  // workgroups of any single-instance launch: at most two per CU (stream) or
  // one per partition of a region (plain, wave64)
  const size_t groups = std::max<size_t>(parts, 4ull * c->prop.multiProcessorCount);
  HIPCHK(c, c->profUseful.alloc(8 * std::max<size_t>(parts, 1)));
  HIPCHK(c, c->profOwner.alloc(4 * std::max<size_t>(parts, 1)));
  HIPCHK(c, c->profWg.alloc(16 * groups));
  HIPCHK(c, hipMemset(c->profOwner.p, 0, c->profOwner.n));
  c->profStart = start;
  c->profCount = count;
  c->partProf = true;
  return TGS_OK;
}

int tgs_get_part_profile(tgs_ctx *c, int64_t *useful, int64_t *total, size_t count, uint32_t *collected) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  const size_t parts = c->design.parts.size();
  if (count > parts || (count && (!useful || !total)))
    return c->fail(TGS_E_ARG, "at most %zu partitions", parts);
  if (collected) *collected = c->partProf ? c->profCollected : 0;
  if (!c->partProf || c->profCollected == 0) {
    for (size_t p = 0; p < count; ++p) useful[p] = total[p] = 0;
    return TGS_OK;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<uint64_t> u(parts), wg(c->profWg.n / 8);
  std::vector<uint32_t> owner(parts);
  HIPCHK(c, hipMemcpy(u.data(), c->profUseful.p, 8 * parts, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(owner.data(), c->profOwner.p, 4 * parts, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(wg.data(), c->profWg.p, c->profWg.n, hipMemcpyDeviceToHost));
  const size_t n0 = c->design.regionPartitionIds[0].size();
  const uint64_t k = c->profCollected;
  for (size_t p = 0; p < count; ++p) {
    const size_t at = 2ull * owner[p] + (p < n0 ? 0 : 1);
    useful[p] = int64_t(u[p] / k);
    total[p] = at < wg.size() ? int64_t(wg[at] / k) : 0;
  }
  return TGS_OK;
}

int tgs_get_program_stats(tgs_ctx *c, tgs_program_stats *out) {
  if (!c || !out) return TGS_E_ARG;
  if (!c->loaded) return c->fail(TGS_E_STATE, "no design loaded");
  *out = c->stats;
  return TGS_OK;
}

}  // extern "C"
