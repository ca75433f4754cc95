// C-ABI runtime of the MI355X Toucan cycle simulator (include/toucan_rt.h).
//
// Replaces the reference's process-global device state and launch code
// (src/SimEval.cu:190-215,1058-1343; src/ToucanSim.cu:16-267) with per-context
// state: one context holds one loaded design, its device images and any
// number of resident instances (register / memory / exchange pools each).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <memory>
#include <string>
#include <vector>

#include "kernel_args.h"
#include "image.hpp"
#include "toucan_rt.h"


using namespace tgs;
using toucanGPUSim::SimDebugInfo;
using toucanGPUSim::SimDesignInfo;

namespace {

constexpr uint32_t kLogCap = 1u << 16;

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

struct Instance {
  DBuf reg, mem, exg, xs, ctl, logCycle, logMsg;
  uint64_t cycles = 0;
};

uint32_t up4(uint32_t x) { return (x + 3u) & ~3u; }
uint32_t up16(uint32_t x) { return (x + 15u) & ~15u; }

}  // namespace

struct tgs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipDeviceProp_t prop{};
  std::string err;

  bool loaded = false;
  SimDesignInfo design;
  SimDebugInfo symbols;
  std::vector<uint8_t> initReg, initMem;
  tgs_program_stats stats{};
  int kernel = TGS_KERNEL_WAVE64;
  int printEnable = 0;
  StimMap stim;
  uint64_t stimSeed = 0;
  std::vector<std::unique_ptr<Instance>> inst;

  DBuf lut, inIds;
  bool plainReady = false, w64Ready = false;
  PlainImage plain;
  DBuf pNet, pParts, pLevels, pMparts, pConsts;
  Wave64Image w64;
  Wave64Schedule sched;
  DBuf wWords, wPartOff, wPartWords, wSchedOff, wSchedPart, wSchedLds, wConsts;
  uint32_t w64Block = 512, w64Groups = 0, w64Lds = 0;
  bool w64Resident = false;
  uint32_t plainGrid = 0, plainBlock = 1024, plainLds = 0;

  tgs_profile prof{};

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

int upload(tgs_ctx *c, DBuf &b, const void *src, size_t n) {
  HIPCHK(c, b.alloc(n));
  if (n) HIPCHK(c, hipMemcpyAsync(b.p, src, n, hipMemcpyHostToDevice, c->stream));
  return TGS_OK;
}

// Reset one instance to the loaded initial state (copy_netlist_to_gpu,
// SimEval.cu:1114-1139: stop flag and cycle counter cleared, exchange zeroed).
int resetInstance(tgs_ctx *c, Instance &in, uint32_t index) {
  const auto &d = c->design;
  HIPCHK(c, in.reg.alloc(up4(d.regPoolSize) + 16));
  HIPCHK(c, in.mem.alloc(size_t(d.memPoolSize) + 16));
  HIPCHK(c, in.exg.alloc(up4(d.exchangePoolSize) + 16));
  HIPCHK(c, in.xs.alloc(8));
  HIPCHK(c, in.ctl.alloc(C_WORDS * 4));
  HIPCHK(c, in.logCycle.alloc(kLogCap * 8));
  HIPCHK(c, in.logMsg.alloc(kLogCap * 4));
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
  // splitmix64(seed + instance), never zero
  uint64_t z = c->stimSeed + index + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  if (z == 0) z = 0x9E3779B97F4A7C15ull;
  std::vector<uint64_t> xs(1, z);
  HIPCHK(c, hipMemcpyAsync(in.xs.p, xs.data(), 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  in.cycles = 0;
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
  const size_t limit = c->prop.sharedMemPerBlock;
  c->w64Resident = resLds <= limit;
  c->w64Lds = c->w64Resident ? resLds : base;
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
  HIPCHK(c, wave64MaxBlocksPerCU(&perCU, c->w64Resident, c->w64Block, c->w64Lds));
  if (perCU < 1) return c->fail(TGS_E_LIMIT, "wave64 kernel cannot be resident with %u B LDS", c->w64Lds);
  if (uint64_t(perCU) * c->prop.multiProcessorCount < groups)
    return c->fail(TGS_E_LIMIT, "%u workgroups exceed the co-resident limit", groups);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->w64Ready = true;
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
  return a;
}

// Largest cycle count per launch that keeps the barrier counter in range.
uint32_t maxCyclesPerLaunch(uint32_t groups) {
  uint64_t m = (0xF0000000ull / (2ull * std::max<uint32_t>(groups, 1)));
  return uint32_t(std::min<uint64_t>(m, 1u << 30));
}

int runInstance(tgs_ctx *c, Instance &in, uint32_t cycles) {
  while (cycles > 0) {
    int rc;
    if (c->kernel == TGS_KERNEL_PLAIN) {
      if ((rc = buildPlain(c))) return rc;
      uint32_t n = std::min(cycles, maxCyclesPerLaunch(c->plainGrid));
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
    } else {
      if ((rc = buildW64(c))) return rc;
      uint32_t n = std::min(cycles, maxCyclesPerLaunch(c->sched.groups));
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
      a.barrierBase = 0;
      HIPCHK(c, hipMemsetAsync(in.ctl.as<uint32_t>() + C_BARRIER, 0, 4, c->stream));
      HIPCHK(c, hipEventRecord(c->ev0, c->stream));
      HIPCHK(c, launchWave64(a, c->w64Resident, c->sched.groups, c->w64Block, c->w64Lds, c->stream));
      HIPCHK(c, hipEventRecord(c->ev1, c->stream));
      c->prof.grid = c->sched.groups;
      c->prof.block = c->w64Block;
      c->prof.lds_bytes = c->w64Lds;
      c->prof.resident = c->w64Resident;
      c->prof.prog_bytes = c->w64.words.size() * 4;
      cycles -= n;
      in.cycles += n;
      c->prof.cycles += n;
    }
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->prof.kernel_ms += ms;
    c->prof.launches += 1;
    c->prof.kernel = c->kernel;
    if (c->kernel == TGS_KERNEL_WAVE64) {
      uint64_t t[2];
      HIPCHK(c, hipMemcpy(t, in.ctl.as<uint32_t>() + C_BARTICKS, 16, hipMemcpyDeviceToHost));
      c->prof.barrier_ticks += t[0];
      c->prof.total_ticks += t[1];
    }
  }
  return TGS_OK;
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

Instance *instanceOf(tgs_ctx *c, uint32_t i) {
  if (i >= c->inst.size()) {
    c->fail(TGS_E_ARG, "instance %u out of range (%zu instances)", i, c->inst.size());
    return nullptr;
  }
  return c->inst[i].get();
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

int tgs_create(int device, tgs_ctx **out) {
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
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return TGS_E_HIP;
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) return TGS_E_HIP;
  const char *blk = std::getenv("TGS_W64_BLOCK");
  if (blk) c->w64Block = std::max(64u, std::min(1024u, uint32_t(std::atoi(blk)) / 64 * 64));
  const char *grp = std::getenv("TGS_W64_GROUPS");
  if (grp) c->w64Groups = uint32_t(std::atoi(grp));
  *out = c.release();
  return TGS_OK;
}

void tgs_destroy(tgs_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->inst.clear();
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
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
  c->plainReady = c->w64Ready = false;
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
  c->stim = StimMap();
  c->printEnable = enablePrint ? 1 : 0;
  programStats(c->design, 0, &c->stats);
  std::vector<uint8_t> lut(wave64LutBytes(), 0);
  std::copy(c->design.lut.begin(), c->design.lut.end(), lut.begin());
  if ((rc = upload(c, c->lut, lut.data(), lut.size()))) return rc;
  if ((rc = upload(c, c->inIds, nullptr, 0))) return rc;
  c->inst.emplace_back(new Instance());
  if ((rc = resetInstance(c, *c->inst[0], 0))) return rc;
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
  if (kernel != TGS_KERNEL_PLAIN && kernel != TGS_KERNEL_WAVE64) {
    if (kernel == TGS_KERNEL_BATCH) return c->fail(TGS_E_UNSUPPORTED, "batched kernel not built into this library");
    return c->fail(TGS_E_ARG, "unknown kernel %d", kernel);
  }
  c->kernel = kernel;
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
  c->inst.clear();
  for (uint32_t i = 0; i < n; ++i) {
    c->inst.emplace_back(new Instance());
    if ((rc = resetInstance(c, *c->inst.back(), i))) return rc;
  }
  return TGS_OK;
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
  c->plainReady = c->w64Ready = false;
  if ((rc = upload(c, c->inIds, c->stim.ids.data(), c->stim.ids.size() * 4))) return rc;
  programStats(c->design, n, &c->stats);
  // every instance restarts its generator from splitmix64(seed + k)
  for (uint32_t i = 0; i < c->inst.size(); ++i) {
    uint64_t z = seed + i + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    if (z == 0) z = 0x9E3779B97F4A7C15ull;
    HIPCHK(c, hipMemcpy(c->inst[i]->xs.p, &z, 8, hipMemcpyHostToDevice));
  }
  return TGS_OK;
}

int tgs_eval_free_running(tgs_ctx *c, uint32_t cycles, int *stopped) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  for (auto &in : c->inst)
    if ((rc = runInstance(c, *in, cycles))) return rc;
  if (stopped) return readStop(c, *c->inst[0], stopped);
  return TGS_OK;
}

int tgs_eval(tgs_ctx *c, int *stopped) { return tgs_eval_free_running(c, 1, stopped); }

int tgs_run_batch(tgs_ctx *c, const uint32_t *cycles, uint32_t n, int *anyStopped) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  if (!cycles || n != c->inst.size()) return c->fail(TGS_E_ARG, "need one cycle count per instance");
  int any = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if ((rc = runInstance(c, *c->inst[i], cycles[i]))) return rc;
    int s = 0;
    if ((rc = readStop(c, *c->inst[i], &s))) return rc;
    any |= s;
  }
  if (anyStopped) *anyStopped = any;
  return TGS_OK;
}

int tgs_read_regs(tgs_ctx *c, uint32_t instance, const uint32_t *ids, uint32_t n, uint8_t *out) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  Instance *in = instanceOf(c, instance);
  if (!in) return TGS_E_ARG;
  if (n && (!ids || !out)) return TGS_E_ARG;
  for (uint32_t i = 0; i < n; ++i)
    if (ids[i] >= c->design.regPoolSize) return c->fail(TGS_E_ARG, "register id %u out of range", ids[i]);
  std::vector<uint8_t> pool(c->design.regPoolSize);
  if (!pool.empty()) HIPCHK(c, hipMemcpy(pool.data(), in->reg.p, pool.size(), hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < n; ++i) out[i] = pool[ids[i]];
  return TGS_OK;
}

int tgs_write_regs(tgs_ctx *c, uint32_t instance, const uint32_t *ids, uint32_t n, const uint8_t *vals) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  Instance *in = instanceOf(c, instance);
  if (!in) return TGS_E_ARG;
  if (n && (!ids || !vals)) return TGS_E_ARG;
  for (uint32_t i = 0; i < n; ++i)
    if (ids[i] >= c->design.regPoolSize) return c->fail(TGS_E_ARG, "register id %u out of range", ids[i]);
  std::vector<uint8_t> pool(c->design.regPoolSize);
  if (!pool.empty()) HIPCHK(c, hipMemcpy(pool.data(), in->reg.p, pool.size(), hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < n; ++i) pool[ids[i]] = vals[i];
  if (!pool.empty()) HIPCHK(c, hipMemcpy(in->reg.p, pool.data(), pool.size(), hipMemcpyHostToDevice));
  return TGS_OK;
}

int tgs_pool_size(const tgs_ctx *c, int pool, uint64_t *size) {
  if (!c || !size || !c->loaded) return TGS_E_ARG;
  switch (pool) {
    case TGS_POOL_REG: *size = c->design.regPoolSize; return TGS_OK;
    case TGS_POOL_MEM: *size = c->design.memPoolSize; return TGS_OK;
    case TGS_POOL_EXCHANGE: *size = c->design.exchangePoolSize; return TGS_OK;
    default: return TGS_E_ARG;
  }
}

static int poolPtr(tgs_ctx *c, Instance *in, int pool, uint64_t off, uint64_t n, uint8_t **p) {
  uint64_t size = 0;
  if (tgs_pool_size(c, pool, &size)) return c->fail(TGS_E_ARG, "unknown pool %d", pool);
  if (off > size || n > size - off) return c->fail(TGS_E_ARG, "pool range out of bounds");
  DBuf &b = pool == TGS_POOL_REG ? in->reg : pool == TGS_POOL_MEM ? in->mem : in->exg;
  *p = b.as<uint8_t>() + off;
  return TGS_OK;
}

int tgs_read_pool(tgs_ctx *c, uint32_t instance, int pool, uint64_t off, uint64_t n, uint8_t *out) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  Instance *in = instanceOf(c, instance);
  if (!in || (n && !out)) return TGS_E_ARG;
  uint8_t *p;
  if ((rc = poolPtr(c, in, pool, off, n, &p))) return rc;
  if (n) HIPCHK(c, hipMemcpy(out, p, n, hipMemcpyDeviceToHost));
  return TGS_OK;
}

int tgs_write_pool(tgs_ctx *c, uint32_t instance, int pool, uint64_t off, uint64_t n, const uint8_t *src) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  Instance *in = instanceOf(c, instance);
  if (!in || (n && !src)) return TGS_E_ARG;
  uint8_t *p;
  if ((rc = poolPtr(c, in, pool, off, n, &p))) return rc;
  if (n) HIPCHK(c, hipMemcpy(p, src, n, hipMemcpyHostToDevice));
  return TGS_OK;
}

int tgs_get_stop(tgs_ctx *c, uint32_t instance, int *stopped) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  Instance *in = instanceOf(c, instance);
  if (!in || !stopped) return TGS_E_ARG;
  return readStop(c, *in, stopped);
}

int tgs_get_cycles(tgs_ctx *c, uint32_t instance, uint64_t *cycles) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  Instance *in = instanceOf(c, instance);
  if (!in || !cycles) return TGS_E_ARG;
  *cycles = in->cycles;
  return TGS_OK;
}

int tgs_get_stop_cycle(tgs_ctx *c, uint32_t instance, uint64_t *cycle) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  Instance *in = instanceOf(c, instance);
  if (!in || !cycle) return TGS_E_ARG;
  HIPCHK(c, hipMemcpy(cycle, in->ctl.as<uint32_t>() + C_STOPCYC, 8, hipMemcpyDeviceToHost));
  return TGS_OK;
}

int tgs_get_print_log(tgs_ctx *c, uint32_t instance, uint64_t *cycles, uint32_t *msgs, uint64_t cap,
                      uint64_t *count) {
  int rc = checkLoaded(c);
  if (rc) return rc;
  Instance *in = instanceOf(c, instance);
  if (!in || !count) return TGS_E_ARG;
  uint32_t n = 0;
  HIPCHK(c, hipMemcpy(&n, in->ctl.as<uint32_t>() + C_LOGCOUNT, 4, hipMemcpyDeviceToHost));
  *count = n;
  uint64_t k = std::min<uint64_t>({cap, n, kLogCap});
  if (k && cycles) HIPCHK(c, hipMemcpy(cycles, in->logCycle.p, k * 8, hipMemcpyDeviceToHost));
  if (k && msgs) HIPCHK(c, hipMemcpy(msgs, in->logMsg.p, k * 4, hipMemcpyDeviceToHost));
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

int tgs_get_program_stats(tgs_ctx *c, tgs_program_stats *out) {
  if (!c || !out) return TGS_E_ARG;
  if (!c->loaded) return c->fail(TGS_E_STATE, "no design loaded");
  *out = c->stats;
  return TGS_OK;
}

}  // extern "C"
