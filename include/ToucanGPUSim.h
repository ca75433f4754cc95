// Simulator API: drop-in for the reference's include/ToucanGPUSim.h:10-74.
//
// Same class, method names, argument meanings and return conventions
// (init returns 0 / -1; eval* return the sticky stop flag). The device side
// is the MI355X runtime behind include/toucan_rt.h; the object owns one
// runtime context instead of the reference's process globals, so several
// simulators may live in one process.
//
// Errors never end the process (the reference asserts on unknown names and
// exits on CUDA errors, include/SimEval.h:11-19, src/ToucanSim.cu:271-289):
// a failing call prints the reason on stderr, returns false / 0 /
// UINT64_MAX, and leaves the status in lastStatus() (a TGS_E_* code of
// include/toucan_rt.h, 0 after a successful call) and the text in
// lastError().
#pragma once

#include <cstddef>
#include <cstdint>
#include <fstream>
#include <string>
#include <tuple>
#include <unordered_map>
#include <utility>
#include <vector>

#include "GPUConsts.h"
#include "SimConfigs.h"
#include "ToucanGPUGenDataTypes.h"

struct tgs_ctx;

namespace toucanGPUSim {

class ToucanSimulator {
 public:
  SimDesignInfo design;
  SimDebugInfo symbols;

  // With ENABLE_SIM_PROFILE defined where this header is included (as in
  // the reference's SimConfigs.h), init also turns on the per-partition
  // profile of cycles PROFILE_START_CYCLE .. + PROFILE_COLLECT_CYCLE of each
  // free-running launch (src/SimEval.cu:894-983) that printProfile reports.
  int init(const int gpu_id, const std::string designBinFilename, const std::string symbolBinFilename,
           bool shouldRandomize, bool enablePrint) {
#ifdef ENABLE_SIM_PROFILE
    return initWithProfile(gpu_id, designBinFilename, symbolBinFilename, shouldRandomize, enablePrint, true);
#else
    return initWithProfile(gpu_id, designBinFilename, symbolBinFilename, shouldRandomize, enablePrint, false);
#endif
  }
  int initWithProfile(const int gpu_id, const std::string &designBinFilename, const std::string &symbolBinFilename,
                      bool shouldRandomize, bool enablePrint, bool profileParts);
  void enableVCDDump(const std::string vcdFilename);
  bool eval();
  bool eval_free_running(uint32_t max_cycles);
  void setRegister(const std::string &signalName, uint64_t signalValue);
  uint64_t readRegister(const std::string &signalName);
  uint64_t readSignal(const std::string &signalName);
  void cleanup();
  // src/ToucanSim.cu:306-325: one line per partition (useful share, useful
  // and total ticks) when the per-partition profile was collected, after a
  // line with the launch shape and the measured kernel time.
  void printProfile();

  // Extension: one testbench step in one runtime call (tgs_step): set the
  // registers `writes` (name, value; as setRegister), run `cycles` cycles,
  // read the registers `reads` into `out` (as readRegister, same order).
  // Returns the sticky stop flag like eval(). With a VCD dump open it runs
  // setRegister / eval / readRegister so every cycle is dumped.
  bool step(const std::vector<std::pair<std::string, uint64_t>> &writes, uint32_t cycles,
            const std::vector<std::string> &reads, std::vector<uint64_t> &out);

  // Extensions: kernel family (TGS_KERNEL_*), stimulus, last runtime error.
  int setKernel(int kernel);
  int setStimulus(const std::vector<uint32_t> &inputIds, uint64_t seed);
  std::string lastError() const;
  int lastStatus() const { return status; }
  tgs_ctx *context() const { return ctx; }

  ~ToucanSimulator() { cleanup(); }

  uint64_t cycle_cnt = 0;

 private:
  tgs_ctx *ctx = nullptr;
  int status = 0;
  bool loadedOk = false;  // the last init succeeded
  std::string errText;
  bool fail(int code, const std::string &what);
  bool ready(const char *call);
  bool dumpVCD = false;
  std::ofstream *vcdOfs = nullptr;
  std::vector<std::string> vcdNames;
  // per VCD variable: (offset in the snapshot, bit width) per fragment
  std::vector<std::vector<std::tuple<uint32_t, uint32_t>>> vcdSrc;
  std::vector<uint32_t> vcdWidth;
  std::vector<std::string> vcdIds;
  std::vector<std::vector<uint8_t>> vcdLast;
  bool vcdPending = false;       // a snapshot started but not yet written
  uint64_t vcdTicket = 0, vcdStamp = 0;
  void dumpVcdWorker(uint64_t cycle);
  void writeVcdSnapshot(uint64_t ticket, uint64_t stamp);
  void closeVcd();
  void flushPrints();
  uint64_t printsSeen = 0;
  // signal probes (one per signal fragment) and where each signal starts
  std::vector<uint32_t> probeParts, probeSlots;
  std::unordered_map<std::string, uint32_t> probeFirst;
  uint64_t probeEpoch = UINT64_MAX;  // cycle_cnt when the probe list last changed
  uint32_t addSignalProbes(const std::vector<std::string> &names);
};

}  // namespace toucanGPUSim
