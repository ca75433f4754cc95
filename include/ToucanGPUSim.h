// Simulator API: drop-in for the reference's include/ToucanGPUSim.h:10-74.
//
// Same class, method names, argument meanings and return conventions
// (init returns 0 / -1; eval* return the sticky stop flag). The device side
// is the MI355X runtime behind include/toucan_rt.h; the object owns one
// runtime context instead of the reference's process globals, so several
// simulators may live in one process.
#pragma once

#include <cstddef>
#include <cstdint>
#include <fstream>
#include <string>
#include <tuple>
#include <unordered_map>
#include <utility>
#include <vector>

#include "SimConfigs.h"
#include "ToucanGPUGenDataTypes.h"

struct tgs_ctx;

namespace toucanGPUSim {

class ToucanSimulator {
 public:
  SimDesignInfo design;
  SimDebugInfo symbols;

  int init(const int gpu_id, const std::string designBinFilename, const std::string symbolBinFilename,
           bool shouldRandomize, bool enablePrint);
  void enableVCDDump(const std::string vcdFilename);
  bool eval();
  bool eval_free_running(uint32_t max_cycles);
  void setRegister(const std::string &signalName, uint64_t signalValue);
  uint64_t readRegister(const std::string &signalName);
  uint64_t readSignal(const std::string &signalName);
  void cleanup();
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
  tgs_ctx *context() const { return ctx; }

  ~ToucanSimulator() { cleanup(); }

  uint64_t cycle_cnt = 0;

 private:
  tgs_ctx *ctx = nullptr;
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
