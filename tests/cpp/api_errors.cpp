// Error paths of toucanGPUSim::ToucanSimulator and the include/SimEval.h
// shim: every misuse returns an observable status instead of ending the
// process (the reference asserts and exits, src/ToucanSim.cu:271-289,
// include/SimEval.h:11-19).
// Usage: api_errors <design.bin> <symbols.bin>
// Prints "loaded" when a GPU took the design, then "ok"; exits 1 on a
// mismatch, naming it.
#include <cstdint>
#include <iostream>
#include <string>

#include "SimEval.h"
#include "ToucanGPUSim.h"
#include "toucan_rt.h"

#define EXPECT(cond)                                                  \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::cout << "FAILED: " #cond " (line " << __LINE__ << ")\n"; \
      return 1;                                                       \
    }                                                                 \
  } while (0)

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  toucanGPUSim::ToucanSimulator sim;
  // before init: every call fails with TGS_E_STATE
  EXPECT(sim.eval() == false && sim.lastStatus() == TGS_E_STATE);
  EXPECT(sim.eval_free_running(10) == false && sim.lastStatus() == TGS_E_STATE);
  EXPECT(sim.readRegister("r") == 0 && sim.lastStatus() == TGS_E_STATE);
  sim.setRegister("r", 1);
  EXPECT(sim.lastStatus() == TGS_E_STATE);
  EXPECT(sim.readSignal("s") == UINT64_MAX && sim.lastStatus() == TGS_E_STATE);
  std::vector<uint64_t> out;
  EXPECT(sim.step({}, 1, {}, out) == false && sim.lastStatus() == TGS_E_STATE);
  EXPECT(!sim.lastError().empty());
  // the shim has no device state yet
  EXPECT(get_eval_done() == false);
  EXPECT(get_real_cycles() == 0);
  EXPECT(read_reg_from_gpu({{0u, 4u}}) == 0);
  // a missing file: -1 and TGS_E_IO (src/ToucanSim.cu:108-122)
  EXPECT(sim.init(0, "/nonexistent/GPUSimDesign.bin", argv[2], false, false) == -1);
  EXPECT(sim.lastStatus() == TGS_E_IO);
  const int rc = sim.init(0, argv[1], argv[2], false, false);
  if (rc != 0) {  // no usable GPU here: init says so and the object stays usable
    EXPECT(sim.lastStatus() == TGS_E_NODEV);
    EXPECT(sim.eval() == false && sim.lastStatus() == TGS_E_STATE);
    std::cout << "ok\n";
    return 0;
  }
  std::cout << "loaded\n";
  // unknown names and values too wide for the register
  EXPECT(sim.readRegister("no.such.register") == 0 && sim.lastStatus() == TGS_E_ARG);
  sim.setRegister("no.such.register", 1);
  EXPECT(sim.lastStatus() == TGS_E_ARG);
  EXPECT(sim.readSignal("no.such.signal") == UINT64_MAX && sim.lastStatus() == TGS_E_ARG);
  std::string narrow;
  for (auto &kv : sim.symbols.regDebugInfo)
    if (kv.second.size() == 1) narrow = kv.first;
  EXPECT(!narrow.empty());
  sim.setRegister(narrow, 0x12);  // one nibble: the low nibble lands, the call reports the overflow
  EXPECT(sim.lastStatus() == TGS_E_ARG);
  EXPECT(sim.readRegister(narrow) == 0x2 && sim.lastStatus() == 0);
  EXPECT(sim.step({{"no.such.register", 1}}, 1, {}, out) == false && sim.lastStatus() == TGS_E_ARG);
  // the simulator still runs after the errors, and the shim follows it
  sim.eval_free_running(3);
  EXPECT(sim.lastStatus() == 0 && sim.cycle_cnt == 3);
  EXPECT(get_real_cycles() == 3);
  std::vector<std::tuple<uint32_t, uint32_t>> locs;
  for (auto &l : sim.symbols.regDebugInfo[narrow]) locs.push_back({std::get<0>(l), std::get<1>(l)});
  EXPECT(read_reg_from_gpu(locs) == sim.readRegister(narrow));
  write_reg_to_gpu(locs, 0x5);
  EXPECT(sim.readRegister(narrow) == 0x5);
  sim.cleanup();
  EXPECT(get_real_cycles() == 0);  // the current context went with the simulator
  std::cout << "ok\n";
  return 0;
}
