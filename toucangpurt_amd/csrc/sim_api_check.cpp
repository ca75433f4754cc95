// Exercises toucanGPUSim::ToucanSimulator the way a reference testbench does
// (main.cpp:43-60 plus setRegister / readRegister, src/ToucanSim.cu:271-289).
// Usage: sim_api_check <design.bin> <symbols.bin> <vcd path> <cycles> <kernel>
// Prints "reg <name> <value>" and "sig <name> <value>" lines after the run for
// the test to compare.
#include <algorithm>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "ToucanGPUSim.h"
#include "toucan_rt.h"

int main(int argc, char **argv) {
  if (argc < 6) {
    std::cerr << "usage: sim_api_check design symbols vcd cycles kernel\n";
    return 2;
  }
  toucanGPUSim::ToucanSimulator sim;
  if (sim.init(0, "/nonexistent/design.bin", argv[2], false, false) != -1) {
    std::cerr << "missing file must return -1\n";
    return 1;
  }
  if (sim.init(0, argv[1], argv[2], true, false) != 0) return 1;
  if (sim.setKernel(std::atoi(argv[5])) != TGS_OK) return 1;
  std::vector<std::string> names;
  for (auto &kv : sim.symbols.regDebugInfo) names.push_back(kv.first);
  std::sort(names.begin(), names.end());
  // write a known value into every word register, read it back
  for (auto &n : names) {
    auto &frags = sim.symbols.regDebugInfo[n];
    if (frags.size() < 2) continue;
    uint64_t v = 0x0123456789ABCDEFull & ((frags.size() >= 16) ? ~0ull : ((1ull << (4 * frags.size())) - 1));
    sim.setRegister(n, v);
    if (sim.readRegister(n) != v) {
      std::cerr << "readRegister mismatch on " << n << "\n";
      return 1;
    }
  }
  std::vector<std::string> sigs;
  for (auto &kv : sim.symbols.signalDebugInfo) sigs.push_back(kv.first);
  std::sort(sigs.begin(), sigs.end());
  // a signal is observable from the first cycle after it is first named
  if (!sigs.empty() && sim.readSignal(sigs[0]) != UINT64_MAX) {
    std::cerr << "readSignal before any probed cycle must return UINT64_MAX\n";
    return 1;
  }
  sim.enableVCDDump(argv[3]);
  int cycles = std::atoi(argv[4]);
  for (int c = 0; c < cycles; ++c) sim.eval();
  sim.eval_free_running(uint32_t(cycles));
  for (auto &n : names) std::cout << "reg " << n << " " << sim.readRegister(n) << "\n";
  for (auto &n : sigs) std::cout << "sig " << n << " " << sim.readSignal(n) << "\n";
  std::cout << "cycles " << sim.cycle_cnt << "\n";
  sim.printProfile();
  sim.cleanup();
  return 0;
}
