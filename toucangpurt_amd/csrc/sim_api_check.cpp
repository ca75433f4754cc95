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
  // step(): one runtime call equals setRegister / eval / readRegister on a
  // second and third simulator of the same design
  {
    std::vector<std::string> words;
    for (auto &n : names)
      if (sim.symbols.regDebugInfo[n].size() >= 2 && sim.symbols.regDebugInfo[n].size() <= 16) words.push_back(n);
    if (words.size() >= 2) {
      toucanGPUSim::ToucanSimulator a, b;
      if (a.init(0, argv[1], argv[2], true, false) != 0 || b.init(0, argv[1], argv[2], true, false) != 0) return 1;
      if (a.setKernel(std::atoi(argv[5])) != TGS_OK || b.setKernel(std::atoi(argv[5])) != TGS_OK) return 1;
      const size_t half = words.size() / 2;
      for (uint32_t s = 0; s < 4; ++s) {
        std::vector<std::pair<std::string, uint64_t>> w;
        for (size_t k = 0; k < half; ++k) {
          const uint32_t nib = uint32_t(a.symbols.regDebugInfo[words[k]].size());
          const uint64_t mask = nib >= 16 ? ~0ull : ((1ull << (4 * nib)) - 1);
          w.emplace_back(words[k], (0x9E3779B97F4A7C15ull * (s + 1) + k) & mask);
        }
        std::vector<std::string> r(words.begin() + half / 2, words.end());
        std::vector<uint64_t> got;
        const bool sa = a.step(w, s, r, got);
        for (auto &x : w) b.setRegister(x.first, x.second);
        bool sb = false;
        for (uint32_t c = 0; c < s; ++c) sb = b.eval();
        for (size_t k = 0; k < r.size(); ++k)
          if (got[k] != b.readRegister(r[k])) {
            std::cerr << "step mismatch on " << r[k] << " at step " << s << "\n";
            return 1;
          }
        if (s > 0 && sa != sb) {
          std::cerr << "step stop flag mismatch at step " << s << "\n";
          return 1;
        }
      }
      if (a.cycle_cnt != b.cycle_cnt) return 1;
      std::cout << "step ok " << a.cycle_cnt << "\n";
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
