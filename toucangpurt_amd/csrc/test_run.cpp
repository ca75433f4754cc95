// Benchmark driver with the reference's CLI (main.cpp:33-68): loads
// GPUSimDesign.bin / GPUSimIOSymbols.bin from the working directory, runs
// eval_free_running(N) (default 1000) and prints the elapsed milliseconds.
// Optional second argument: kernel family (0 plain, 1 wave64).
#include <chrono>
#include <cstdlib>
#include <iostream>

#include "ToucanGPUSim.h"
#include "toucan_rt.h"

int main(int argc, char **argv) {
  int sim_cycles = 1000;
  if (argc >= 2) sim_cycles = std::atoi(argv[1]);
  toucanGPUSim::ToucanSimulator sim;
  std::cout << "Loading design" << std::endl;
  if (sim.init(0, "GPUSimDesign.bin", "GPUSimIOSymbols.bin", false, false) != 0) return 1;
  if (argc >= 3) sim.setKernel(std::atoi(argv[2]));
  std::cout << "Loading done" << std::endl;
  auto start = std::chrono::system_clock::now();
  sim.eval_free_running(static_cast<uint32_t>(sim_cycles));
  auto end = std::chrono::system_clock::now();
  auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(end - start).count();
  std::cout << "Free running for " << sim_cycles << " cycles in " << ms << " ms\n";
  sim.printProfile();
  return 0;
}
