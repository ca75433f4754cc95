// Host shim of the simulation kernels: drop-in for the reference's
// include/SimEval.h:1-39, host-only.
//
// The reference header pulls in the CUDA runtime and declares the three
// __global__ cycle kernels (evalSingleCycle, evalFreeRunningNCycles,
// evalFreeRunningNCycles_Large) beside host functions that act on the
// process-global device state. Here the kernels are the gfx950 kernels behind
// include/toucan_rt.h, launched by tgs_eval / tgs_eval_free_running, so they
// are not declared; the host functions keep their names and signatures and
// act on the process's current simulator: the context of the
// ToucanSimulator that most recently returned from a successful init(), or
// the context copy_netlist_to_gpu created. Errors are reported on stderr and
// leave the state unchanged (the reference's gpuErrchk exits the process,
// include/SimEval.h:11-19).
#pragma once

#include <cstddef>
#include <cstdint>
#include <tuple>
#include <vector>

#include "SimConfigs.h"
#include "ToucanGPUGenDataTypes.h"

// read_reg_from_gpu (src/SimEval.cu:1058-1069): r = r << 4 | nibble over the
// (regId, width) fragments in list order. Fragments past the pool read 0.
uint64_t read_reg_from_gpu(const std::vector<std::tuple<uint32_t, uint32_t>> &signalLocs);

// write_reg_to_gpu (src/SimEval.cu:1072-1081): the last fragment takes the
// low nibble; the value must fit the fragments.
void write_reg_to_gpu(const std::vector<std::tuple<uint32_t, uint32_t>> &signalLocs, uint64_t signalValue);

// copy_netlist_to_gpu (src/SimEval.cu:1114-1343): upload a design with its
// current register and memory pools (the caller has run Init / Randomize),
// zero the exchange pool and the cycle counter. Creates the current context
// on device 0 when there is none.
void copy_netlist_to_gpu(toucanGPUSim::SimDesignInfo &design);

// get_eval_done (src/SimEval.cu:1083-1087): the sticky stop flag.
bool get_eval_done();

// get_real_cycles (src/SimEval.cu:1089-1093): cycles run since the upload.
uint32_t get_real_cycles();

// setEnablePrint (src/SimEval.cu:1095-1097).
void setEnablePrint(bool print_en);

#ifdef ENABLE_SIM_PROFILE
// copy_profile_data (src/SimEval.cu:1100-1104): per-partition average
// useful and total clock ticks of the profile window.
void copy_profile_data(int64_t *usefulTicks, int64_t *totalTicks, size_t count);
#endif
