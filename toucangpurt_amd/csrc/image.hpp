// Host-side program checks, statistics and device-image builders.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "ToucanGPUGenDataTypes.h"
#include "device_layout.h"
#include "toucan_rt.h"

namespace tgs {

// Format invariants the reference asserts (SimEval.cu:92-97,1122,1157-1160,
// 1183,1292-1311; ToucanSim.cu:153) plus index ranges and the exec-level
// hazard rule: micropartitions of one exec level run concurrently, so none may
// write a slot another one reads or writes. Returns TGS_OK or TGS_E_INVALID.
int validateProgram(const toucanGPUSim::SimDesignInfo &d, std::string *err);

// Record and nibble counts; b_prog / b_state follow the F1 record sizes.
void programStats(const toucanGPUSim::SimDesignInfo &d, uint64_t nInputs, tgs_program_stats *s);

// Stimulus: regPool id -> input index (or -1). Empty when stimulus is off.
struct StimMap {
  std::vector<uint32_t> ids;
  std::vector<int32_t> inputOf;  // sized regPoolSize when ids non-empty
  int32_t lookup(uint32_t reg) const { return inputOf.empty() ? -1 : inputOf[reg]; }
};
// Rejects ids out of range, duplicates, and inputs covered by a register write.
int buildStimMap(const toucanGPUSim::SimDesignInfo &d, const uint32_t *ids, uint32_t n, StimMap *m,
                 std::string *err);

struct PlainImage {
  std::vector<uint8_t> net;
  std::vector<PlainPart> parts;
  std::vector<uint32_t> levels;   // pairs {nMParts, mpartBase}
  std::vector<uint32_t> mparts;   // byte offsets within the partition stream
  std::vector<uint8_t> consts;
  uint32_t maxVp = 0, maxRegion = 0;
  uint32_t n0 = 0, nTotal = 0;
};
PlainImage buildPlainImage(const toucanGPUSim::SimDesignInfo &d, const StimMap &stim);

struct Wave64Image {
  std::vector<uint32_t> words;
  std::vector<uint32_t> partOff, partWords;
  std::vector<uint64_t> partCost;
  std::vector<uint8_t> consts;
  uint32_t maxVp = 0;
  uint32_t n0 = 0, nTotal = 0;
  uint32_t fastTasks = 0, rowTasks = 0;  // batched image: LUT pairs as T_LUTFAST / T_LUTROWS
};
// vpRows > 0: the batched kernel's image, LUT pairs as T_LUTROWS tasks for
// waves whose value pool has vpRows rows (level rows follow them); every
// task starts at an even word. fastInst > 0 (instances per wave; only for
// programs that pass nibbleProgram and whose rows times fastInst fit in 16
// bits): pairs whose ops allow it become T_LUTFAST tasks for that count.
Wave64Image buildWave64Image(const toucanGPUSim::SimDesignInfo &d, const StimMap &stim, uint32_t nWaves,
                             uint32_t vpRows = 0, uint32_t fastInst = 0);
// True when no LUT entry, constant-vector nibble or program immediate that
// can become a value exceeds 15, so that every value the program computes
// from nibble-valued pools is a nibble (LUT indices then stay below 0x1000).
bool nibbleProgram(const toucanGPUSim::SimDesignInfo &d);

struct Wave64Schedule {
  uint32_t groups = 0;
  std::vector<uint32_t> off;    // groups*2+1 entries into parts/ldsOff
  std::vector<uint32_t> parts;  // partition ids
  std::vector<uint32_t> ldsOff; // word offset of each scheduled partition in the group's LDS program area
  uint32_t maxGroupWords = 0;   // largest per-group program (words)
  uint64_t maxGroupCost = 0;
};
// Longest-processing-time assignment of partitions to `groups` workgroups per region.
Wave64Schedule scheduleWave64(const Wave64Image &img, uint32_t groups);

// Stream image: 64-word rows plus, per workgroup, the step stream of one
// cycle (partitions assigned per region to workgroup slots by longest
// processing time, the tasks of every exec level assigned to the slot's
// compute waves the same way).
struct StreamImage {
  std::vector<uint32_t> words;
  std::vector<uint32_t> steps;                // step records, 1 + 2 * nC words each
  std::vector<uint32_t> stepBase, stepCount;  // per group: first record, records per cycle
  std::vector<uint8_t> consts;
  uint32_t maxVp = 0, n0 = 0, nTotal = 0, groups = 0, nC = 0, slots = 1, maxSteps = 0;
  uint32_t regionSlots[2] = {1, 1};  // slots used in each region (1 or `slots`)
  uint64_t maxGroupCost = 0, tasks = 0;
  double estCycle = 0;  // estimated cycle time, in LUT exec-level steps
};
// Up to `slots` partitions run side by side in a workgroup, on nC / slots
// compute waves each (the partition-begin, level and partition-end steps of
// the partitions assigned to the slots are merged). With `forceSlots` every
// region is laid out for `slots`; otherwise each region uses `slots` or one
// slot, whichever gives it the shorter estimated critical path.
StreamImage buildStreamImage(const toucanGPUSim::SimDesignInfo &d, const StimMap &stim, uint32_t nC,
                             uint32_t groups, uint32_t slots, bool forceSlots = false);

}  // namespace tgs
