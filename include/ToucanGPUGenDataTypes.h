// Program records and containers of the Toucan serialized simulation program.
//
// Drop-in counterpart of the reference's include/ToucanGPUGenDataTypes.h:13-294.
// The record structs are raw-copied to and from GPUSimDesign.bin, so their
// names, field order, sizes and offsets must stay identical to the reference
// (x86-64 SysV layout); every one is pinned by a static_assert below. The
// containers keep the reference's public member names because external
// testbenches read `sim.design` and `sim.symbols` directly.
//
// Host-only header: the MI355X kernels never see these records, they run on
// the re-laid device images built in toucangpurt_amd/csrc/image.cpp.
#pragma once

#include <cassert>
#include <cstddef>
#include <cstdint>
#include <istream>
#include <ostream>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace toucanGPUSim {

// ---- L0 (first level) records: include/ToucanGPUGenDataTypes.h:18-31 ----
struct CGRegReadMetaInfo {
  uint32_t reg;     // nibble index into regPool
  uint16_t result;  // value-pool slot written
};

struct CGExchangeReadMetaInfo {
  uint32_t exchange;  // nibble index into exchangePool
  uint16_t result;
};

struct CGExchangeWriteMetaInfo {
  uint32_t exchange;  // first exchange nibble (copied in 4-nibble words)
  uint16_t dat;       // first value-pool slot
  uint16_t count;     // nibbles; 0 means "no exchange write"
};

// ---- last level records: include/ToucanGPUGenDataTypes.h:34-56 ----
struct CGPrintMetaInfo {
  uint16_t en;   // slot, or immediate when < 16
  uint16_t msg;  // index into SimDesignInfo::printMsgs
};

struct CGStopMetaInfo {
  uint16_t en;
};

struct CGRegWriteMetaInfo {
  uint32_t reg;
  uint16_t dat;
  uint16_t count;  // 0 means "no register write"
};

struct CGMemWriteMetaInfo {
  bool hasMultipleWriter;  // address is scaled by 4 when set
  uint64_t memBase;
  uint16_t addrVec;  // 8 consecutive address-nibble slots
  uint16_t dat;
  uint16_t en;
};

// ---- LUT micropartition records: include/ToucanGPUGenDataTypes.h:61-146 ----
struct CGMicroPartLUTTopLevelOp {
  uint16_t lutIndex;  // 0 = pass op2 through
  uint16_t op0;       // slot, or immediate when < 16
  uint16_t op1;
  uint16_t op2;
};

// 32-bit packed middle-level op: lut<<18 | op0<<12 | op1<<6 | op2.
// Operands 0..31 are immediates, 32..63 name a lane of the previous level.
#pragma pack(push, 1)
struct CGMicroPartLUTMiddleLevelOp {
 private:
  uint32_t packed_data;

 public:
  CGMicroPartLUTMiddleLevelOp() : packed_data(0) {}
  explicit CGMicroPartLUTMiddleLevelOp(uint16_t lutIndex, uint8_t op0, uint8_t op1, uint8_t op2)
      : packed_data(0) {
    setLutIndex(lutIndex);
    setOp0(op0);
    setOp1(op1);
    setOp2(op2);
  }
  uint16_t lutIndex() const { return (packed_data >> 18) & 0x3FFFu; }
  uint8_t op0() const { return (packed_data >> 12) & 0x3Fu; }
  uint8_t op1() const { return (packed_data >> 6) & 0x3Fu; }
  uint8_t op2() const { return packed_data & 0x3Fu; }
  void setLutIndex(uint16_t v) {
    assert(v <= 0x3FFF);
    packed_data = (packed_data & ~(0x3FFFu << 18)) | ((uint32_t(v) & 0x3FFFu) << 18);
  }
  void setOp0(uint8_t v) {
    assert(v <= 0x3F);
    packed_data = (packed_data & ~(0x3Fu << 12)) | ((uint32_t(v) & 0x3Fu) << 12);
  }
  void setOp1(uint8_t v) {
    assert(v <= 0x3F);
    packed_data = (packed_data & ~(0x3Fu << 6)) | ((uint32_t(v) & 0x3Fu) << 6);
  }
  void setOp2(uint8_t v) {
    assert(v <= 0x3F);
    packed_data = (packed_data & ~0x3Fu) | (uint32_t(v) & 0x3Fu);
  }
  uint32_t getPacked() const { return packed_data; }
  void setPacked(uint32_t v) { packed_data = v; }
};
#pragma pack(pop)

struct CGMicroPartLUTLastLevelWriteBack {
  uint8_t shuffleId;  // < 32: immediate; else lane (shuffleId mod 32)
  uint16_t result;
};

struct CGMicroPartVecRead {
  uint16_t vecBase;
  uint16_t vecLength;
  uint16_t offset;
  bool isConstVec;
  uint16_t index0;  // each index: slot, or immediate when < 16
  uint16_t index1;
  uint16_t index2;
  uint16_t index3;
  uint16_t outRangeValue;
  uint16_t result;
};

#define VEC_ARITH_ADD 0
#define VEC_ARITH_SUB 1
#define VEC_ARITH_MUL 2
#define VEC_LOGIC_EQ 3
#define VEC_LOGIC_LT 4
#define VEC_LOGIC_LE 5

struct CGMicroPartVecArithOrLogic {
  uint16_t vec1Base;
  uint16_t vec2Base;
  uint16_t vecLength;   // nibbles
  uint8_t opName;       // VEC_ARITH_* / VEC_LOGIC_*
  uint8_t isV1V2Const;  // bit1: v1 from constVecPool, bit0: v2 from constVecPool
  uint16_t result;
};

struct CGMicroPartMemRead {
  bool hasMultipleWriter;
  uint64_t memBase;
  uint16_t en;
  uint16_t addrVec;
  uint16_t result;
};

static_assert(sizeof(CGRegReadMetaInfo) == 8 && offsetof(CGRegReadMetaInfo, result) == 4, "layout");
static_assert(sizeof(CGExchangeReadMetaInfo) == 8 && offsetof(CGExchangeReadMetaInfo, result) == 4, "layout");
static_assert(sizeof(CGExchangeWriteMetaInfo) == 8 && offsetof(CGExchangeWriteMetaInfo, dat) == 4 &&
                  offsetof(CGExchangeWriteMetaInfo, count) == 6, "layout");
static_assert(sizeof(CGRegWriteMetaInfo) == 8 && offsetof(CGRegWriteMetaInfo, dat) == 4 &&
                  offsetof(CGRegWriteMetaInfo, count) == 6, "layout");
static_assert(sizeof(CGMemWriteMetaInfo) == 24 && offsetof(CGMemWriteMetaInfo, memBase) == 8 &&
                  offsetof(CGMemWriteMetaInfo, addrVec) == 16 && offsetof(CGMemWriteMetaInfo, dat) == 18 &&
                  offsetof(CGMemWriteMetaInfo, en) == 20, "layout");
static_assert(sizeof(CGPrintMetaInfo) == 4 && sizeof(CGStopMetaInfo) == 2, "layout");
static_assert(sizeof(CGMicroPartLUTTopLevelOp) == 8, "layout");
static_assert(sizeof(CGMicroPartLUTMiddleLevelOp) == 4, "layout");
static_assert(sizeof(CGMicroPartLUTLastLevelWriteBack) == 4 &&
                  offsetof(CGMicroPartLUTLastLevelWriteBack, result) == 2, "layout");
static_assert(sizeof(CGMicroPartVecRead) == 20 && offsetof(CGMicroPartVecRead, isConstVec) == 6 &&
                  offsetof(CGMicroPartVecRead, index0) == 8 && offsetof(CGMicroPartVecRead, outRangeValue) == 16 &&
                  offsetof(CGMicroPartVecRead, result) == 18, "layout");
static_assert(sizeof(CGMicroPartVecArithOrLogic) == 10 && offsetof(CGMicroPartVecArithOrLogic, opName) == 6 &&
                  offsetof(CGMicroPartVecArithOrLogic, isV1V2Const) == 7 &&
                  offsetof(CGMicroPartVecArithOrLogic, result) == 8, "layout");
static_assert(sizeof(CGMicroPartMemRead) == 24 && offsetof(CGMicroPartMemRead, memBase) == 8 &&
                  offsetof(CGMicroPartMemRead, en) == 16 && offsetof(CGMicroPartMemRead, addrVec) == 18 &&
                  offsetof(CGMicroPartMemRead, result) == 20, "layout");

// ---- containers: include/ToucanGPUGenDataTypes.h:213-282 ----
struct CGMicroPartInfo {
  bool isLUTPart = false;
  std::vector<CGMicroPartLUTTopLevelOp> topLevel;
  std::vector<std::vector<CGMicroPartLUTMiddleLevelOp>> middleLevels;
  std::vector<CGMicroPartLUTLastLevelWriteBack> lastLevel;
  std::vector<CGMicroPartVecRead> vecRead;
  std::vector<CGMicroPartVecArithOrLogic> vecArithAndLogic;
  std::vector<CGMicroPartMemRead> memRead;
};

struct SimDebugInfo {
  // register name -> fragments (regPool id, bit width), most significant first
  std::unordered_map<std::string, std::vector<std::tuple<uint32_t, uint32_t>>> regDebugInfo;
  // signal name -> fragments (partition, value-pool slot, bit width)
  std::unordered_map<std::string, std::vector<std::tuple<uint32_t, uint32_t, uint32_t>>> signalDebugInfo;
  // memory name -> planes (memPool start, bit width, depth)
  std::unordered_map<std::string, std::vector<std::tuple<uint32_t, uint32_t, uint32_t>>> memDebugInfo;
};

struct SimPartitionInfo {
  std::vector<uint8_t> valuePool;
  uint32_t valuePoolSize = 0;
  std::vector<uint8_t> constVecPool;
  std::vector<CGRegReadMetaInfo> ops_l0_regRead;
  std::vector<CGExchangeReadMetaInfo> ops_l0_exchangeRead;
  std::vector<std::vector<CGMicroPartInfo>> exec_mParts;
  CGRegWriteMetaInfo op_last_regWrite{};
  std::vector<CGMemWriteMetaInfo> ops_last_memWrite;
  std::vector<CGPrintMetaInfo> ops_last_print;
  std::vector<CGStopMetaInfo> ops_last_stop;
  CGExchangeWriteMetaInfo op_last_exchangeWrite{};
};

struct SimDesignInfo {
  std::vector<uint8_t> lut;
  std::vector<uint8_t> regPool;
  std::vector<uint8_t> memPool;
  uint32_t regPoolSize = 0;
  uint32_t memPoolSize = 0;
  uint32_t exchangePoolSize = 0;
  std::vector<SimPartitionInfo> parts;
  std::vector<std::vector<uint32_t>> regionPartitionIds;
  bool shouldStop = false;
  std::vector<std::string> printMsgs;

  // src/SimInit.cpp:16-26: zero register and memory pools, size value pools.
  void Init();
  // src/SimInit.cpp:28-64: seeded nibbles into every register and memory plane
  // named by `symbols`.
  void Randomize(uint32_t seed, SimDebugInfo &symbols);
};

// Stream entry points of src/Serializer.cpp:199-354. Writers emit zeroed
// padding (the reference writes whatever the padding bytes hold); readers take
// the raw records. The deserializers set the stream's failbit on truncation.
void serializeSimDesignInfo(std::ostream &out, const SimDesignInfo &info);
void deserializeSimDesignInfo(std::istream &in, SimDesignInfo &info);
void serializeSimDebugInfo(std::ostream &out, const SimDebugInfo &info);
void deserializeSimDebugInfo(std::istream &in, SimDebugInfo &info);

// src/Serializer.cpp:400-443. Compares field values (padding excluded) and
// returns false on the first difference instead of asserting.
bool isSimDesignInfoIdentical(const SimDesignInfo &a, const SimDesignInfo &b);

// Buffer forms used by the runtime: parse returns false (and an error text)
// on truncated or oversized input instead of reading past the end.
bool parseSimDesignInfo(const uint8_t *data, size_t size, SimDesignInfo &info, std::string *err);
bool parseSimDebugInfo(const uint8_t *data, size_t size, SimDebugInfo &info, std::string *err);
void writeSimDesignInfo(std::vector<uint8_t> &out, const SimDesignInfo &info);
void writeSimDebugInfo(std::vector<uint8_t> &out, const SimDebugInfo &info);

}  // namespace toucanGPUSim
