// Serialized program format (GPUSimDesign.bin / GPUSimIOSymbols.bin) and the
// initial-state routines.
//
// Format: src/Serializer.cpp:25-354 of the reference. All integers are raw
// little-endian; a vector is a u64 element count followed by the raw element
// images (Serializer.cpp:39-55); a string is a vector<char> (:58-66). Design
// file order is Serializer.cpp:199-223 (partition :147-166, micropart :106-121),
// symbol file order is :264-300. The order is fixed by the format, so the
// reader/writer below walk it in the same sequence; the code is a bounded
// buffer parser (no per-element istream reads) that zeroes padding on write.
#include "ToucanGPUGenDataTypes.h"

#include <cstring>
#include <iterator>
#include <random>
#include <array>

namespace toucanGPUSim {
namespace {

struct Writer {
  std::vector<uint8_t> &b;
  template <typename T>
  void raw(const T &v) {
    const uint8_t *p = reinterpret_cast<const uint8_t *>(&v);
    b.insert(b.end(), p, p + sizeof(T));
  }
  void u64(uint64_t v) { raw(v); }
  template <typename T>
  void at(size_t base, size_t off, const T &v) {
    std::memcpy(b.data() + base + off, &v, sizeof(T));
  }
  size_t zeros(size_t n) {
    size_t base = b.size();
    b.resize(base + n, 0);
    return base;
  }
};

// Record images with zeroed padding. Offsets are the static_assert'ed ones.
void emit(Writer &w, const CGRegReadMetaInfo &r) {
  size_t o = w.zeros(8); w.at(o, 0, r.reg); w.at(o, 4, r.result);
}
void emit(Writer &w, const CGExchangeReadMetaInfo &r) {
  size_t o = w.zeros(8); w.at(o, 0, r.exchange); w.at(o, 4, r.result);
}
void emit(Writer &w, const CGExchangeWriteMetaInfo &r) {
  size_t o = w.zeros(8); w.at(o, 0, r.exchange); w.at(o, 4, r.dat); w.at(o, 6, r.count);
}
void emit(Writer &w, const CGRegWriteMetaInfo &r) {
  size_t o = w.zeros(8); w.at(o, 0, r.reg); w.at(o, 4, r.dat); w.at(o, 6, r.count);
}
void emit(Writer &w, const CGMemWriteMetaInfo &r) {
  size_t o = w.zeros(24);
  uint8_t m = r.hasMultipleWriter ? 1 : 0;
  w.at(o, 0, m); w.at(o, 8, r.memBase); w.at(o, 16, r.addrVec); w.at(o, 18, r.dat); w.at(o, 20, r.en);
}
void emit(Writer &w, const CGPrintMetaInfo &r) {
  size_t o = w.zeros(4); w.at(o, 0, r.en); w.at(o, 2, r.msg);
}
void emit(Writer &w, const CGStopMetaInfo &r) {
  size_t o = w.zeros(2); w.at(o, 0, r.en);
}
void emit(Writer &w, const CGMicroPartLUTTopLevelOp &r) {
  size_t o = w.zeros(8); w.at(o, 0, r.lutIndex); w.at(o, 2, r.op0); w.at(o, 4, r.op1); w.at(o, 6, r.op2);
}
void emit(Writer &w, const CGMicroPartLUTMiddleLevelOp &r) {
  size_t o = w.zeros(4); w.at(o, 0, r.getPacked());
}
void emit(Writer &w, const CGMicroPartLUTLastLevelWriteBack &r) {
  size_t o = w.zeros(4); w.at(o, 0, r.shuffleId); w.at(o, 2, r.result);
}
void emit(Writer &w, const CGMicroPartVecRead &r) {
  size_t o = w.zeros(20);
  uint8_t c = r.isConstVec ? 1 : 0;
  w.at(o, 0, r.vecBase); w.at(o, 2, r.vecLength); w.at(o, 4, r.offset); w.at(o, 6, c);
  w.at(o, 8, r.index0); w.at(o, 10, r.index1); w.at(o, 12, r.index2); w.at(o, 14, r.index3);
  w.at(o, 16, r.outRangeValue); w.at(o, 18, r.result);
}
void emit(Writer &w, const CGMicroPartVecArithOrLogic &r) {
  size_t o = w.zeros(10);
  w.at(o, 0, r.vec1Base); w.at(o, 2, r.vec2Base); w.at(o, 4, r.vecLength); w.at(o, 6, r.opName);
  w.at(o, 7, r.isV1V2Const); w.at(o, 8, r.result);
}
void emit(Writer &w, const CGMicroPartMemRead &r) {
  size_t o = w.zeros(24);
  uint8_t m = r.hasMultipleWriter ? 1 : 0;
  w.at(o, 0, m); w.at(o, 8, r.memBase); w.at(o, 16, r.en); w.at(o, 18, r.addrVec); w.at(o, 20, r.result);
}
void emit(Writer &w, uint8_t v) { w.raw(v); }
void emit(Writer &w, uint32_t v) { w.raw(v); }
void emit(Writer &w, char v) { w.raw(v); }

template <typename T>
void emitVec(Writer &w, const std::vector<T> &v) {
  w.u64(v.size());
  for (const auto &e : v) emit(w, e);
}

void emitString(Writer &w, const std::string &s) {
  w.u64(s.size());
  w.b.insert(w.b.end(), s.begin(), s.end());
}

struct Reader {
  const uint8_t *p;
  size_t n;
  size_t pos = 0;
  bool ok = true;
  std::string err;
  bool need(size_t k, const char *what) {
    if (!ok) return false;
    if (k > n - pos) {
      ok = false;
      err = std::string("truncated input while reading ") + what + " at byte " + std::to_string(pos);
      return false;
    }
    return true;
  }
  template <typename T>
  bool raw(T &v, const char *what) {
    if (!need(sizeof(T), what)) return false;
    std::memcpy(&v, p + pos, sizeof(T));
    pos += sizeof(T);
    return true;
  }
  bool count(uint64_t &c, size_t elem, const char *what) {
    if (!raw(c, what)) return false;
    // A count whose payload cannot fit in the rest of the input is corrupt;
    // reject it before allocating.
    if (elem != 0 && c > (n - pos) / elem) {
      ok = false;
      err = std::string("element count ") + std::to_string(c) + " of " + what + " exceeds input size";
      return false;
    }
    return true;
  }
  template <typename T>
  bool vec(std::vector<T> &v, const char *what) {
    uint64_t c;
    if (!count(c, sizeof(T), what)) return false;
    v.resize(c);
    if (c) std::memcpy(static_cast<void *>(v.data()), p + pos, c * sizeof(T));
    pos += c * sizeof(T);
    return true;
  }
  bool str(std::string &s, const char *what) {
    uint64_t c;
    if (!count(c, 1, what)) return false;
    s.assign(reinterpret_cast<const char *>(p + pos), c);
    pos += c;
    return true;
  }
};

bool readMicroPart(Reader &r, CGMicroPartInfo &m) {
  uint8_t isLut;
  if (!r.raw(isLut, "micropart kind")) return false;
  m.isLUTPart = isLut != 0;
  if (m.isLUTPart) {
    if (!r.vec(m.topLevel, "top level")) return false;
    uint64_t nMid;
    if (!r.count(nMid, 8, "middle level count")) return false;
    m.middleLevels.resize(nMid);
    for (auto &lv : m.middleLevels)
      if (!r.vec(lv, "middle level")) return false;
    if (!r.vec(m.lastLevel, "last level")) return false;
    m.vecRead.clear();
    m.vecArithAndLogic.clear();
    m.memRead.clear();
  } else {
    if (!r.vec(m.vecRead, "vecRead")) return false;
    if (!r.vec(m.vecArithAndLogic, "vecArithAndLogic")) return false;
    if (!r.vec(m.memRead, "memRead")) return false;
  }
  return true;
}

bool readPartition(Reader &r, SimPartitionInfo &p) {
  if (!r.vec(p.valuePool, "valuePool")) return false;
  if (!r.raw(p.valuePoolSize, "valuePoolSize")) return false;
  if (!r.vec(p.constVecPool, "constVecPool")) return false;
  if (!r.vec(p.ops_l0_regRead, "regRead")) return false;
  if (!r.vec(p.ops_l0_exchangeRead, "exchangeRead")) return false;
  uint64_t nLevels;
  if (!r.count(nLevels, 8, "exec level count")) return false;
  p.exec_mParts.resize(nLevels);
  for (auto &lv : p.exec_mParts) {
    uint64_t nM;
    if (!r.count(nM, 1, "micropart count")) return false;
    lv.resize(nM);
    for (auto &m : lv)
      if (!readMicroPart(r, m)) return false;
  }
  if (!r.raw(p.op_last_regWrite, "regWrite")) return false;
  if (!r.vec(p.ops_last_memWrite, "memWrite")) return false;
  if (!r.vec(p.ops_last_print, "print")) return false;
  if (!r.vec(p.ops_last_stop, "stop")) return false;
  if (!r.raw(p.op_last_exchangeWrite, "exchangeWrite")) return false;
  return true;
}

template <typename Tuple>
bool readSymbolMap(Reader &r, std::unordered_map<std::string, std::vector<Tuple>> &m, const char *what) {
  uint64_t nKeys;
  if (!r.count(nKeys, 16, what)) return false;
  for (uint64_t i = 0; i < nKeys; ++i) {
    std::string key;
    if (!r.str(key, what)) return false;
    uint64_t nT;
    if (!r.count(nT, sizeof(uint32_t) * std::tuple_size<Tuple>::value, what)) return false;
    std::vector<Tuple> v;
    v.reserve(nT);
    for (uint64_t j = 0; j < nT; ++j) {
      uint32_t f[3] = {0, 0, 0};
      for (size_t k = 0; k < std::tuple_size<Tuple>::value; ++k)
        if (!r.raw(f[k], what)) return false;
      if constexpr (std::tuple_size<Tuple>::value == 2)
        v.emplace_back(f[0], f[1]);
      else
        v.emplace_back(f[0], f[1], f[2]);
    }
    m[key] = std::move(v);
  }
  return true;
}

template <typename Tuple>
void writeSymbolMap(Writer &w, const std::unordered_map<std::string, std::vector<Tuple>> &m) {
  w.u64(m.size());
  for (const auto &kv : m) {
    emitString(w, kv.first);
    w.u64(kv.second.size());
    for (const auto &t : kv.second) {
      w.raw(std::get<0>(t));
      w.raw(std::get<1>(t));
      if constexpr (std::tuple_size<Tuple>::value == 3) w.raw(std::get<2>(t));
    }
  }
}

template <typename T>
bool podEq(const T &a, const T &b);

// Field-wise equality (padding excluded).
bool eq(const CGRegReadMetaInfo &a, const CGRegReadMetaInfo &b) { return a.reg == b.reg && a.result == b.result; }
bool eq(const CGExchangeReadMetaInfo &a, const CGExchangeReadMetaInfo &b) {
  return a.exchange == b.exchange && a.result == b.result;
}
bool eq(const CGExchangeWriteMetaInfo &a, const CGExchangeWriteMetaInfo &b) {
  return a.exchange == b.exchange && a.dat == b.dat && a.count == b.count;
}
bool eq(const CGRegWriteMetaInfo &a, const CGRegWriteMetaInfo &b) {
  return a.reg == b.reg && a.dat == b.dat && a.count == b.count;
}
bool eq(const CGMemWriteMetaInfo &a, const CGMemWriteMetaInfo &b) {
  return a.hasMultipleWriter == b.hasMultipleWriter && a.memBase == b.memBase && a.addrVec == b.addrVec &&
         a.dat == b.dat && a.en == b.en;
}
bool eq(const CGPrintMetaInfo &a, const CGPrintMetaInfo &b) { return a.en == b.en && a.msg == b.msg; }
bool eq(const CGStopMetaInfo &a, const CGStopMetaInfo &b) { return a.en == b.en; }
bool eq(const CGMicroPartLUTTopLevelOp &a, const CGMicroPartLUTTopLevelOp &b) {
  return a.lutIndex == b.lutIndex && a.op0 == b.op0 && a.op1 == b.op1 && a.op2 == b.op2;
}
bool eq(const CGMicroPartLUTMiddleLevelOp &a, const CGMicroPartLUTMiddleLevelOp &b) {
  return a.getPacked() == b.getPacked();
}
bool eq(const CGMicroPartLUTLastLevelWriteBack &a, const CGMicroPartLUTLastLevelWriteBack &b) {
  return a.shuffleId == b.shuffleId && a.result == b.result;
}
bool eq(const CGMicroPartVecRead &a, const CGMicroPartVecRead &b) {
  return a.vecBase == b.vecBase && a.vecLength == b.vecLength && a.offset == b.offset &&
         a.isConstVec == b.isConstVec && a.index0 == b.index0 && a.index1 == b.index1 && a.index2 == b.index2 &&
         a.index3 == b.index3 && a.outRangeValue == b.outRangeValue && a.result == b.result;
}
bool eq(const CGMicroPartVecArithOrLogic &a, const CGMicroPartVecArithOrLogic &b) {
  return a.vec1Base == b.vec1Base && a.vec2Base == b.vec2Base && a.vecLength == b.vecLength &&
         a.opName == b.opName && a.isV1V2Const == b.isV1V2Const && a.result == b.result;
}
bool eq(const CGMicroPartMemRead &a, const CGMicroPartMemRead &b) {
  return a.hasMultipleWriter == b.hasMultipleWriter && a.memBase == b.memBase && a.en == b.en &&
         a.addrVec == b.addrVec && a.result == b.result;
}
bool eq(uint8_t a, uint8_t b) { return a == b; }
bool eq(uint32_t a, uint32_t b) { return a == b; }

template <typename T>
bool vecEq(const std::vector<T> &a, const std::vector<T> &b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (!eq(a[i], b[i])) return false;
  return true;
}

bool microPartEq(const CGMicroPartInfo &a, const CGMicroPartInfo &b) {
  if (a.isLUTPart != b.isLUTPart || !vecEq(a.topLevel, b.topLevel) || a.middleLevels.size() != b.middleLevels.size())
    return false;
  for (size_t i = 0; i < a.middleLevels.size(); ++i)
    if (!vecEq(a.middleLevels[i], b.middleLevels[i])) return false;
  return vecEq(a.lastLevel, b.lastLevel) && vecEq(a.vecRead, b.vecRead) &&
         vecEq(a.vecArithAndLogic, b.vecArithAndLogic) && vecEq(a.memRead, b.memRead);
}

}  // namespace

void writeSimDesignInfo(std::vector<uint8_t> &out, const SimDesignInfo &info) {
  Writer w{out};
  w.raw(info.regPoolSize);
  w.raw(info.memPoolSize);
  w.raw(info.exchangePoolSize);
  uint8_t st = info.shouldStop ? 1 : 0;
  w.raw(st);
  emitVec(w, info.lut);
  emitVec(w, info.regPool);
  emitVec(w, info.memPool);
  w.u64(info.parts.size());
  for (const auto &p : info.parts) {
    emitVec(w, p.valuePool);
    w.raw(p.valuePoolSize);
    emitVec(w, p.constVecPool);
    emitVec(w, p.ops_l0_regRead);
    emitVec(w, p.ops_l0_exchangeRead);
    w.u64(p.exec_mParts.size());
    for (const auto &lv : p.exec_mParts) {
      w.u64(lv.size());
      for (const auto &m : lv) {
        uint8_t isLut = m.isLUTPart ? 1 : 0;
        w.raw(isLut);
        if (m.isLUTPart) {
          emitVec(w, m.topLevel);
          w.u64(m.middleLevels.size());
          for (const auto &ml : m.middleLevels) emitVec(w, ml);
          emitVec(w, m.lastLevel);
        } else {
          emitVec(w, m.vecRead);
          emitVec(w, m.vecArithAndLogic);
          emitVec(w, m.memRead);
        }
      }
    }
    emit(w, p.op_last_regWrite);
    emitVec(w, p.ops_last_memWrite);
    emitVec(w, p.ops_last_print);
    emitVec(w, p.ops_last_stop);
    emit(w, p.op_last_exchangeWrite);
  }
  w.u64(info.regionPartitionIds.size());
  for (const auto &r : info.regionPartitionIds) emitVec(w, r);
  w.u64(info.printMsgs.size());
  for (const auto &s : info.printMsgs) emitString(w, s);
}

bool parseSimDesignInfo(const uint8_t *data, size_t size, SimDesignInfo &info, std::string *err) {
  Reader r{data, size};
  uint8_t st = 0;
  r.raw(info.regPoolSize, "regPoolSize");
  r.raw(info.memPoolSize, "memPoolSize");
  r.raw(info.exchangePoolSize, "exchangePoolSize");
  r.raw(st, "shouldStop");
  info.shouldStop = st != 0;
  r.vec(info.lut, "lut");
  r.vec(info.regPool, "regPool");
  r.vec(info.memPool, "memPool");
  uint64_t nParts = 0;
  if (r.count(nParts, 8, "partition count")) {
    info.parts.resize(nParts);
    for (auto &p : info.parts)
      if (!readPartition(r, p)) break;
  }
  uint64_t nRegions = 0;
  if (r.count(nRegions, 8, "region count")) {
    info.regionPartitionIds.resize(nRegions);
    for (auto &reg : info.regionPartitionIds)
      if (!r.vec(reg, "region")) break;
  }
  uint64_t nMsg = 0;
  if (r.count(nMsg, 8, "message count")) {
    info.printMsgs.resize(nMsg);
    for (auto &s : info.printMsgs)
      if (!r.str(s, "message")) break;
  }
  if (!r.ok && err) *err = r.err;
  return r.ok;
}

void writeSimDebugInfo(std::vector<uint8_t> &out, const SimDebugInfo &info) {
  Writer w{out};
  writeSymbolMap(w, info.regDebugInfo);
  writeSymbolMap(w, info.signalDebugInfo);
  writeSymbolMap(w, info.memDebugInfo);
}

bool parseSimDebugInfo(const uint8_t *data, size_t size, SimDebugInfo &info, std::string *err) {
  Reader r{data, size};
  bool ok = readSymbolMap(r, info.regDebugInfo, "register symbols") &&
            readSymbolMap(r, info.signalDebugInfo, "signal symbols") &&
            readSymbolMap(r, info.memDebugInfo, "memory symbols");
  if (!ok && err) *err = r.err;
  return ok;
}

static std::vector<uint8_t> slurp(std::istream &in) {
  return std::vector<uint8_t>(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
}

void serializeSimDesignInfo(std::ostream &out, const SimDesignInfo &info) {
  std::vector<uint8_t> b;
  writeSimDesignInfo(b, info);
  out.write(reinterpret_cast<const char *>(b.data()), static_cast<std::streamsize>(b.size()));
}

void deserializeSimDesignInfo(std::istream &in, SimDesignInfo &info) {
  auto b = slurp(in);
  if (!parseSimDesignInfo(b.data(), b.size(), info, nullptr)) in.setstate(std::ios::failbit);
}

void serializeSimDebugInfo(std::ostream &out, const SimDebugInfo &info) {
  std::vector<uint8_t> b;
  writeSimDebugInfo(b, info);
  out.write(reinterpret_cast<const char *>(b.data()), static_cast<std::streamsize>(b.size()));
}

void deserializeSimDebugInfo(std::istream &in, SimDebugInfo &info) {
  auto b = slurp(in);
  if (!parseSimDebugInfo(b.data(), b.size(), info, nullptr)) in.setstate(std::ios::failbit);
}

bool isSimDesignInfoIdentical(const SimDesignInfo &a, const SimDesignInfo &b) {
  if (a.regPoolSize != b.regPoolSize || a.memPoolSize != b.memPoolSize || a.exchangePoolSize != b.exchangePoolSize ||
      a.shouldStop != b.shouldStop)
    return false;
  if (a.lut != b.lut || a.regPool != b.regPool || a.memPool != b.memPool || a.printMsgs != b.printMsgs) return false;
  if (a.parts.size() != b.parts.size()) return false;
  for (size_t i = 0; i < a.parts.size(); ++i) {
    const auto &pa = a.parts[i];
    const auto &pb = b.parts[i];
    if (pa.valuePool != pb.valuePool || pa.valuePoolSize != pb.valuePoolSize || pa.constVecPool != pb.constVecPool ||
        !vecEq(pa.ops_l0_regRead, pb.ops_l0_regRead) || !vecEq(pa.ops_l0_exchangeRead, pb.ops_l0_exchangeRead))
      return false;
    if (pa.exec_mParts.size() != pb.exec_mParts.size()) return false;
    for (size_t j = 0; j < pa.exec_mParts.size(); ++j) {
      if (pa.exec_mParts[j].size() != pb.exec_mParts[j].size()) return false;
      for (size_t k = 0; k < pa.exec_mParts[j].size(); ++k)
        if (!microPartEq(pa.exec_mParts[j][k], pb.exec_mParts[j][k])) return false;
    }
    if (!eq(pa.op_last_regWrite, pb.op_last_regWrite) || !vecEq(pa.ops_last_memWrite, pb.ops_last_memWrite) ||
        !vecEq(pa.ops_last_print, pb.ops_last_print) || !vecEq(pa.ops_last_stop, pb.ops_last_stop) ||
        !eq(pa.op_last_exchangeWrite, pb.op_last_exchangeWrite))
      return false;
  }
  if (a.regionPartitionIds != b.regionPartitionIds) return false;
  return true;
}

// src/SimInit.cpp:16-26.
void SimDesignInfo::Init() {
  shouldStop = false;
  regPool.resize(regPoolSize, 0);
  memPool.resize(memPoolSize, 0);
  for (auto &p : parts) p.valuePool.resize(p.valuePoolSize);
}

// src/SimInit.cpp:28-64. The nibble table comes from libstdc++'s mt19937 and
// uniform_int_distribution<uint8_t>(0, 15), called exactly as the reference
// calls them, so a given seed gives the reference's table on this toolchain.
void SimDesignInfo::Randomize(uint32_t seed, SimDebugInfo &symbols) {
  constexpr size_t kTable = 1024;
  std::mt19937 gen(seed);
  std::uniform_int_distribution<uint8_t> dist(0, 0xF);
  std::array<uint8_t, kTable> nib;
  for (auto &e : nib) e = dist(gen);
  for (const auto &kv : symbols.regDebugInfo) {
    for (const auto &frag : kv.second) {
      uint32_t id = std::get<0>(frag);
      uint32_t width = std::get<1>(frag);
      int mask = (1 << width) - 1;
      if (id < regPool.size()) regPool[id] = static_cast<uint8_t>(nib[id % kTable] & mask);
    }
  }
  for (const auto &kv : symbols.memDebugInfo) {
    for (const auto &plane : kv.second) {
      uint32_t pos = std::get<0>(plane);
      uint32_t width = std::get<1>(plane);
      uint32_t depth = std::get<2>(plane);
      int mask = (1 << width) - 1;
      for (size_t i = 0; i < depth; ++i) {
        size_t idx = size_t(pos) + i;
        if (idx < memPool.size()) memPool[idx] = static_cast<uint8_t>(nib[idx % kTable] & mask);
      }
    }
  }
}

}  // namespace toucanGPUSim
