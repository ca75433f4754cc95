// Seeded synthetic design generator and mini-compiler.
//
// The reference ships only the runtime; its designs come from an external
// LLVM-based compiler that is not part of it. This file builds seeded random
// designs directly in the reference's serialized program form
// (include/ToucanGPUGenDataTypes.h, src/Serializer.cpp) and writes a logical
// netlist beside them so tests can evaluate the design independently of the
// program.
//
// Logical design (all values are nibbles in bytes, as in the reference):
//   * 1-bit state registers, 1-bit primary inputs, 32/64-bit word registers,
//     SRAMs with one read and one write port, LUT4s with uint16 truth tables.
//   * Region 0 ("stage A") computes from the current register state: LUT
//     clusters over state bits and inputs, word ops over word registers, SRAM
//     reads, next SRAM addresses. Its results go to the exchange pool.
//   * Region 1 ("stage B") computes from the exchange pool: LUT clusters that
//     drive the next state bits, word-register next states (nibble muxes),
//     SRAM writes and the SRAM address registers.
//   * A LUT at level L takes each of its 4 inputs from a level drawn uniformly
//     from [0, L) of its own cluster (level 0 = the cluster's sources), so a
//     cluster is a random levelized DAG; clusters are the atoms partitions are
//     packed from, which keeps the logical design independent of part_luts.
//
// LUT4 lowering (a build decision; the reference has no per-LUT truth
// tables, only 3-operand nibble lookups into one 4,642-byte table,
// src/SimEval.cu:33,353): with t = a<<2|b<<1|c,
//   top    : t = PACK3(a,b,c), pass d
//   middle : hi = BSEL8(t, TT[15:12], TT[11:8]), lo = BSEL8(t, TT[7:4], TT[3:0]), pass d
//   middle : y = MUX(d, hi, lo)
//   last   : write y
// so y = TT[d<<3 | t]. Truth-table nibbles ride as middle-level immediates.
#include "ToucanGPUGenDataTypes.h"
#include "toucan_gen.h"
#include "toucan_rt.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

using namespace toucanGPUSim;

namespace {

constexpr uint32_t kLutSize = 4642;
// Table bases inside the design-wide LUT (index 0 is the pass-through op).
constexpr uint16_t T_PACK3 = 1;     // (o0&1)<<2 | (o1&1)<<1 | (o2&1), 274 entries
constexpr uint16_t T_BSEL8 = 275;   // ((o1<<4|o2) >> o0) & 1, o0 in 0..7, 2048 entries
constexpr uint16_t T_MUX = 2323;    // (o0&1) ? o1 : o2, o0 in 0..1, 512 entries
constexpr uint16_t T_XOR = 2835;    // o1 ^ o2, 256 entries
constexpr uint16_t T_AND = 3091;    // o1 & o2, 256 entries (ends at 3347)

constexpr uint32_t NONE = 0xFFFFFFFFu;

enum BitKind : uint32_t { K_REG = 0, K_IN = 1, K_LUT = 2, K_CMP = 3 };
enum WordOp : uint32_t { W_ADD = 0, W_SUB = 1, W_EQ = 2, W_LT = 3, W_SHR = 4, W_MUX = 5, W_RDATA = 6 };

struct BitNode {
  uint32_t kind, a, b, c, d, tt, lvl;
};

struct WordRes {
  uint32_t op, akind, a, bkind, b, nib, sel;  // akind/bkind: 0 = word reg, 1 = word result
  uint32_t part;                                // region-0 word partition (NONE for RDATA)
  uint32_t level;                               // op depth inside its partition (1-based)
  uint32_t cmpNode;                             // bit node for EQ/LT
};

struct WordReg {
  uint32_t nib, sel, x, y;  // next = MUX(sel, x, y); sel NONE means always x
  uint32_t pos;
};

struct Sram {
  uint32_t depth, nib, k;
  uint32_t raMul, raAdd, waMul, waAdd;
  uint32_t wdata, we;  // word result, bit node (NONE = always)
  uint32_t raPos, waPos, memBase, rdataRes;
};

struct Cluster {
  std::vector<uint32_t> sources;
  std::vector<std::vector<uint32_t>> levels;  // LUT node ids per level 1..D
  std::vector<uint32_t> exports;              // region 0 clusters
  uint32_t regBegin = 0, regEnd = 0;          // region 1 clusters: owned state regs
  std::vector<uint32_t> drivers;              // region 1: LUT node driving each owned reg
};

struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t s) : g(s) {}
  uint64_t next() { return g(); }
  uint32_t below(uint64_t n) { return n ? static_cast<uint32_t>(g() % n) : 0; }
};

std::vector<uint8_t> buildLutTable() {
  std::vector<uint8_t> t(kLutSize, 0);
  auto fill = [&](uint16_t base, uint32_t n, auto f) {
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t o0 = i >> 8, o1 = (i >> 4) & 15, o2 = i & 15;
      t[base + i] = static_cast<uint8_t>(f(o0, o1, o2) & 15);
    }
  };
  fill(T_PACK3, 274, [](uint32_t a, uint32_t b, uint32_t c) { return (a & 1) << 2 | (b & 1) << 1 | (c & 1); });
  fill(T_BSEL8, 2048, [](uint32_t s, uint32_t h, uint32_t l) { return ((h << 4 | l) >> s) & 1; });
  fill(T_MUX, 512, [](uint32_t s, uint32_t h, uint32_t l) { return (s & 1) ? h : l; });
  fill(T_XOR, 256, [](uint32_t, uint32_t h, uint32_t l) { return h ^ l; });
  fill(T_AND, 256, [](uint32_t, uint32_t h, uint32_t l) { return h & l; });
  return t;
}

// Keys of values that live in value-pool slots or exchange positions.
inline uint64_t keyBit(uint32_t n) { return (1ull << 32) | n; }
inline uint64_t keyRes(uint32_t r) { return (2ull << 32) | r; }
inline uint64_t keyWReg(uint32_t w) { return (3ull << 32) | w; }
inline uint64_t keyWaCur(uint32_t s) { return (5ull << 32) | s; }
inline uint64_t keyRaNext(uint32_t s) { return (6ull << 32) | s; }
inline uint64_t keyWaNext(uint32_t s) { return (7ull << 32) | s; }

struct Gen;

// Builds one partition: slot allocation, micropartitions per exec level,
// the output block that feeds the last-level word copy.
struct PartBuilder {
  SimPartitionInfo P;
  uint32_t next = 16;  // slots 0..15 read as immediates, never allocate them
  std::unordered_map<uint64_t, uint32_t> slot;
  uint32_t outStart = 0, outCount = 0;
  std::vector<uint32_t> zeroSlots;

  uint32_t alloc(uint32_t n) {
    uint32_t s = next;
    next += n;
    return s;
  }
  void beginOut(uint32_t n) {
    next = (next + 3) & ~3u;
    outStart = next;
    outCount = (n + 3) & ~3u;
    for (uint32_t i = n; i < outCount; ++i) zeroSlots.push_back(outStart + i);
    next += outCount;
  }
  std::vector<CGMicroPartInfo> &level(uint32_t e) {
    if (P.exec_mParts.size() <= e) P.exec_mParts.resize(e + 1);
    return P.exec_mParts[e];
  }
  void finish() {
    // Zero-fill writes for the padding of the output block.
    for (size_t i = 0; i < zeroSlots.size(); i += 32) {
      CGMicroPartInfo m;
      m.isLUTPart = true;
      for (size_t j = i; j < std::min(zeroSlots.size(), i + 32); ++j) {
        CGMicroPartLUTLastLevelWriteBack w{};
        w.shuffleId = 0;
        w.result = static_cast<uint16_t>(zeroSlots[j]);
        m.lastLevel.push_back(w);
      }
      level(0).push_back(std::move(m));
    }
    if (next > 65535) throw std::runtime_error("partition value pool exceeds 65535 slots");
    P.valuePoolSize = (next + 3) & ~3u;
  }
};

CGMicroPartLUTTopLevelOp top(uint16_t lut, uint16_t a, uint16_t b, uint16_t c) {
  CGMicroPartLUTTopLevelOp o{};
  o.lutIndex = lut;
  o.op0 = a;
  o.op1 = b;
  o.op2 = c;
  return o;
}

CGMicroPartLUTLastLevelWriteBack last(uint8_t sid, uint32_t slot) {
  CGMicroPartLUTLastLevelWriteBack w{};
  w.shuffleId = sid;
  w.result = static_cast<uint16_t>(slot);
  return w;
}

struct Gen {
  tgs_gen_params p;
  Rng rng;
  std::vector<BitNode> bits;
  std::vector<uint32_t> regNode, inNode;  // bit node per state reg / input
  std::vector<WordRes> wres;
  std::vector<WordReg> wregs;
  std::vector<Sram> srams;
  std::vector<Cluster> clA, clB;
  std::vector<uint32_t> regPos, inPos;  // regPool position of state regs / inputs
  std::vector<uint32_t> regNext;        // driver node per state reg
  std::vector<uint32_t> aExports;       // bit nodes available to region 1
  std::vector<uint32_t> cmpNodes;
  uint32_t maxLevelA = 0;

  SimDesignInfo D;
  SimDebugInfo S;
  std::unordered_map<uint64_t, uint32_t> xpos;  // exchange position of exported values
  uint64_t nLutOps = 0, nMicroparts = 0, nVecOpsEmitted = 0, nLutsEmitted = 0;

  // Partition plans.
  struct WaPart { std::vector<uint32_t> regs; std::vector<uint32_t> ops; };
  struct WbPart { std::vector<uint32_t> regs; };
  std::vector<WaPart> waParts;
  std::vector<WbPart> wbParts;
  std::vector<std::vector<uint32_t>> lutPartsA, lutPartsB;  // cluster ids
  std::vector<std::vector<uint32_t>> sramGroups;
  std::vector<uint32_t> stopPrintPart;  // B-LUT partition index per print/stop

  explicit Gen(const tgs_gen_params &pp) : p(pp), rng(pp.seed * 0x9E3779B97F4A7C15ull + 0x1234567ull) {}

  uint32_t newBit(uint32_t kind, uint32_t a = 0) {
    bits.push_back(BitNode{kind, a, 0, 0, 0, 0, 0});
    return static_cast<uint32_t>(bits.size() - 1);
  }

  void makeClusters(std::vector<Cluster> &cls, uint64_t nLuts, uint32_t depth, bool stageA) {
    if (nLuts == 0) return;
    uint64_t csize = std::max<uint32_t>(p.cluster_luts, depth);
    uint64_t n = std::max<uint64_t>(1, (nLuts + csize / 2) / csize);
    cls.resize(n);
    for (uint64_t c = 0; c < n; ++c) {
      uint64_t lo = nLuts * c / n, hi = nLuts * (c + 1) / n;
      uint64_t cnt = hi - lo;
      if (cnt < depth) throw std::runtime_error("cluster smaller than its depth");
      auto &cl = cls[c];
      cl.levels.resize(depth);
      for (uint32_t l = 0; l < depth; ++l) {
        uint64_t a = cnt * l / depth, b = cnt * (l + 1) / depth;
        cl.levels[l].resize(b - a);
      }
    }
    (void)stageA;
  }

  void fillCluster(Cluster &cl, uint32_t lvlBase) {
    for (uint32_t l = 0; l < cl.levels.size(); ++l) {
      for (auto &id : cl.levels[l]) {
        BitNode n{K_LUT, 0, 0, 0, 0, static_cast<uint32_t>(rng.next() & 0xFFFF), lvlBase + l + 1};
        uint32_t in[4];
        for (auto &x : in) {
          uint32_t src = rng.below(l + 1);  // 0 = sources, else level src-1
          const auto &pool = src == 0 ? cl.sources : cl.levels[src - 1];
          x = pool[rng.below(pool.size())];
        }
        n.a = in[0];
        n.b = in[1];
        n.c = in[2];
        n.d = in[3];
        bits.push_back(n);
        id = static_cast<uint32_t>(bits.size() - 1);
        ++nLutsEmitted;
      }
    }
  }

  void plan() {
    if (p.depth < 2) throw std::runtime_error("depth must be >= 2 (one LUT level per region at least)");
    if (p.n_luts && p.n_regs > p.n_luts / 2) throw std::runtime_error("n_regs exceeds the region-1 LUT count");
    if (p.n_srams && (p.sram_depth == 0 || (p.sram_depth & (p.sram_depth - 1)) || p.sram_depth > 65536 ||
                      p.sram_width == 0 || p.sram_width % 4 || p.sram_width > 128))
      throw std::runtime_error("bad SRAM geometry");
    for (uint32_t i = 0; i < p.n_regs; ++i) regNode.push_back(newBit(K_REG, i));
    for (uint32_t i = 0; i < p.n_inputs; ++i) inNode.push_back(newBit(K_IN, i));

    // ---- word registers and region-0 word ops ----
    uint32_t nW = p.n_vecops ? std::max<uint32_t>(2, p.n_vecops / 4) : 0;
    uint32_t nOpsA = p.n_vecops > nW ? p.n_vecops - nW : 0;
    if (nW) {
      uint32_t opsPer = std::max<uint32_t>(8, p.word_part_ops);
      uint32_t nParts = std::max<uint32_t>(1, (nOpsA + opsPer - 1) / opsPer);
      waParts.resize(nParts);
      // Homogeneous-width word-register groups, one per region-0 word partition.
      for (uint32_t q = 0; q < nParts; ++q) {
        uint32_t lo = uint64_t(nW) * q / nParts, hi = uint64_t(nW) * (q + 1) / nParts;
        bool wide = (uint64_t(q) * 100 + 50) / nParts < p.word_bits_mix;
        for (uint32_t w = lo; w < hi; ++w) {
          wregs.push_back(WordReg{wide ? 16u : 8u, NONE, NONE, NONE, 0});
          waParts[q].regs.push_back(w);
        }
      }
      for (uint32_t q = 0; q < nParts; ++q) {
        auto &wp = waParts[q];
        if (wp.regs.empty()) continue;
        uint32_t lo = uint64_t(nOpsA) * q / nParts, hi = uint64_t(nOpsA) * (q + 1) / nParts;
        std::vector<uint32_t> wordResults;  // results of this partition usable as operands
        uint32_t nib = wregs[wp.regs[0]].nib;
        for (uint32_t i = lo; i < hi; ++i) {
          WordRes r{};
          r.op = static_cast<uint32_t>(rng.below(6));
          r.nib = nib;
          r.part = q;
          r.sel = NONE;
          r.cmpNode = NONE;
          auto pick = [&](uint32_t &kind, uint32_t &id, uint32_t &lvl) {
            if (!wordResults.empty() && rng.below(3) == 0) {
              uint32_t rr = wordResults[rng.below(wordResults.size())];
              if (wres[rr].level < 4) {
                kind = 1;
                id = rr;
                lvl = wres[rr].level;
                return;
              }
            }
            kind = 0;
            id = wp.regs[rng.below(wp.regs.size())];
            lvl = 0;
          };
          uint32_t la = 0, lb = 0;
          pick(r.akind, r.a, la);
          pick(r.bkind, r.b, lb);
          if (r.op == W_MUX) {
            // select from a state bit (region 0 reads registers)
            r.sel = p.n_regs ? regNode[rng.below(p.n_regs)] : NONE;
            if (r.sel == NONE) r.op = W_ADD;
          }
          if (r.op == W_SHR && r.bkind != 0) {
            // shift amount: low nibble of a word register
            r.bkind = 0;
            r.b = wp.regs[rng.below(wp.regs.size())];
            lb = 0;
          }
          r.level = std::max(la, lb) + 1;
          wres.push_back(r);
          uint32_t id = static_cast<uint32_t>(wres.size() - 1);
          wp.ops.push_back(id);
          if (r.op == W_EQ || r.op == W_LT) {
            wres[id].cmpNode = newBit(K_CMP, id);
            cmpNodes.push_back(wres[id].cmpNode);
          } else {
            wordResults.push_back(id);
          }
          ++nVecOpsEmitted;
        }
      }
    }

    // ---- SRAMs ----
    for (uint32_t s = 0; s < p.n_srams; ++s) {
      Sram m{};
      m.depth = p.sram_depth;
      m.nib = p.sram_width / 4;
      uint32_t bitsA = 0;
      while ((1u << bitsA) < m.depth) ++bitsA;
      m.k = std::max<uint32_t>(1, (bitsA + 3) / 4);
      m.raMul = (static_cast<uint32_t>(rng.next()) | 1u) & 0xFFFFu;
      m.raMul = (m.raMul & ~3u) | 1u;  // a = 1 mod 4: full-period LCG modulo 16^k
      m.raAdd = (static_cast<uint32_t>(rng.next()) | 1u) & 0xFFFFu;
      m.waMul = ((static_cast<uint32_t>(rng.next()) & 0xFFFFu) & ~3u) | 1u;
      m.waAdd = (static_cast<uint32_t>(rng.next()) | 1u) & 0xFFFFu;
      m.we = NONE;
      WordRes r{};
      r.op = W_RDATA;
      r.a = s;
      r.nib = m.nib;
      r.part = NONE;
      r.sel = NONE;
      r.cmpNode = NONE;
      r.level = 2;
      wres.push_back(r);
      m.rdataRes = static_cast<uint32_t>(wres.size() - 1);
      srams.push_back(m);
    }
    for (uint32_t s = 0; s < p.n_srams; s += 8) {
      std::vector<uint32_t> g;
      for (uint32_t t = s; t < std::min(p.n_srams, s + 8); ++t) g.push_back(t);
      sramGroups.push_back(g);
    }

    // ---- LUT clusters ----
    uint64_t nA = (uint64_t(p.n_luts) + 1) / 2, nB = p.n_luts - nA;
    uint32_t dA = (p.depth + 1) / 2, dB = p.depth - dA;
    makeClusters(clA, nA, dA, true);
    makeClusters(clB, nB, dB, false);
    maxLevelA = dA;
    if (!clA.empty()) {
      std::vector<uint32_t> srcs;
      for (auto n : regNode) srcs.push_back(n);
      for (auto n : inNode) srcs.push_back(n);
      std::vector<uint32_t> order(srcs);
      for (size_t i = order.size(); i > 1; --i) std::swap(order[i - 1], order[rng.below(i)]);
      for (size_t i = 0; i < order.size(); ++i) clA[i % clA.size()].sources.push_back(order[i]);
      for (auto &cl : clA) {
        while (cl.sources.size() < p.sources_a && !srcs.empty()) cl.sources.push_back(srcs[rng.below(srcs.size())]);
        if (cl.sources.empty()) throw std::runtime_error("region-0 cluster without sources (no registers or inputs)");
        fillCluster(cl, 0);
        cl.exports = cl.levels.back();
        for (auto n : cl.exports) aExports.push_back(n);
      }
    }
    if (!clB.empty()) {
      std::vector<uint32_t> srcs(aExports);
      for (auto n : cmpNodes) srcs.push_back(n);
      if (srcs.empty()) throw std::runtime_error("region-1 cluster without sources");
      // state registers: contiguous blocks per cluster
      for (size_t c = 0; c < clB.size(); ++c) {
        auto &cl = clB[c];
        cl.regBegin = static_cast<uint32_t>(uint64_t(p.n_regs) * c / clB.size());
        cl.regEnd = static_cast<uint32_t>(uint64_t(p.n_regs) * (c + 1) / clB.size());
        while (cl.sources.size() < std::max<uint32_t>(p.sources_b, 4)) cl.sources.push_back(srcs[rng.below(srcs.size())]);
        fillCluster(cl, dA);
        std::vector<uint32_t> cand;
        for (size_t l = cl.levels.size(); l-- > 0;)
          for (auto n : cl.levels[l]) cand.push_back(n);
        if (cand.size() < cl.regEnd - cl.regBegin) throw std::runtime_error("cluster owns more registers than LUTs");
        for (uint32_t r = cl.regBegin; r < cl.regEnd; ++r) cl.drivers.push_back(cand[r - cl.regBegin]);
      }
    }
    regNext.assign(p.n_regs, NONE);
    for (auto &cl : clB)
      for (uint32_t r = cl.regBegin; r < cl.regEnd; ++r) regNext[r] = cl.drivers[r - cl.regBegin];
    for (uint32_t r = 0; r < p.n_regs; ++r)
      if (regNext[r] == NONE) throw std::runtime_error("state register without a driver");

    // ---- word-register next states and SRAM write ports ----
    std::map<uint32_t, std::vector<uint32_t>> resByNib;
    for (uint32_t i = 0; i < wres.size(); ++i)
      if (wres[i].op != W_EQ && wres[i].op != W_LT) resByNib[wres[i].nib].push_back(i);
    for (auto &w : wregs) {
      auto &v = resByNib[w.nib];
      if (v.empty()) throw std::runtime_error("no word result of matching width");
      w.x = v[rng.below(v.size())];
      w.y = v[rng.below(v.size())];
      w.sel = aExports.empty() ? NONE : aExports[rng.below(aExports.size())];
      ++nVecOpsEmitted;
    }
    for (auto &m : srams) {
      auto &v = resByNib[m.nib];
      m.wdata = v[rng.below(v.size())];
      m.we = aExports.empty() ? NONE : aExports[rng.below(aExports.size())];
    }
    if (!wregs.empty()) {
      uint32_t per = std::max<uint32_t>(8, p.word_part_ops / 2);
      for (uint32_t w = 0; w < wregs.size(); w += per) {
        WbPart b;
        for (uint32_t k = w; k < std::min<uint32_t>(wregs.size(), w + per); ++k) b.regs.push_back(k);
        wbParts.push_back(b);
      }
    }

    // ---- pack clusters into partitions ----
    auto pack = [&](std::vector<Cluster> &cls, std::vector<std::vector<uint32_t>> &parts) {
      uint64_t acc = 0;
      for (uint32_t c = 0; c < cls.size(); ++c) {
        uint64_t sz = 0;
        for (auto &l : cls[c].levels) sz += l.size();
        if (parts.empty() || acc + sz > std::max<uint32_t>(p.part_luts, 1) + sz / 2) {
          parts.emplace_back();
          acc = 0;
        }
        parts.back().push_back(c);
        acc += sz;
      }
    };
    pack(clA, lutPartsA);
    pack(clB, lutPartsB);
    for (uint32_t i = 0; i < p.n_prints + p.n_stops; ++i)
      stopPrintPart.push_back(lutPartsB.empty() ? NONE : rng.below(lutPartsB.size()));
  }

  // ---------------- layout ----------------
  // regPool: [inputs][per region-1 partition owned registers] (4-aligned blocks)
  uint32_t regTop = 0;
  std::vector<uint32_t> lutBRegBase, wbRegBase, swRegBase;

  void layoutRegs() {
    regPos.assign(p.n_regs, NONE);
    inPos.assign(p.n_inputs, NONE);
    uint32_t at = 0;
    for (uint32_t i = 0; i < p.n_inputs; ++i) inPos[i] = at++;
    at = (at + 3) & ~3u;
    for (auto &part : lutPartsB) {
      lutBRegBase.push_back(at);
      for (auto c : part)
        for (uint32_t r = clB[c].regBegin; r < clB[c].regEnd; ++r) regPos[r] = at++;
      at = (at + 3) & ~3u;
    }
    for (auto &wb : wbParts) {
      wbRegBase.push_back(at);
      for (auto w : wb.regs) {
        wregs[w].pos = at;
        at += wregs[w].nib;
      }
      at = (at + 3) & ~3u;
    }
    for (auto &g : sramGroups) {
      swRegBase.push_back(at);
      for (auto s : g) {
        srams[s].raPos = at;
        at += srams[s].k;
        srams[s].waPos = at;
        at += srams[s].k;
      }
      at = (at + 3) & ~3u;
    }
    regTop = at;
    uint32_t m = 0;
    for (auto &s : srams) {
      s.memBase = m;
      m += s.depth * s.nib;
    }
    D.regPoolSize = regTop;
    D.memPoolSize = m;
  }

  // ---------------- micropart emitters ----------------
  void emitLut4Level(PartBuilder &B, uint32_t e, const std::vector<uint32_t> &luts) {
    for (size_t i = 0; i < luts.size(); i += 10) {
      CGMicroPartInfo m;
      m.isLUTPart = true;
      size_t K = std::min<size_t>(10, luts.size() - i);
      std::vector<CGMicroPartLUTMiddleLevelOp> m0, m1;
      for (size_t k = 0; k < K; ++k) {
        const auto &n = bits[luts[i + k]];
        uint16_t sa = static_cast<uint16_t>(B.slot.at(keyBit(n.a)));
        uint16_t sb = static_cast<uint16_t>(B.slot.at(keyBit(n.b)));
        uint16_t sc = static_cast<uint16_t>(B.slot.at(keyBit(n.c)));
        uint16_t sd = static_cast<uint16_t>(B.slot.at(keyBit(n.d)));
        uint8_t t = static_cast<uint8_t>(32 + 2 * k), dl = static_cast<uint8_t>(32 + 2 * k + 1);
        m.topLevel.push_back(top(T_PACK3, sa, sb, sc));
        m.topLevel.push_back(top(0, 0, 0, sd));
        m0.emplace_back(T_BSEL8, t, (n.tt >> 12) & 15, (n.tt >> 8) & 15);
        m0.emplace_back(T_BSEL8, t, (n.tt >> 4) & 15, n.tt & 15);
        m0.emplace_back(0, 0, 0, dl);
        m1.emplace_back(T_MUX, static_cast<uint8_t>(32 + 3 * k + 2), static_cast<uint8_t>(32 + 3 * k),
                        static_cast<uint8_t>(32 + 3 * k + 1));
        m.lastLevel.push_back(last(static_cast<uint8_t>(32 + k), B.slot.at(keyBit(luts[i + k]))));
      }
      nLutOps += m.topLevel.size() + m0.size() + m1.size();
      m.middleLevels.push_back(std::move(m0));
      m.middleLevels.push_back(std::move(m1));
      B.level(e).push_back(std::move(m));
      ++nMicroparts;
    }
  }

  // Top-level-only LUT micropartitions: one op per lane, each written to a slot.
  struct TopOp {
    uint16_t lut, a, b, c;
    uint32_t out;
  };
  void emitTopOps(PartBuilder &B, uint32_t e, const std::vector<TopOp> &ops) {
    for (size_t i = 0; i < ops.size(); i += 32) {
      CGMicroPartInfo m;
      m.isLUTPart = true;
      for (size_t j = i; j < std::min(ops.size(), i + 32); ++j) {
        m.topLevel.push_back(top(ops[j].lut, ops[j].a, ops[j].b, ops[j].c));
        m.lastLevel.push_back(last(static_cast<uint8_t>(32 + (j - i)), ops[j].out));
      }
      nLutOps += m.topLevel.size();
      B.level(e).push_back(std::move(m));
      ++nMicroparts;
    }
  }

  template <typename T, typename F>
  void emitChunks(PartBuilder &B, uint32_t e, const std::vector<T> &ops, F put) {
    for (size_t i = 0; i < ops.size(); i += 32) {
      CGMicroPartInfo m;
      m.isLUTPart = false;
      for (size_t j = i; j < std::min(ops.size(), i + 32); ++j) put(m, ops[j]);
      B.level(e).push_back(std::move(m));
      ++nMicroparts;
    }
  }

  void addrFormer(PartBuilder &B, uint32_t e, const Sram &m, uint32_t firstSlot, uint32_t av,
                  std::vector<TopOp> &ops) {
    for (uint32_t j = 0; j < 8; ++j) {
      TopOp o{0, 0, 0, 0, av + j};
      if (j < m.k) {
        uint16_t s = static_cast<uint16_t>(firstSlot + j);
        uint32_t bitsUsed = 0;
        while ((1u << bitsUsed) < m.depth) ++bitsUsed;
        uint32_t topBits = bitsUsed - 4 * (m.k - 1);
        if (j == m.k - 1 && topBits < 4)
          o = TopOp{T_AND, 0, s, static_cast<uint16_t>((1u << topBits) - 1), av + j};
        else
          o = TopOp{0, 0, 0, s, av + j};
      }
      ops.push_back(o);
    }
    (void)B;
    (void)e;
  }

  // ---------------- region 0 partitions ----------------
  void exportBlock(PartBuilder &B, const std::vector<std::pair<uint64_t, uint32_t>> &items, uint32_t &xat) {
    // items: (key, nibbles); assigns output-block slots and exchange positions
    uint32_t n = 0;
    for (auto &it : items) n += it.second;
    B.beginOut(n);
    uint32_t s = B.outStart;
    xat = (xat + 3) & ~3u;
    uint32_t xbase = xat;
    for (auto &it : items) {
      B.slot[it.first] = s;
      xpos[it.first] = xbase + (s - B.outStart);
      s += it.second;
    }
    xat = xbase + B.outCount;
    CGExchangeWriteMetaInfo w{};
    w.exchange = xbase;
    w.dat = static_cast<uint16_t>(B.outStart);
    w.count = static_cast<uint16_t>(B.outCount);
    if (B.outCount > 65535) throw std::runtime_error("exchange block too large");
    B.P.op_last_exchangeWrite = w;
  }

  uint32_t regPosOf(uint32_t node) const {
    const auto &n = bits[node];
    return n.kind == K_REG ? regPos[n.a] : inPos[n.a];
  }

  void compileLutA(PartBuilder &B, const std::vector<uint32_t> &cls, uint32_t &xat) {
    std::vector<std::pair<uint64_t, uint32_t>> ex;
    for (auto c : cls)
      for (auto n : clA[c].exports) ex.emplace_back(keyBit(n), 1);
    exportBlock(B, ex, xat);
    for (auto c : cls)
      for (auto n : clA[c].sources)
        if (!B.slot.count(keyBit(n))) {
          uint32_t s = B.alloc(1);
          B.slot[keyBit(n)] = s;
          B.P.ops_l0_regRead.push_back(CGRegReadMetaInfo{regPosOf(n), static_cast<uint16_t>(s)});
        }
    for (auto c : cls)
      for (auto &lv : clA[c].levels)
        for (auto n : lv)
          if (!B.slot.count(keyBit(n))) B.slot[keyBit(n)] = B.alloc(1);
    uint32_t depth = 0;
    for (auto c : cls) depth = std::max<uint32_t>(depth, clA[c].levels.size());
    for (uint32_t l = 0; l < depth; ++l) {
      std::vector<uint32_t> luts;
      for (auto c : cls)
        if (l < clA[c].levels.size())
          for (auto n : clA[c].levels[l]) luts.push_back(n);
      emitLut4Level(B, l, luts);
    }
  }

  void compileWordA(PartBuilder &B, const WaPart &wp, uint32_t &xat) {
    std::vector<std::pair<uint64_t, uint32_t>> ex;
    for (auto r : wp.ops) ex.emplace_back(keyRes(r), (wres[r].op == W_EQ || wres[r].op == W_LT) ? 1 : wres[r].nib);
    exportBlock(B, ex, xat);
    for (auto r : wp.ops)
      if (wres[r].cmpNode != NONE) xpos[keyBit(wres[r].cmpNode)] = xpos.at(keyRes(r));
    for (auto w : wp.regs) {
      uint32_t s = B.alloc(wregs[w].nib);
      B.slot[keyWReg(w)] = s;
      for (uint32_t j = 0; j < wregs[w].nib; ++j)
        B.P.ops_l0_regRead.push_back(CGRegReadMetaInfo{wregs[w].pos + j, static_cast<uint16_t>(s + j)});
    }
    for (auto r : wp.ops)
      if (wres[r].op == W_MUX && !B.slot.count(keyBit(wres[r].sel))) {
        uint32_t s = B.alloc(1);
        B.slot[keyBit(wres[r].sel)] = s;
        B.P.ops_l0_regRead.push_back(CGRegReadMetaInfo{regPosOf(wres[r].sel), static_cast<uint16_t>(s)});
      }
    uint32_t maxL = 0;
    for (auto r : wp.ops) maxL = std::max(maxL, wres[r].level);
    for (uint32_t L = 1; L <= maxL; ++L) {
      std::vector<CGMicroPartVecArithOrLogic> vops;
      std::vector<CGMicroPartVecRead> vreads;
      std::vector<TopOp> muxes;
      for (auto r : wp.ops) {
        const auto &o = wres[r];
        if (o.level != L) continue;
        uint32_t sa = B.slot.at(o.akind ? keyRes(o.a) : keyWReg(o.a));
        uint32_t sb = B.slot.at(o.bkind ? keyRes(o.b) : keyWReg(o.b));
        uint32_t out = B.slot.at(keyRes(r));
        if (o.op == W_ADD || o.op == W_SUB || o.op == W_EQ || o.op == W_LT) {
          CGMicroPartVecArithOrLogic v{};
          v.vec1Base = static_cast<uint16_t>(sa);
          v.vec2Base = static_cast<uint16_t>(sb);
          v.vecLength = static_cast<uint16_t>(o.nib);
          v.opName = o.op == W_ADD ? VEC_ARITH_ADD : o.op == W_SUB ? VEC_ARITH_SUB : o.op == W_EQ ? VEC_LOGIC_EQ : VEC_LOGIC_LT;
          v.isV1V2Const = 0;
          v.result = static_cast<uint16_t>(out);
          vops.push_back(v);
        } else if (o.op == W_SHR) {
          for (uint32_t j = 0; j < o.nib; ++j) {
            CGMicroPartVecRead v{};
            v.vecBase = static_cast<uint16_t>(sa);
            v.vecLength = static_cast<uint16_t>(o.nib);
            v.offset = static_cast<uint16_t>(j);
            v.isConstVec = false;
            v.index0 = v.index1 = v.index2 = 0;
            v.index3 = static_cast<uint16_t>(sb);  // low nibble of the amount register
            v.outRangeValue = 0;
            v.result = static_cast<uint16_t>(out + j);
            vreads.push_back(v);
          }
        } else {  // W_MUX
          uint16_t ss = static_cast<uint16_t>(B.slot.at(keyBit(o.sel)));
          for (uint32_t j = 0; j < o.nib; ++j)
            muxes.push_back(TopOp{T_MUX, ss, static_cast<uint16_t>(sa + j), static_cast<uint16_t>(sb + j), out + j});
        }
      }
      emitChunks(B, L - 1, vops, [](CGMicroPartInfo &m, const CGMicroPartVecArithOrLogic &v) { m.vecArithAndLogic.push_back(v); });
      emitChunks(B, L - 1, vreads, [](CGMicroPartInfo &m, const CGMicroPartVecRead &v) { m.vecRead.push_back(v); });
      emitTopOps(B, L - 1, muxes);
    }
  }

  void compileSramRead(PartBuilder &B, const std::vector<uint32_t> &g, uint32_t &xat) {
    std::vector<std::pair<uint64_t, uint32_t>> ex;
    for (auto s : g) {
      const auto &m = srams[s];
      ex.emplace_back(keyRes(m.rdataRes), m.nib);
      ex.emplace_back(keyWaCur(s), m.k);
      ex.emplace_back(keyRaNext(s), m.k);
      ex.emplace_back(keyWaNext(s), m.k);
    }
    exportBlock(B, ex, xat);
    std::vector<TopOp> former;
    std::vector<CGMicroPartVecArithOrLogic> muls, adds;
    std::vector<CGMicroPartMemRead> reads;
    for (auto s : g) {
      const auto &m = srams[s];
      uint32_t ra = B.alloc(m.k);
      for (uint32_t j = 0; j < m.k; ++j)
        B.P.ops_l0_regRead.push_back(CGRegReadMetaInfo{m.raPos + j, static_cast<uint16_t>(ra + j)});
      uint32_t wa = B.slot.at(keyWaCur(s));
      for (uint32_t j = 0; j < m.k; ++j)
        B.P.ops_l0_regRead.push_back(CGRegReadMetaInfo{m.waPos + j, static_cast<uint16_t>(wa + j)});
      uint32_t av = B.alloc(8);
      addrFormer(B, 0, m, ra, av, former);
      // constants: [raMul][raAdd][waMul][waAdd], k nibbles each
      uint32_t cbase = static_cast<uint32_t>(B.P.constVecPool.size());
      for (uint32_t c : {m.raMul, m.raAdd, m.waMul, m.waAdd})
        for (uint32_t j = 0; j < m.k; ++j) B.P.constVecPool.push_back(static_cast<uint8_t>((c >> (4 * j)) & 15));
      uint32_t tRa = B.alloc(m.k), tWa = B.alloc(m.k);
      auto vop = [&](uint32_t v1, uint32_t c, uint8_t op, uint32_t out) {
        CGMicroPartVecArithOrLogic v{};
        v.vec1Base = static_cast<uint16_t>(v1);
        v.vec2Base = static_cast<uint16_t>(c);
        v.vecLength = static_cast<uint16_t>(m.k);
        v.opName = op;
        v.isV1V2Const = 1;
        v.result = static_cast<uint16_t>(out);
        return v;
      };
      muls.push_back(vop(ra, cbase + 0 * m.k, VEC_ARITH_MUL, tRa));
      muls.push_back(vop(wa, cbase + 2 * m.k, VEC_ARITH_MUL, tWa));
      adds.push_back(vop(tRa, cbase + 1 * m.k, VEC_ARITH_ADD, B.slot.at(keyRaNext(s))));
      adds.push_back(vop(tWa, cbase + 3 * m.k, VEC_ARITH_ADD, B.slot.at(keyWaNext(s))));
      uint32_t rd = B.slot.at(keyRes(m.rdataRes));
      for (uint32_t j = 0; j < m.nib; ++j) {
        CGMicroPartMemRead r{};
        r.hasMultipleWriter = false;
        r.memBase = m.memBase + uint64_t(j) * m.depth;
        r.en = 1;
        r.addrVec = static_cast<uint16_t>(av);
        r.result = static_cast<uint16_t>(rd + j);
        reads.push_back(r);
      }
    }
    emitTopOps(B, 0, former);
    emitChunks(B, 0, muls, [](CGMicroPartInfo &mp, const CGMicroPartVecArithOrLogic &v) { mp.vecArithAndLogic.push_back(v); });
    emitChunks(B, 1, adds, [](CGMicroPartInfo &mp, const CGMicroPartVecArithOrLogic &v) { mp.vecArithAndLogic.push_back(v); });
    emitChunks(B, 1, reads, [](CGMicroPartInfo &mp, const CGMicroPartMemRead &v) { mp.memRead.push_back(v); });
    nVecOpsEmitted += 0;
  }

  // ---------------- region 1 partitions ----------------
  uint32_t exgSlot(PartBuilder &B, uint64_t key, uint32_t nib) {
    auto it = B.slot.find(key);
    if (it != B.slot.end()) return it->second;
    uint32_t s = B.alloc(nib);
    B.slot[key] = s;
    uint32_t x = xpos.at(key);
    for (uint32_t j = 0; j < nib; ++j)
      B.P.ops_l0_exchangeRead.push_back(CGExchangeReadMetaInfo{x + j, static_cast<uint16_t>(s + j)});
    return s;
  }

  void regWriteBlock(PartBuilder &B, uint32_t rbase, uint32_t nRegs) {
    B.beginOut(nRegs);
    CGRegWriteMetaInfo w{};
    w.reg = rbase;
    w.dat = static_cast<uint16_t>(B.outStart);
    w.count = static_cast<uint16_t>(B.outCount);
    if (B.outCount > 65535) throw std::runtime_error("register block too large");
    B.P.op_last_regWrite = w;
  }

  void compileLutB(PartBuilder &B, uint32_t pi, uint32_t rbase) {
    const auto &cls = lutPartsB[pi];
    uint32_t nRegs = 0;
    for (auto c : cls) nRegs += clB[c].regEnd - clB[c].regBegin;
    regWriteBlock(B, rbase, nRegs);
    for (auto c : cls)
      for (uint32_t r = clB[c].regBegin; r < clB[c].regEnd; ++r)
        B.slot[keyBit(clB[c].drivers[r - clB[c].regBegin])] = B.outStart + (regPos[r] - rbase);
    for (auto c : cls)
      for (auto n : clB[c].sources) exgSlot(B, keyBit(n), 1);
    for (auto c : cls)
      for (auto &lv : clB[c].levels)
        for (auto n : lv)
          if (!B.slot.count(keyBit(n))) B.slot[keyBit(n)] = B.alloc(1);
    uint32_t depth = 0;
    for (auto c : cls) depth = std::max<uint32_t>(depth, clB[c].levels.size());
    for (uint32_t l = 0; l < depth; ++l) {
      std::vector<uint32_t> luts;
      for (auto c : cls)
        if (l < clB[c].levels.size())
          for (auto n : clB[c].levels[l]) luts.push_back(n);
      emitLut4Level(B, l, luts);
    }
    // print / stop ops enabled by LUT outputs of this partition
    std::vector<uint32_t> pool;
    for (auto c : cls)
      for (auto &lv : clB[c].levels)
        for (auto n : lv) pool.push_back(n);
    for (uint32_t i = 0; i < stopPrintPart.size(); ++i) {
      if (stopPrintPart[i] != pi) continue;
      uint16_t en = static_cast<uint16_t>(B.slot.at(keyBit(pool[rng.below(pool.size())])));
      if (i < p.n_prints) {
        CGPrintMetaInfo pr{};
        pr.en = en;
        pr.msg = static_cast<uint16_t>(i);
        B.P.ops_last_print.push_back(pr);
      } else {
        CGStopMetaInfo st{};
        st.en = en;
        B.P.ops_last_stop.push_back(st);
      }
    }
  }

  void compileWordB(PartBuilder &B, const WbPart &wb, uint32_t rbase) {
    uint32_t n = 0;
    for (auto w : wb.regs) n = wregs[w].pos + wregs[w].nib - rbase;
    regWriteBlock(B, rbase, n);
    std::vector<TopOp> muxes;
    for (auto w : wb.regs) {
      const auto &r = wregs[w];
      uint32_t sx = exgSlot(B, keyRes(r.x), r.nib);
      uint32_t sy = exgSlot(B, keyRes(r.y), r.nib);
      uint16_t ss = 1;  // immediate 1: always x
      if (r.sel != NONE) ss = static_cast<uint16_t>(exgSlot(B, keyBit(r.sel), 1));
      uint32_t out = B.outStart + (r.pos - rbase);
      for (uint32_t j = 0; j < r.nib; ++j)
        muxes.push_back(TopOp{T_MUX, ss, static_cast<uint16_t>(sx + j), static_cast<uint16_t>(sy + j), out + j});
    }
    emitTopOps(B, 0, muxes);
  }

  void compileSramWrite(PartBuilder &B, const std::vector<uint32_t> &g, uint32_t rbase) {
    uint32_t n = 0;
    for (auto s : g) n = srams[s].waPos + srams[s].k - rbase;
    regWriteBlock(B, rbase, n);
    std::vector<TopOp> former;
    for (auto s : g) {
      const auto &m = srams[s];
      // next address registers straight from the exchange into the output block
      uint32_t xr = xpos.at(keyRaNext(s)), xw = xpos.at(keyWaNext(s));
      for (uint32_t j = 0; j < m.k; ++j) {
        B.P.ops_l0_exchangeRead.push_back(
            CGExchangeReadMetaInfo{xr + j, static_cast<uint16_t>(B.outStart + (m.raPos - rbase) + j)});
        B.P.ops_l0_exchangeRead.push_back(
            CGExchangeReadMetaInfo{xw + j, static_cast<uint16_t>(B.outStart + (m.waPos - rbase) + j)});
      }
      uint32_t wa = exgSlot(B, keyWaCur(s), m.k);
      uint32_t wd = exgSlot(B, keyRes(m.wdata), m.nib);
      uint16_t en = 1;
      if (m.we != NONE) en = static_cast<uint16_t>(exgSlot(B, keyBit(m.we), 1));
      uint32_t av = B.alloc(8);
      addrFormer(B, 0, m, wa, av, former);
      for (uint32_t j = 0; j < m.nib; ++j) {
        CGMemWriteMetaInfo w{};
        w.hasMultipleWriter = false;
        w.memBase = m.memBase + uint64_t(j) * m.depth;
        w.addrVec = static_cast<uint16_t>(av);
        w.dat = static_cast<uint16_t>(wd + j);
        w.en = en;
        B.P.ops_last_memWrite.push_back(w);
      }
    }
    emitTopOps(B, 0, former);
  }

  void compile() {
    D.lut = buildLutTable();
    layoutRegs();
    uint32_t xat = 0;
    std::vector<uint32_t> r0, r1;
    auto addPart = [&](PartBuilder &B, std::vector<uint32_t> &reg) {
      B.finish();
      reg.push_back(static_cast<uint32_t>(D.parts.size()));
      D.parts.push_back(std::move(B.P));
    };
    for (auto &cls : lutPartsA) {
      PartBuilder B;
      compileLutA(B, cls, xat);
      addPart(B, r0);
    }
    for (auto &wp : waParts) {
      if (wp.ops.empty()) continue;
      PartBuilder B;
      compileWordA(B, wp, xat);
      addPart(B, r0);
    }
    for (auto &g : sramGroups) {
      PartBuilder B;
      compileSramRead(B, g, xat);
      addPart(B, r0);
    }
    D.exchangePoolSize = (xat + 3) & ~3u;
    for (uint32_t i = 0; i < lutPartsB.size(); ++i) {
      PartBuilder B;
      compileLutB(B, i, lutBRegBase[i]);
      addPart(B, r1);
    }
    for (uint32_t i = 0; i < wbParts.size(); ++i) {
      PartBuilder B;
      compileWordB(B, wbParts[i], wbRegBase[i]);
      addPart(B, r1);
    }
    for (uint32_t i = 0; i < sramGroups.size(); ++i) {
      PartBuilder B;
      compileSramWrite(B, sramGroups[i], swRegBase[i]);
      addPart(B, r1);
    }
    if (r0.empty() || r1.empty()) throw std::runtime_error("design needs partitions in both regions");
    D.regionPartitionIds = {r0, r1};
    for (uint32_t i = 0; i < p.n_prints; ++i) D.printMsgs.push_back("print " + std::to_string(i) + " fired\n");
    D.shouldStop = false;

    // symbols
    for (uint32_t r = 0; r < p.n_regs; ++r) S.regDebugInfo["r" + std::to_string(r)] = {{regPos[r], 1u}};
    for (uint32_t i = 0; i < p.n_inputs; ++i) S.regDebugInfo["in" + std::to_string(i)] = {{inPos[i], 1u}};
    for (uint32_t w = 0; w < wregs.size(); ++w) {
      std::vector<std::tuple<uint32_t, uint32_t>> f;
      for (uint32_t j = wregs[w].nib; j-- > 0;) f.emplace_back(wregs[w].pos + j, 4u);
      S.regDebugInfo["w" + std::to_string(w)] = f;
    }
    for (uint32_t s = 0; s < srams.size(); ++s) {
      std::vector<std::tuple<uint32_t, uint32_t>> fr, fw;
      for (uint32_t j = srams[s].k; j-- > 0;) {
        fr.emplace_back(srams[s].raPos + j, 4u);
        fw.emplace_back(srams[s].waPos + j, 4u);
      }
      S.regDebugInfo["ra" + std::to_string(s)] = fr;
      S.regDebugInfo["wa" + std::to_string(s)] = fw;
      std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> planes;
      for (uint32_t j = 0; j < srams[s].nib; ++j) planes.emplace_back(srams[s].memBase + j * srams[s].depth, 4u, srams[s].depth);
      S.memDebugInfo["sram" + std::to_string(s)] = planes;
    }
    // a few named LUT signals (partition, slot, width) of region-0 exports
    uint32_t nSig = 0;
    for (uint32_t pi = 0; pi < lutPartsA.size() && nSig < 64; ++pi) {
      for (auto c : lutPartsA[pi]) {
        for (auto n : clA[c].exports) {
          if (nSig >= 64) break;
          // exported slots are the output block: exchange position minus block base
          uint32_t part = D.regionPartitionIds[0][pi];
          uint32_t slot = D.parts[part].op_last_exchangeWrite.dat +
                          (xpos.at(keyBit(n)) - D.parts[part].op_last_exchangeWrite.exchange);
          S.signalDebugInfo["sig" + std::to_string(nSig++)] = {{part, slot, 1u}};
        }
      }
    }
  }

  std::vector<uint8_t> netlist() const {
    std::vector<uint32_t> w;
    auto put = [&](uint32_t v) { w.push_back(v); };
    put(0x4C4E4754u);
    put(1);
    put(static_cast<uint32_t>(bits.size()));
    for (auto &n : bits) {
      put(n.kind); put(n.a); put(n.b); put(n.c); put(n.d); put(n.tt); put(n.lvl);
    }
    put(p.n_regs);
    for (uint32_t r = 0; r < p.n_regs; ++r) { put(regPos[r]); put(regNext[r]); }
    put(p.n_inputs);
    for (uint32_t i = 0; i < p.n_inputs; ++i) put(inPos[i]);
    put(static_cast<uint32_t>(wres.size()));
    for (auto &r : wres) {
      put(r.op); put(r.akind); put(r.a); put(r.bkind); put(r.b); put(r.nib); put(r.sel);
    }
    put(static_cast<uint32_t>(wregs.size()));
    for (auto &r : wregs) { put(r.pos); put(r.nib); put(r.sel); put(r.x); put(r.y); }
    put(static_cast<uint32_t>(srams.size()));
    for (auto &m : srams) {
      put(m.depth); put(m.nib); put(m.k); put(m.raPos); put(m.waPos); put(m.raMul); put(m.raAdd);
      put(m.waMul); put(m.waAdd); put(m.wdata); put(m.we); put(m.memBase); put(m.rdataRes);
    }
    put(D.regPoolSize);
    put(D.memPoolSize);
    std::vector<uint8_t> out(w.size() * 4);
    std::memcpy(out.data(), w.data(), out.size());
    return out;
  }
};

}  // namespace

struct tgs_gen_result {
  std::vector<uint8_t> design, symbols, netlist, inputs;
  tgs_gen_stats stats{};
  std::string error;
};

extern "C" {

int tgs_gen_default_params(tgs_gen_params *p) {
  if (!p) return TGS_E_ARG;
  std::memset(p, 0, sizeof(*p));
  p->seed = 1;
  p->n_luts = 20000;
  p->depth = 32;
  p->n_regs = 256;
  p->n_inputs = 64;
  p->n_vecops = 0;
  p->word_bits_mix = 50;
  p->n_srams = 0;
  p->sram_depth = 4096;
  p->sram_width = 64;
  p->part_luts = 2048;
  p->cluster_luts = 256;
  p->sources_a = 32;
  p->sources_b = 32;
  p->word_part_ops = 128;
  return TGS_OK;
}

int tgs_gen_build(const tgs_gen_params *p, tgs_gen_result **out) {
  if (!p || !out) return TGS_E_ARG;
  auto res = std::make_unique<tgs_gen_result>();
  try {
    Gen g(*p);
    g.plan();
    g.compile();
    writeSimDesignInfo(res->design, g.D);
    writeSimDebugInfo(res->symbols, g.S);
    res->netlist = g.netlist();
    res->inputs.resize(g.inPos.size() * 4);
    std::memcpy(res->inputs.data(), g.inPos.data(), res->inputs.size());
    auto &s = res->stats;
    s.n_parts[0] = g.D.regionPartitionIds[0].size();
    s.n_parts[1] = g.D.regionPartitionIds[1].size();
    s.n_luts = g.nLutsEmitted;
    s.n_vecops = g.nVecOpsEmitted;
    s.n_lut_ops = g.nLutOps;
    s.n_microparts = g.nMicroparts;
    s.reg_pool = g.D.regPoolSize;
    s.mem_pool = g.D.memPoolSize;
    s.exchange_pool = g.D.exchangePoolSize;
    uint64_t mx = 0;
    for (auto &pp : g.D.parts) mx = std::max<uint64_t>(mx, pp.valuePoolSize);
    s.max_value_pool = mx;
    s.design_bytes = res->design.size();
    s.symbol_bytes = res->symbols.size();
    s.netlist_bytes = res->netlist.size();
    s.n_inputs = g.inPos.size();
  } catch (const std::exception &e) {
    res->error = e.what();
    *out = res.release();
    return TGS_E_ARG;
  }
  *out = res.release();
  return TGS_OK;
}

int tgs_gen_get(const tgs_gen_result *r, int which, const uint8_t **data, size_t *size) {
  if (!r || !data || !size) return TGS_E_ARG;
  const std::vector<uint8_t> *v = nullptr;
  switch (which) {
    case 0: v = &r->design; break;
    case 1: v = &r->symbols; break;
    case 2: v = &r->netlist; break;
    case 3: v = &r->inputs; break;
    default: return TGS_E_ARG;
  }
  *data = v->data();
  *size = v->size();
  return TGS_OK;
}

int tgs_gen_get_stats(const tgs_gen_result *r, tgs_gen_stats *s) {
  if (!r || !s) return TGS_E_ARG;
  *s = r->stats;
  return TGS_OK;
}

int tgs_gen_error(const tgs_gen_result *r, char *buf, size_t len) {
  if (!r || !buf || !len) return TGS_E_ARG;
  std::snprintf(buf, len, "%s", r->error.c_str());
  return TGS_OK;
}

void tgs_gen_free(tgs_gen_result *r) { delete r; }

}  // extern "C"
