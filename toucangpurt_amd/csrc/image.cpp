// Host-side program checks, statistics and device-image builders.
#include "image.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <queue>
#include <unordered_map>

using namespace toucanGPUSim;

namespace tgs {
namespace {

uint32_t up4(uint32_t x) { return (x + 3u) & ~3u; }

bool setErr(std::string *err, const std::string &m) {
  if (err) *err = m;
  return false;
}

// Non-LUT micropart kind by the reference's priority (SimEval.cu:129-147):
// vecRead, then vecArithAndLogic, else memRead.
int otherKind(const CGMicroPartInfo &m) {
  if (!m.vecRead.empty()) return 1;
  if (!m.vecArithAndLogic.empty()) return 2;
  if (!m.memRead.empty()) return 3;
  return 0;
}

// Slots read / written by one micropartition, per lane (for the hazard rule).
struct Access {
  std::vector<std::pair<uint32_t, uint32_t>> reads, writes;  // (slot, lane)
};

void accessOf(const CGMicroPartInfo &m, Access &a) {
  auto rd = [&](uint32_t s, uint32_t lane) { a.reads.emplace_back(s, lane); };
  auto rdImm = [&](uint32_t s, uint32_t lane) {
    if (s >= 16) rd(s, lane);
  };
  if (m.isLUTPart) {
    for (uint32_t i = 0; i < m.topLevel.size(); ++i) {
      const auto &o = m.topLevel[i];
      rdImm(o.op2, i);
      if (o.lutIndex != 0) {
        rdImm(o.op0, i);
        rdImm(o.op1, i);
      }
    }
    // A LUT micropart reads all its top operands before its last level writes.
    for (uint32_t i = 0; i < m.lastLevel.size(); ++i) a.writes.emplace_back(m.lastLevel[i].result, 0);
    return;
  }
  switch (otherKind(m)) {
    case 1:
      for (uint32_t i = 0; i < m.vecRead.size(); ++i) {
        const auto &o = m.vecRead[i];
        rdImm(o.index0, i); rdImm(o.index1, i); rdImm(o.index2, i); rdImm(o.index3, i);
        rdImm(o.outRangeValue, i);
        if (!o.isConstVec)
          for (uint32_t k = 0; k < o.vecLength; ++k) rd(uint32_t(o.vecBase) + k, i);
        a.writes.emplace_back(o.result, i);
      }
      break;
    case 2:
      for (uint32_t i = 0; i < m.vecArithAndLogic.size(); ++i) {
        const auto &o = m.vecArithAndLogic[i];
        if (!(o.isV1V2Const & 2))
          for (uint32_t k = 0; k < o.vecLength; ++k) rd(uint32_t(o.vec1Base) + k, i);
        if (!(o.isV1V2Const & 1))
          for (uint32_t k = 0; k < o.vecLength; ++k) rd(uint32_t(o.vec2Base) + k, i);
        if (o.opName <= VEC_ARITH_MUL)
          for (uint32_t k = 0; k < o.vecLength; ++k) a.writes.emplace_back(uint16_t(o.result + k), i);
        else
          a.writes.emplace_back(o.result, i);
      }
      break;
    case 3:
      for (uint32_t i = 0; i < m.memRead.size(); ++i) {
        const auto &o = m.memRead[i];
        rdImm(o.en, i);
        for (uint32_t k = 0; k < 8; ++k) rd(uint32_t(o.addrVec) + k, i);
        a.writes.emplace_back(o.result, i);
      }
      break;
    default:
      break;
  }
}

}  // namespace

int validateProgram(const SimDesignInfo &d, std::string *err) {
  auto bad = [&](const std::string &m) {
    setErr(err, m);
    return TGS_E_INVALID;
  };
  if (d.lut.size() != kLutSize) return bad("LUT table must have 4642 entries (SimEval.cu:1122)");
  if (d.regionPartitionIds.size() != 2) return bad("exactly 2 regions are supported (SimEval.cu:1292-1295)");
  uint32_t expect = 0;
  for (const auto &r : d.regionPartitionIds)
    for (uint32_t id : r)
      if (id != expect++) return bad("region partition ids must be contiguous from 0 (SimEval.cu:1298-1302)");
  if (expect != d.parts.size()) return bad("regions must cover every partition");
  if (d.regionPartitionIds[0].empty() && d.regionPartitionIds[1].empty()) return bad("design has no partitions");
  const uint32_t regN = up4(d.regPoolSize), exgN = up4(d.exchangePoolSize);

  for (size_t pi = 0; pi < d.parts.size(); ++pi) {
    const auto &p = d.parts[pi];
    const std::string at = "partition " + std::to_string(pi) + ": ";
    if (p.valuePoolSize > 65535) {
      setErr(err, at + "value pool over 65535 slots (ToucanSim.cu:153)");
      return TGS_E_LIMIT;
    }
    const uint32_t vp = p.valuePoolSize, vpw = up4(vp);
    auto slot = [&](uint32_t s) { return s < vp; };
    auto opnd = [&](uint32_t s) { return s < 16 || s < vp; };
    if (p.ops_l0_regRead.empty() == p.ops_l0_exchangeRead.empty())
      return bad(at + "exactly one of regRead / exchangeRead must be non-empty (SimEval.cu:1183)");
    std::unordered_map<uint32_t, uint64_t> l0;
    for (const auto &o : p.ops_l0_regRead) {
      if (o.reg >= d.regPoolSize || !slot(o.result)) return bad(at + "register read out of range");
      auto ins = l0.emplace(o.result, (uint64_t(1) << 32) | o.reg);
      if (!ins.second && ins.first->second != ((uint64_t(1) << 32) | o.reg)) return bad(at + "two L0 reads write one slot");
    }
    for (const auto &o : p.ops_l0_exchangeRead) {
      if (o.exchange >= d.exchangePoolSize || !slot(o.result)) return bad(at + "exchange read out of range");
      auto ins = l0.emplace(o.result, (uint64_t(2) << 32) | o.exchange);
      if (!ins.second && ins.first->second != ((uint64_t(2) << 32) | o.exchange)) return bad(at + "two L0 reads write one slot");
    }
    for (size_t li = 0; li < p.exec_mParts.size(); ++li) {
      const auto &lv = p.exec_mParts[li];
      std::unordered_map<uint32_t, uint64_t> writer;  // slot -> (mpart << 32 | lane)
      std::vector<Access> acc(lv.size());
      for (size_t mi = 0; mi < lv.size(); ++mi) {
        const auto &m = lv[mi];
        const std::string mat = at + "level " + std::to_string(li) + " micropart " + std::to_string(mi) + ": ";
        if (m.isLUTPart) {
          if (m.topLevel.size() > 32 || m.lastLevel.size() > 32) return bad(mat + "more than 32 lanes (SimEval.cu:93-94)");
          if (m.middleLevels.size() + 2 > 160) return bad(mat + "more than 158 middle levels (SimEval.cu:92)");
          for (const auto &lvl : m.middleLevels)
            if (lvl.size() > 32) return bad(mat + "more than 32 lanes (SimEval.cu:95-97)");
          for (const auto &o : m.topLevel)
            if (!opnd(o.op2) || (o.lutIndex && (!opnd(o.op0) || !opnd(o.op1)))) return bad(mat + "top operand out of range");
          for (const auto &o : m.lastLevel)
            if (!slot(o.result)) return bad(mat + "last-level slot out of range");
        } else {
          int k = otherKind(m);
          if (k == 0) return bad(mat + "empty micropart (SimEval.cu:143)");
          size_t n = k == 1 ? m.vecRead.size() : k == 2 ? m.vecArithAndLogic.size() : m.memRead.size();
          if (n > 32) return bad(mat + "more than 32 ops in one micropart");
          if (k == 1)
            for (const auto &o : m.vecRead) {
              if (!opnd(o.index0) || !opnd(o.index1) || !opnd(o.index2) || !opnd(o.index3) || !opnd(o.outRangeValue) ||
                  !slot(o.result))
                return bad(mat + "vecRead operand out of range");
              uint32_t end = uint32_t(o.vecBase) + o.vecLength;
              if (o.vecLength && (o.isConstVec ? end > p.constVecPool.size() : end > vp))
                return bad(mat + "vecRead vector out of range");
            }
          if (k == 2)
            for (const auto &o : m.vecArithAndLogic) {
              if (o.vecLength > 32) return bad(mat + "vector op longer than 32 nibbles (128 bits)");
              uint32_t e1 = uint32_t(o.vec1Base) + o.vecLength, e2 = uint32_t(o.vec2Base) + o.vecLength;
              if ((o.isV1V2Const & 2) ? e1 > p.constVecPool.size() : e1 > vp) return bad(mat + "vector operand out of range");
              if ((o.isV1V2Const & 1) ? e2 > p.constVecPool.size() : e2 > vp) return bad(mat + "vector operand out of range");
              uint32_t we = o.opName <= VEC_ARITH_MUL ? uint32_t(o.result) + o.vecLength : uint32_t(o.result) + 1;
              if (we > vp) return bad(mat + "vector result out of range");
            }
          if (k == 3)
            for (const auto &o : m.memRead)
              if (!opnd(o.en) || uint32_t(o.addrVec) + 8 > vp || !slot(o.result)) return bad(mat + "memRead operand out of range");
        }
        accessOf(m, acc[mi]);
        for (const auto &w : acc[mi].writes) {
          uint64_t tag = (uint64_t(mi) << 32) | w.second;
          auto ins = writer.emplace(w.first, tag);
          if (!ins.second && ins.first->second != tag)
            return bad(mat + "slot " + std::to_string(w.first) + " written twice in one exec level");
        }
      }
      for (size_t mi = 0; mi < lv.size(); ++mi)
        for (const auto &r : acc[mi].reads) {
          auto it = writer.find(r.first);
          if (it == writer.end()) continue;
          uint32_t wm = uint32_t(it->second >> 32), wl = uint32_t(it->second);
          if (wm != mi || (!lv[mi].isLUTPart && wl != r.second))
            return bad(at + "level " + std::to_string(li) + ": slot " + std::to_string(r.first) +
                       " read and written by concurrent ops");
        }
    }
    const auto &ew = p.op_last_exchangeWrite;
    if (ew.count) {
      uint32_t w = (uint32_t(ew.count) + 3) / 4;
      if (((ew.dat >> 2) + w) * 4 > vpw || ((ew.exchange >> 2) + w) * 4 > exgN) return bad(at + "exchange write out of range");
    } else {
      const auto &rw = p.op_last_regWrite;
      if (rw.count) {
        uint32_t w = (uint32_t(rw.count) + 3) / 4;
        if (((rw.dat >> 2) + w) * 4 > vpw || ((rw.reg >> 2) + w) * 4 > regN) return bad(at + "register write out of range");
      }
      for (const auto &o : p.ops_last_memWrite)
        if (!opnd(o.en) || !opnd(o.dat) || uint32_t(o.addrVec) + 8 > vp) return bad(at + "memWrite operand out of range");
      for (const auto &o : p.ops_last_print)
        if (!opnd(o.en) || o.msg >= d.printMsgs.size()) return bad(at + "print operand out of range");
      for (const auto &o : p.ops_last_stop)
        if (!opnd(o.en)) return bad(at + "stop operand out of range");
    }
  }
  return TGS_OK;
}

void programStats(const SimDesignInfo &d, uint64_t nInputs, tgs_program_stats *s) {
  std::memset(s, 0, sizeof(*s));
  if (d.regionPartitionIds.size() == 2) {
    s->n_parts[0] = d.regionPartitionIds[0].size();
    s->n_parts[1] = d.regionPartitionIds[1].size();
  }
  uint64_t bs = 0;
  for (const auto &p : d.parts) {
    s->n_regread += p.ops_l0_regRead.size();
    s->n_exgread += p.ops_l0_exchangeRead.size();
    s->n_exec_levels += p.exec_mParts.size();
    s->max_value_pool = std::max<uint64_t>(s->max_value_pool, p.valuePoolSize);
    for (const auto &lv : p.exec_mParts)
      for (const auto &m : lv) {
        ++s->n_microparts;
        if (m.isLUTPart) {
          s->n_top += m.topLevel.size();
          for (const auto &l : m.middleLevels) s->n_mid += l.size();
          s->n_last += m.lastLevel.size();
        } else {
          int k = otherKind(m);
          if (k == 1) s->n_vecread += m.vecRead.size();
          if (k == 2) s->n_vecop += m.vecArithAndLogic.size();
          if (k == 3) s->n_memread += m.memRead.size();
        }
      }
    if (p.op_last_exchangeWrite.count) {
      ++s->n_exgwrite;
      bs += p.op_last_exchangeWrite.count;
    } else {
      if (p.op_last_regWrite.count) {
        ++s->n_regwrite;
        bs += p.op_last_regWrite.count;
      }
      s->n_memwrite += p.ops_last_memWrite.size();
      s->n_print += p.ops_last_print.size();
      s->n_stop += p.ops_last_stop.size();
    }
  }
  s->b_prog = 8 * s->n_top + 4 * s->n_mid + 4 * s->n_last + 20 * s->n_vecread + 10 * s->n_vecop + 24 * s->n_memread +
              8 * s->n_regread + 8 * s->n_exgread + 8 * s->n_regwrite + 8 * s->n_exgwrite + 24 * s->n_memwrite +
              4 * s->n_print + 2 * s->n_stop;
  s->b_state = s->n_regread + s->n_exgread + bs + s->n_memread + s->n_memwrite + nInputs;
  s->reg_pool = d.regPoolSize;
  s->mem_pool = d.memPoolSize;
  s->exchange_pool = d.exchangePoolSize;
}

int buildStimMap(const SimDesignInfo &d, const uint32_t *ids, uint32_t n, StimMap *m, std::string *err) {
  m->ids.clear();
  m->inputOf.clear();
  if (n == 0) return TGS_OK;
  m->inputOf.assign(d.regPoolSize, -1);
  for (uint32_t i = 0; i < n; ++i) {
    if (ids[i] >= d.regPoolSize) {
      setErr(err, "stimulus input id out of range");
      return TGS_E_ARG;
    }
    if (m->inputOf[ids[i]] >= 0) {
      setErr(err, "duplicate stimulus input id");
      return TGS_E_ARG;
    }
    m->inputOf[ids[i]] = static_cast<int32_t>(i);
    m->ids.push_back(ids[i]);
  }
  for (const auto &p : d.parts) {
    if (p.op_last_exchangeWrite.count || !p.op_last_regWrite.count) continue;
    uint32_t w = (uint32_t(p.op_last_regWrite.count) + 3) / 4;
    uint32_t lo = (p.op_last_regWrite.reg >> 2) * 4, hi = lo + 4 * w;
    for (uint32_t r = lo; r < hi && r < d.regPoolSize; ++r)
      if (m->inputOf[r] >= 0) {
        setErr(err, "a stimulus input is overwritten by a register write");
        m->ids.clear();
        m->inputOf.clear();
        return TGS_E_ARG;
      }
  }
  return TGS_OK;
}

// ---------------------------------------------------------------------------
// Plain image (reference layout, SimEval.cu:69-148,1143-1281)

namespace {
void align16(std::vector<uint8_t> &b) {
  while (b.size() % 16) b.push_back(0);
}
template <typename T>
void put(std::vector<uint8_t> &b, const T &v) {
  const uint8_t *p = reinterpret_cast<const uint8_t *>(&v);
  b.insert(b.end(), p, p + sizeof(T));
}
template <typename T>
void putVec(std::vector<uint8_t> &b, const std::vector<T> &v) {
  for (const auto &e : v) put(b, e);
}
}  // namespace

PlainImage buildPlainImage(const SimDesignInfo &d, const StimMap &stim) {
  PlainImage img;
  img.n0 = d.regionPartitionIds[0].size();
  img.nTotal = d.parts.size();
  img.maxRegion = std::max<uint32_t>(d.regionPartitionIds[0].size(), d.regionPartitionIds[1].size());
  for (const auto &p : d.parts) {
    PlainPart pp{};
    align16(img.net);
    pp.net = static_cast<uint32_t>(img.net.size());
    std::vector<uint8_t> b;
    img.maxVp = std::max(img.maxVp, p.valuePoolSize);
    std::vector<CGRegReadMetaInfo> rr;
    std::vector<std::pair<uint32_t, uint32_t>> sr;
    for (const auto &o : p.ops_l0_regRead) {
      int32_t in = stim.lookup(o.reg);
      if (in >= 0) sr.emplace_back(uint32_t(in), o.result);
      else rr.push_back(o);
    }
    pp.nRegRead = rr.size();
    pp.offRegRead = b.size();
    putVec(b, rr);
    align16(b);
    pp.nExgRead = p.ops_l0_exchangeRead.size();
    pp.offExgRead = b.size();
    putVec(b, p.ops_l0_exchangeRead);
    align16(b);
    pp.nStimRead = sr.size();
    pp.offStimRead = b.size();
    for (auto &s : sr) {
      put(b, s.first);
      put(b, s.second);
    }
    align16(b);
    pp.nLevels = p.exec_mParts.size();
    pp.levelBase = img.levels.size() / 2;
    for (const auto &lv : p.exec_mParts) {
      img.levels.push_back(lv.size());
      img.levels.push_back(img.mparts.size());
      for (const auto &m : lv) {
        img.mparts.push_back(b.size());
        if (m.isLUTPart) {
          put(b, kMagicLut);
          put(b, uint32_t(m.middleLevels.size()));
          uint8_t sizes[160] = {0};
          sizes[0] = static_cast<uint8_t>(m.topLevel.size());
          for (size_t l = 0; l < m.middleLevels.size(); ++l) sizes[1 + l] = static_cast<uint8_t>(m.middleLevels[l].size());
          sizes[1 + m.middleLevels.size()] = static_cast<uint8_t>(m.lastLevel.size());
          b.insert(b.end(), sizes, sizes + 160);
          for (auto o : m.topLevel) {
            if (o.lutIndex == 0) o.op0 = o.op1 = 0;  // unused operands of a pass-through op
            put(b, o);
          }
          for (const auto &l : m.middleLevels) putVec(b, l);
          putVec(b, m.lastLevel);
        } else {
          int k = otherKind(m);
          if (k == 1) {
            put(b, kMagicVecRead);
            put(b, uint32_t(m.vecRead.size()));
            putVec(b, m.vecRead);
          } else if (k == 2) {
            put(b, kMagicVecOp);
            put(b, uint32_t(m.vecArithAndLogic.size()));
            putVec(b, m.vecArithAndLogic);
          } else {
            put(b, kMagicMemRead);
            put(b, uint32_t(m.memRead.size()));
            putVec(b, m.memRead);
          }
        }
        align16(b);
      }
    }
    pp.offLast = b.size();
    pp.hasExgWrite = p.op_last_exchangeWrite.count != 0;
    pp.hasRegWrite = !pp.hasExgWrite && p.op_last_regWrite.count != 0;
    pp.offRegWrite = b.size();
    put(b, p.op_last_regWrite);
    align16(b);
    pp.offExgWrite = b.size();
    put(b, p.op_last_exchangeWrite);
    align16(b);
    pp.nMemWrite = p.ops_last_memWrite.size();
    pp.offMemWrite = b.size();
    putVec(b, p.ops_last_memWrite);
    align16(b);
    pp.nPrint = p.ops_last_print.size();
    pp.offPrint = b.size();
    putVec(b, p.ops_last_print);
    align16(b);
    pp.nStop = p.ops_last_stop.size();
    pp.offStop = b.size();
    putVec(b, p.ops_last_stop);
    align16(b);
    pp.constBase = img.consts.size();
    pp.nConst = p.constVecPool.size();
    img.consts.insert(img.consts.end(), p.constVecPool.begin(), p.constVecPool.end());
    img.net.insert(img.net.end(), b.begin(), b.end());
    img.parts.push_back(pp);
  }
  align16(img.net);
  if (img.consts.empty()) img.consts.push_back(0);
  return img;
}

// ---------------------------------------------------------------------------
// Wave64 image

// A T_LUT2 pair (A, B may be null) as a T_LUTROWS task for a value pool of V
// rows. Lanes at or past the previous level's size read 0 (SimEval.cu:
// 372-375,406-408, 441-446) and become the constant row of 0; a pass-through
// op's unused op0/op1 read the constant rows too.
static std::vector<uint32_t> lutRowsTask(const CGMicroPartInfo *A, const CGMicroPartInfo *B, uint32_t V) {
  const uint32_t nMid = A->middleLevels.size();
  const CGMicroPartInfo *H[2] = {A, B};
  auto prevSize = [](const CGMicroPartInfo *m, uint32_t l) -> uint32_t {
    return uint32_t(l == 0 ? m->topLevel.size() : m->middleLevels[l - 1].size());
  };
  // level row of reference lane `o` (< 32 immediate, else lane o mod 32) read
  // from buffer base `rb` by half h
  auto laneRow = [&](uint32_t o, uint32_t h, uint32_t rb, uint32_t nPrev) -> uint32_t {
    if (o < 32) return kConstRow + o;
    return (o & 31u) < nPrev ? rb + 32 * h + (o & 31u) : kConstRow;
  };
  std::vector<uint32_t> t;
  uint32_t nt = 0, nl = 0;
  for (auto *m : H)
    if (m) {
      nt += m->topLevel.size();
      nl += m->lastLevel.size();
    }
  t.push_back(T_LUTROWS | nMid << 8);
  t.push_back(nt | nl << 16);
  t.push_back(0);
  t.push_back(0);
  for (uint32_t l = 0; l < nMid; l += 4) {
    uint32_t v = 0;
    for (uint32_t k = 0; k < 4 && l + k < nMid; ++k) {
      uint32_t n = 0;
      for (auto *m : H)
        if (m) n += m->middleLevels[l + k].size();
      v |= n << (8 * k);
    }
    t.push_back(v);
  }
  if (t.size() & 1) t.push_back(0);
  // top level: value-pool operands (slots below 16 are immediates: constant rows)
  auto topRow = [&](uint32_t s) -> uint32_t { return s < 16 ? V + kConstRow + s : s; };
  for (uint32_t h = 0; h < 2; ++h) {
    if (!H[h]) continue;
    uint32_t k = 0;
    for (const auto &o : H[h]->topLevel) {
      const uint32_t r0 = o.lutIndex ? topRow(o.op0) : V + kConstRow, r1 = o.lutIndex ? topRow(o.op1) : V + kConstRow;
      t.push_back(uint32_t(o.lutIndex) | r0 << 16);
      t.push_back(r1 | topRow(o.op2) << 13 | (32 * h + k) << 26);
      ++k;
    }
  }
  t[2] = t.size();
  // A lane of the last middle level that exactly one last-level op copies
  // (the usual case) is written straight to that op's value-pool slot, and
  // the op is dropped; fan-outs and immediates stay last-level ops.
  std::vector<int32_t> direct[2];  // per half and lane: the one copying last op, -1 none, -2 several
  for (uint32_t h = 0; h < 2 && nMid; ++h) {
    if (!H[h]) continue;
    const uint32_t nPrev = prevSize(H[h], nMid);
    direct[h].assign(32, -1);
    for (uint32_t q = 0; q < H[h]->lastLevel.size(); ++q) {
      const uint32_t sid = H[h]->lastLevel[q].shuffleId;
      if (sid < 32 || (sid & 31u) >= nPrev) continue;
      int32_t &e = direct[h][sid & 31u];
      e = e == -1 ? int32_t(q) : -2;
    }
  }
  for (uint32_t l = 0; l < nMid; ++l) {
    const uint32_t rb = (l & 1) ? 64 : 0, wb = (l & 1) ? 0 : 64;
    for (uint32_t h = 0; h < 2; ++h) {
      if (!H[h]) continue;
      const uint32_t nPrev = prevSize(H[h], l);
      uint32_t k = 0;
      for (const auto &o : H[h]->middleLevels[l]) {
        const uint32_t lt = o.lutIndex();
        const uint32_t r0 = lt ? laneRow(o.op0(), h, rb, nPrev) : kConstRow;
        const uint32_t r1 = lt ? laneRow(o.op1(), h, rb, nPrev) : kConstRow;
        uint32_t dst = V + wb + 32 * h + k;
        if (l + 1 == nMid && k < 32 && direct[h][k] >= 0) dst = H[h]->lastLevel[direct[h][k]].result;
        t.push_back(lt | r0 << 14 | r1 << 22);
        t.push_back(laneRow(o.op2(), h, rb, nPrev) | dst << 8);
        ++k;
      }
    }
  }
  t[3] = t.size();
  const uint32_t fb = (nMid & 1) ? 64 : 0;
  uint32_t nlKept = 0;
  for (uint32_t h = 0; h < 2; ++h) {
    if (!H[h]) continue;
    const uint32_t nPrev = prevSize(H[h], nMid);
    for (uint32_t q = 0; q < H[h]->lastLevel.size(); ++q) {
      const auto &o = H[h]->lastLevel[q];
      const uint32_t sid = o.shuffleId;
      if (nMid && sid >= 32 && (sid & 31u) < nPrev && direct[h][sid & 31u] == int32_t(q)) continue;
      t.push_back(uint32_t(o.result) | laneRow(sid, h, fb, nPrev) << 16);
      ++nlKept;
    }
  }
  t[1] = nt | nlKept << 16;
  if (t.size() & 1) t.push_back(0);
  return t;
}

// Level sizes of LUT micropartitions packed into one T_LUTFAST task, levels
// aligned at the end: with nMid the largest middle-level count, part h
// (middle-level count m_h) has its top level on task level nMid - m_h and
// its last middle level on task level nMid. Entry L is the ops on task
// level L (0 = the first).
static std::vector<uint32_t> packedSizes(const std::vector<const CGMicroPartInfo *> &H) {
  uint32_t nMid = 0;
  for (auto *m : H) nMid = std::max<uint32_t>(nMid, m->middleLevels.size());
  std::vector<uint32_t> tot(nMid + 1, 0);
  for (auto *m : H) {
    const uint32_t s0 = nMid - m->middleLevels.size();
    tot[s0] += m->topLevel.size();
    for (size_t l = 0; l < m->middleLevels.size(); ++l) tot[s0 + 1 + l] += m->middleLevels[l].size();
  }
  return tot;
}

// LUT micropartitions of one exec level (at most 64 ops on any task level
// together, levels aligned at the end as in packedSizes) as one T_LUTFAST
// task for I instances per wave and a value pool of V rows; false (and no
// words) when an op needs the bounds-checked lookup of T_LUTROWS. On every
// task level the parts' ops follow each other (part h's after parts
// 0..h-1's); top and middle ops have one form, so a part's top level may sit
// on any task level.
static bool lutFastTask(const std::vector<const CGMicroPartInfo *> &H, uint32_t V, uint32_t I,
                        std::vector<uint32_t> &t) {
  t.clear();
  const uint32_t nh = H.size();
  const std::vector<uint32_t> tot = packedSizes(H);
  const uint32_t nMid = tot.size() - 1;
  std::vector<uint32_t> start(nh);  // task level of part h's top level
  for (uint32_t h = 0; h < nh; ++h) start[h] = nMid - H[h]->middleLevels.size();
  // ops of part h on task level L (0 before its top level)
  auto sizeAt = [&](uint32_t h, uint32_t L) -> uint32_t {
    if (L < start[h]) return 0;
    return uint32_t(L == start[h] ? H[h]->topLevel.size() : H[h]->middleLevels[L - start[h] - 1].size());
  };
  std::vector<std::vector<uint32_t>> first(nMid + 1, std::vector<uint32_t>(nh + 1, 0));  // part h's op 0
  for (uint32_t L = 0; L <= nMid; ++L) {
    for (uint32_t h = 0; h < nh; ++h) first[L][h + 1] = first[L][h] + sizeAt(h, L);
    if (first[L][nh] > 64) return false;
  }
  auto off = [&](uint32_t row) -> uint32_t { return row * I; };
  const uint32_t zeroRow = V + kConstRow;
  // level buffer written by task level L (0: X; then alternately Y, X, ...)
  auto bufOf = [](uint32_t L) -> uint32_t { return L == 0 ? 0u : ((L - 1) & 1u) ? 0u : kFastBufY; };
  // row read through reference lane o by an op of part h whose sources are
  // task level L (lanes at or past that level's size read 0; o < 32 is an
  // immediate)
  auto laneRow = [&](uint32_t o, uint32_t h, uint32_t L) -> uint32_t {
    if (o < 32) return V + kConstRow + o;
    const uint32_t lane = o & 31u;
    if (lane >= sizeAt(h, L)) return zeroRow;
    return V + bufOf(L) + first[L][h] + lane;
  };
  auto lutPos = [](uint32_t lt) -> uint32_t { return lt == 0 ? kLutIdBase : std::min(lt, kLutSize); };
  auto topRow = [&](uint32_t s) -> uint32_t { return s < 16 ? V + kConstRow + s : s; };
  uint32_t nl = 0;
  t.push_back(T_LUTFAST | nMid << 8);
  t.push_back(0);
  t.push_back(0);
  t.push_back(0);
  for (uint32_t l = 0; l < nMid; l += 4) {
    uint32_t v = 0;
    for (uint32_t k = 0; k < 4 && l + k < nMid; ++k) v |= tot[l + k + 1] << (8 * k);
    t.push_back(v);
  }
  if (t.size() & 1) t.push_back(0);
  // lanes of a part's last level copied by exactly one last-level op are
  // written straight to that op's slot when that level is a middle task
  // level (it has a destination list)
  std::vector<std::vector<int32_t>> direct(nh);
  for (uint32_t h = 0; h < nh && nMid; ++h) {
    const uint32_t nPrev = sizeAt(h, nMid);
    direct[h].assign(32, -1);
    for (uint32_t q = 0; q < H[h]->lastLevel.size(); ++q) {
      const uint32_t sid = H[h]->lastLevel[q].shuffleId;
      if (sid < 32 || (sid & 31u) >= nPrev) continue;
      int32_t &e = direct[h][sid & 31u];
      e = e == -1 ? int32_t(q) : -2;
    }
  }
  for (uint32_t L = 0; L <= nMid; ++L) {
    if (L == 1) t[2] = t.size();
    std::vector<uint32_t> dsts;
    for (uint32_t h = 0; h < nh; ++h) {
      if (L < start[h]) continue;
      uint32_t q = 0;
      auto push = [&](uint32_t lutp, uint32_t r0, uint32_t r1, uint32_t r2) {
        t.push_back(lutp | off(r0) << 16);
        t.push_back(off(r1) | off(r2) << 16);
        if (L == nMid && L > 0) {
          uint32_t dst = V + bufOf(L) + first[L][h] + q;
          if (q < 32 && direct[h][q] >= 0) dst = H[h]->lastLevel[direct[h][q]].result;
          dsts.push_back(off(dst));
        }
        ++q;
      };
      if (L == start[h]) {
        for (const auto &o : H[h]->topLevel) {
          const uint32_t lt = o.lutIndex;
          if (lt > 65536u - 4096u) return false;  // the 16-bit index could wrap
          push(lutPos(lt), lt ? topRow(o.op0) : zeroRow, lt ? topRow(o.op1) : zeroRow, topRow(o.op2));
        }
      } else {
        for (const auto &o : H[h]->middleLevels[L - start[h] - 1]) {
          const uint32_t lt = o.lutIndex();
          const uint32_t o0 = o.op0(), o1 = o.op1(), o2 = o.op2();
          if (lt && ((o0 < 32 && o0 >= 16) || (o1 < 32 && o1 >= 16) || (o2 < 32 && o2 >= 16))) return false;
          push(lutPos(lt), lt ? laneRow(o0, h, L - 1) : zeroRow, lt ? laneRow(o1, h, L - 1) : zeroRow,
               laneRow(o2, h, L - 1));
        }
      }
    }
    for (size_t j = 0; j < dsts.size(); j += 2)
      t.push_back(dsts[j] | (j + 1 < dsts.size() ? dsts[j + 1] : dsts[j]) << 16);
  }
  if (nMid == 0) t[2] = t.size();
  t[3] = t.size();
  for (uint32_t h = 0; h < nh; ++h) {
    const uint32_t nPrev = sizeAt(h, nMid);
    for (uint32_t q = 0; q < H[h]->lastLevel.size(); ++q) {
      const auto &o = H[h]->lastLevel[q];
      const uint32_t sid = o.shuffleId;
      if (nMid && sid >= 32 && (sid & 31u) < nPrev && direct[h][sid & 31u] == int32_t(q)) continue;
      uint32_t src;
      if (sid < 32) src = V + kConstRow + sid;
      else if ((sid & 31u) >= nPrev) src = zeroRow;
      else src = V + bufOf(nMid) + first[nMid][h] + (sid & 31u);
      t.push_back(off(o.result) | off(src) << 16);
      ++nl;
    }
  }
  t[1] = tot[0] | nl << 16;
  if (t.size() & 1) t.push_back(0);
  return true;
}

bool nibbleProgram(const SimDesignInfo &d) {
  for (uint8_t v : d.lut)
    if (v > 15) return false;
  for (const auto &p : d.parts) {
    for (uint8_t v : p.constVecPool)
      if (v > 15) return false;
    for (const auto &lv : p.exec_mParts)
      for (const auto &m : lv) {
        if (!m.isLUTPart) continue;
        // an immediate of 16..31 passed through or copied out becomes a value
        for (const auto &ml : m.middleLevels)
          for (const auto &o : ml)
            if (o.lutIndex() == 0 && o.op2() >= 16 && o.op2() < 32) return false;
        for (const auto &o : m.lastLevel)
          if (o.shuffleId >= 16 && o.shuffleId < 32) return false;
      }
  }
  return true;
}

Wave64Image buildWave64Image(const SimDesignInfo &d, const StimMap &stim, uint32_t nWaves, uint32_t vpRows,
                             uint32_t fastInst) {
  Wave64Image img;
  img.n0 = d.regionPartitionIds[0].size();
  img.nTotal = d.parts.size();
  for (const auto &p : d.parts) {
    std::vector<uint32_t> w(H_WORDS, 0);
    img.maxVp = std::max(img.maxVp, p.valuePoolSize);
    w[H_VP] = p.valuePoolSize;
    // L0: register reads, exchange reads, stimulus reads; 2 words each
    std::vector<std::pair<uint32_t, uint32_t>> rr, sr;
    for (const auto &o : p.ops_l0_regRead) {
      int32_t in = stim.lookup(o.reg);
      if (in >= 0) sr.emplace_back(uint32_t(in), o.result);
      else rr.emplace_back(o.reg, o.result);
    }
    w[H_NRR] = rr.size();
    w[H_NER] = p.ops_l0_exchangeRead.size();
    w[H_NSR] = sr.size();
    w[H_OFF_L0] = w.size();
    for (auto &e : rr) { w.push_back(e.first); w.push_back(e.second); }
    for (const auto &o : p.ops_l0_exchangeRead) { w.push_back(o.exchange); w.push_back(o.result); }
    for (auto &e : sr) { w.push_back(e.first); w.push_back(e.second); }

    // tasks per exec level
    std::vector<std::vector<uint32_t>> tasks(p.exec_mParts.size());  // task streams (relative words)
    std::vector<std::vector<uint32_t>> levelTasks;                   // per level: task word vectors
    std::vector<std::vector<std::vector<uint32_t>>> streams(p.exec_mParts.size());
    uint64_t cost = 4 + (rr.size() + sr.size() + p.ops_l0_exchangeRead.size()) / 64;
    for (size_t li = 0; li < p.exec_mParts.size(); ++li) {
      const auto &lv = p.exec_mParts[li];
      auto &st = streams[li];
      std::vector<const CGMicroPartInfo *> luts;
      std::vector<const CGMicroPartVecRead *> vr;
      std::vector<const CGMicroPartVecArithOrLogic *> vo;
      std::vector<const CGMicroPartMemRead *> mr;
      for (const auto &m : lv) {
        if (m.isLUTPart) luts.push_back(&m);
        else {
          int k = otherKind(m);
          if (k == 1) for (const auto &o : m.vecRead) vr.push_back(&o);
          if (k == 2) for (const auto &o : m.vecArithAndLogic) vo.push_back(&o);
          if (k == 3) for (const auto &o : m.memRead) mr.push_back(&o);
        }
      }
      std::stable_sort(luts.begin(), luts.end(), [](const CGMicroPartInfo *a, const CGMicroPartInfo *b) {
        return a->middleLevels.size() > b->middleLevels.size();
      });
      uint64_t levelCost = 0;
      for (size_t i = 0; i < luts.size();) {
        const CGMicroPartInfo *A = luts[i];
        const CGMicroPartInfo *B = nullptr;
        if (i + 1 < luts.size() && luts[i + 1]->middleLevels.size() == A->middleLevels.size()) B = luts[i + 1];
        i += B ? 2 : 1;
        uint32_t nMid = A->middleLevels.size();
        if (vpRows && fastInst) {
          // T_LUTFAST: this part and the following ones (fewer middle
          // levels: aligned at the end) while every task level holds at
          // most 64 ops
          std::vector<const CGMicroPartInfo *> grp{A};
          if (B) grp.push_back(B);
          while (i < luts.size()) {
            grp.push_back(luts[i]);
            const auto tot = packedSizes(grp);
            if (*std::max_element(tot.begin(), tot.end()) > 64) {
              grp.pop_back();
              break;
            }
            ++i;
          }
          std::vector<uint32_t> fast;
          if (lutFastTask(grp, vpRows, fastInst, fast)) {
            st.push_back(std::move(fast));
            ++img.fastTasks;
          } else {
            for (size_t j = 0; j < grp.size();) {
              const CGMicroPartInfo *b = j + 1 < grp.size() &&
                                                 grp[j + 1]->middleLevels.size() == grp[j]->middleLevels.size()
                                             ? grp[j + 1]
                                             : nullptr;
              st.push_back(lutRowsTask(grp[j], b, vpRows));
              ++img.rowTasks;
              j += b ? 2 : 1;
            }
          }
          levelCost += nMid + 2;
          continue;
        }
        if (vpRows) {
          st.push_back(lutRowsTask(A, B, vpRows));
          ++img.rowTasks;
          levelCost += nMid + 2;
          continue;
        }
        std::vector<uint32_t> t;
        t.push_back(T_LUT2 | nMid << 8);
        uint32_t ntA = A->topLevel.size(), ntB = B ? B->topLevel.size() : 0;
        uint32_t nlA = A->lastLevel.size(), nlB = B ? B->lastLevel.size() : 0;
        t.push_back(ntA | ntB << 8 | nlA << 16 | nlB << 24);
        t.push_back(0);  // word offset of the last-level ops, filled below
        for (uint32_t l = 0; l < nMid; l += 2) {
          uint32_t v = 0;
          for (uint32_t k = 0; k < 2 && l + k < nMid; ++k) {
            uint32_t na = A->middleLevels[l + k].size(), nb = B ? B->middleLevels[l + k].size() : 0;
            v |= (na | nb << 8) << (16 * k);
          }
          t.push_back(v);
        }
        auto tops = [&](const CGMicroPartInfo *m) {
          if (!m) return;
          for (const auto &o : m->topLevel) {
            // a pass-through op ignores op0/op1: store immediates so no kernel reads an unchecked slot
            const uint32_t o0 = o.lutIndex ? o.op0 : 0u, o1 = o.lutIndex ? o.op1 : 0u;
            t.push_back(uint32_t(o.lutIndex) | o0 << 16);
            t.push_back(o1 | uint32_t(o.op2) << 16);
          }
        };
        tops(A);
        tops(B);
        // Operands that name a lane at or past the previous level's size read
        // 0 (SimEval.cu:372-375,406-408): they are stored as the immediate 0,
        // and a pass-through op's unused op0/op1 as immediates too, so the
        // kernels never need the previous level's size to read an operand.
        auto prevSize = [](const CGMicroPartInfo *m, uint32_t l) -> uint32_t {
          return uint32_t(l == 0 ? m->topLevel.size() : m->middleLevels[l - 1].size());
        };
        auto lane = [](uint32_t o, uint32_t nPrev) -> uint32_t { return o >= 32 && (o & 31u) >= nPrev ? 0u : o; };
        auto mids = [&](const CGMicroPartInfo *m, uint32_t l) {
          if (!m) return;
          const uint32_t nPrev = prevSize(m, l);
          for (const auto &o : m->middleLevels[l]) {
            const uint32_t w = o.getPacked(), lt = w >> 18;
            const uint32_t o0 = lt ? lane((w >> 12) & 63u, nPrev) : 0u, o1 = lt ? lane((w >> 6) & 63u, nPrev) : 0u;
            t.push_back(lt << 18 | o0 << 12 | o1 << 6 | lane(w & 63u, nPrev));
          }
        };
        for (uint32_t l = 0; l < nMid; ++l) {
          mids(A, l);
          mids(B, l);
        }
        auto lasts = [&](const CGMicroPartInfo *m) {
          if (!m) return;
          const uint32_t nPrev = prevSize(m, nMid);
          for (const auto &o : m->lastLevel)
            t.push_back(uint32_t(o.result) | lane(uint32_t(o.shuffleId), nPrev) << 16);
        };
        t[2] = uint32_t(t.size());
        lasts(A);
        lasts(B);
        st.push_back(std::move(t));
        levelCost += nMid + 2;
      }
      for (size_t i = 0; i < vr.size(); i += 64) {
        std::vector<uint32_t> t;
        uint32_t n = std::min<size_t>(64, vr.size() - i);
        t.push_back(T_VECREAD | n << 8);
        for (uint32_t k = 0; k < n; ++k) {
          const auto &o = *vr[i + k];
          t.push_back(uint32_t(o.vecBase) | uint32_t(o.vecLength) << 16);
          t.push_back(uint32_t(o.offset) | uint32_t(o.isConstVec ? 1 : 0) << 16);
          t.push_back(uint32_t(o.index0) | uint32_t(o.index1) << 16);
          t.push_back(uint32_t(o.index2) | uint32_t(o.index3) << 16);
          t.push_back(uint32_t(o.outRangeValue) | uint32_t(o.result) << 16);
        }
        st.push_back(std::move(t));
        levelCost += 4;
      }
      for (size_t i = 0; i < vo.size(); i += 64) {
        std::vector<uint32_t> t;
        uint32_t n = std::min<size_t>(64, vo.size() - i);
        t.push_back(T_VECOP | n << 8);
        uint32_t maxLen = 0;
        for (uint32_t k = 0; k < n; ++k) {
          const auto &o = *vo[i + k];
          t.push_back(uint32_t(o.vec1Base) | uint32_t(o.vec2Base) << 16);
          t.push_back(uint32_t(o.vecLength) | uint32_t(o.opName) << 16 | uint32_t(o.isV1V2Const) << 24);
          t.push_back(o.result);
          maxLen = std::max<uint32_t>(maxLen, o.vecLength);
        }
        st.push_back(std::move(t));
        levelCost += 4 + maxLen;
      }
      for (size_t i = 0; i < mr.size(); i += 64) {
        std::vector<uint32_t> t;
        uint32_t n = std::min<size_t>(64, mr.size() - i);
        t.push_back(T_MEMREAD | n << 8);
        for (uint32_t k = 0; k < n; ++k) {
          const auto &o = *mr[i + k];
          t.push_back(uint32_t(o.memBase));
          t.push_back(uint32_t(o.memBase >> 32));
          t.push_back(uint32_t(o.en) | uint32_t(o.addrVec) << 16);
          t.push_back(uint32_t(o.result) | uint32_t(o.hasMultipleWriter ? 1 : 0) << 16);
        }
        st.push_back(std::move(t));
        levelCost += 8;
      }
      uint64_t nt = st.size();
      cost += 2 + (nt ? ((nt + nWaves - 1) / nWaves) * (levelCost / std::max<uint64_t>(nt, 1) + 1) : 0);
    }
    w[H_NLEV] = p.exec_mParts.size();
    w[H_OFF_LEV] = w.size();
    uint32_t taskIdx = 0;
    for (size_t li = 0; li < streams.size(); ++li) {
      w.push_back(taskIdx);
      taskIdx += streams[li].size();
      w.push_back(taskIdx);
    }
    w[H_OFF_TASK] = w.size();
    size_t taskTab = w.size();
    w.resize(w.size() + taskIdx, 0);
    uint32_t ti = 0;
    for (auto &lvStreams : streams)
      for (auto &t : lvStreams) {
        if (vpRows && (w.size() & 1)) w.push_back(0);  // batched: every task at an even word
        w[taskTab + ti++] = w.size();
        w.insert(w.end(), t.begin(), t.end());
      }
    // last level
    const auto &ew = p.op_last_exchangeWrite;
    const auto &rw = p.op_last_regWrite;
    w[H_EW_EXG] = ew.exchange;
    w[H_EW_DC] = uint32_t(ew.dat) | uint32_t(ew.count) << 16;
    bool exg = ew.count != 0;
    w[H_RW_REG] = rw.reg;
    w[H_RW_DC] = exg ? 0 : (uint32_t(rw.dat) | uint32_t(rw.count) << 16);
    w[H_NMW] = exg ? 0 : p.ops_last_memWrite.size();
    w[H_NPR] = exg ? 0 : p.ops_last_print.size();
    w[H_NST] = exg ? 0 : p.ops_last_stop.size();
    w[H_OFF_LAST] = w.size();
    if (!exg) {
      for (const auto &o : p.ops_last_memWrite) {
        w.push_back(uint32_t(o.memBase));
        w.push_back(uint32_t(o.memBase >> 32));
        w.push_back(uint32_t(o.addrVec) | uint32_t(o.dat) << 16);
        w.push_back(uint32_t(o.en) | uint32_t(o.hasMultipleWriter ? 1 : 0) << 16);
      }
      for (const auto &o : p.ops_last_print) w.push_back(uint32_t(o.en) | uint32_t(o.msg) << 16);
      for (const auto &o : p.ops_last_stop) w.push_back(o.en);
    }
    w[H_CONST_BASE] = img.consts.size();
    w[H_NCONST] = p.constVecPool.size();
    img.consts.insert(img.consts.end(), p.constVecPool.begin(), p.constVecPool.end());
    cost += 2 + ((exg ? ew.count : rw.count) / 4 + p.ops_last_memWrite.size()) / 64;
    img.partOff.push_back(img.words.size());
    img.partWords.push_back(w.size());
    img.partCost.push_back(cost);
    img.words.insert(img.words.end(), w.begin(), w.end());
    while (img.words.size() % 4) img.words.push_back(0);  // 16-byte aligned partitions
  }
  if (img.consts.empty()) img.consts.push_back(0);
  return img;
}

Wave64Schedule scheduleWave64(const Wave64Image &img, uint32_t groups) {
  Wave64Schedule s;
  s.groups = groups;
  std::vector<std::vector<uint32_t>> per(groups * 2);
  std::vector<uint64_t> load(groups, 0);
  std::vector<uint32_t> words(groups, 0);
  for (int r = 0; r < 2; ++r) {
    uint32_t lo = r == 0 ? 0 : img.n0, hi = r == 0 ? img.n0 : img.nTotal;
    std::vector<uint32_t> ids(hi - lo);
    std::iota(ids.begin(), ids.end(), lo);
    std::stable_sort(ids.begin(), ids.end(), [&](uint32_t a, uint32_t b) { return img.partCost[a] > img.partCost[b]; });
    // min-heap on (region load, total words, group)
    using E = std::tuple<uint64_t, uint32_t, uint32_t>;
    std::priority_queue<E, std::vector<E>, std::greater<E>> pq;
    for (uint32_t g = 0; g < groups; ++g) pq.emplace(0, words[g], g);
    for (uint32_t id : ids) {
      auto [l, wd, g] = pq.top();
      pq.pop();
      per[g * 2 + r].push_back(id);
      l += img.partCost[id];
      words[g] += img.partWords[id];
      load[g] = std::max<uint64_t>(load[g], 0) + img.partCost[id];
      pq.emplace(l, words[g], g);
    }
  }
  s.off.push_back(0);
  for (uint32_t g = 0; g < groups; ++g) {
    uint32_t lds = 0;
    for (int r = 0; r < 2; ++r) {
      for (uint32_t id : per[g * 2 + r]) {
        s.parts.push_back(id);
        s.ldsOff.push_back(lds);
        // whole 16-byte units: the kernel copies a program in 16-byte
        // pieces (its last piece padded), so the next program must start
        // past that piece, 16-byte aligned
        lds += (img.partWords[id] + 3) & ~3u;
      }
      s.off.push_back(s.parts.size());
    }
    s.maxGroupWords = std::max(s.maxGroupWords, lds);
    s.maxGroupCost = std::max(s.maxGroupCost, load[g]);
  }
  return s;
}


// ---------------------------------------------------------------------------
// Stream image

namespace {

struct StreamTask {
  uint32_t off, y;  // entry words (type filled in per wave)
  uint64_t cost;
};

// Append `rows` zeroed 64-word rows; returns the word offset of the first.
uint32_t addRows(std::vector<uint32_t> &w, uint32_t rows) {
  const uint32_t off = uint32_t(w.size());
  w.resize(w.size() + size_t(rows) * kRow, 0);
  return off;
}

// Longest-processing-time assignment of weighted items to `bins`; returns
// the bin of each item and the largest bin load.
uint64_t lpt(const std::vector<uint64_t> &cost, uint32_t bins, std::vector<uint32_t> &binOf) {
  std::vector<uint32_t> ids(cost.size());
  std::iota(ids.begin(), ids.end(), 0u);
  std::stable_sort(ids.begin(), ids.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  using E = std::pair<uint64_t, uint32_t>;
  std::priority_queue<E, std::vector<E>, std::greater<E>> pq;
  for (uint32_t b = 0; b < bins; ++b) pq.emplace(0, b);
  binOf.assign(cost.size(), 0);
  uint64_t mx = 0;
  for (uint32_t id : ids) {
    auto [l, b] = pq.top();
    pq.pop();
    binOf[id] = b;
    l += cost[id];
    mx = std::max(mx, l);
    pq.emplace(l, b);
  }
  return mx;
}

// Last level of a LUT2 task indexed by the lane that holds each value
// (the kLastDirect row): lane (shuffleId mod 32) of the micropart's half
// writes its own final value to the op's slot, so no shuffle is needed.
// An immediate shuffle id, or a lane past the last level's size (whose
// value is 0: its rows are zero words, which evaluate to 0), becomes a
// constant write placed on a lane with no value to write. Returns false,
// leaving `row` alone, when some lane's value goes to more than one slot or
// two ops of the level write one slot (their order then stays the op order).
bool directLastRow(const CGMicroPartInfo *A, const CGMicroPartInfo *B, uint32_t *row) {
  uint32_t out[64];
  std::fill(out, out + 64, 0xFFFFFFFFu);
  std::vector<std::pair<uint32_t, uint32_t>> consts;  // (slot, value)
  std::vector<uint32_t> slots;
  for (const CGMicroPartInfo *m : {A, B})
    if (m)
      for (const auto &o : m->lastLevel) slots.push_back(o.result);
  std::sort(slots.begin(), slots.end());
  if (std::adjacent_find(slots.begin(), slots.end()) != slots.end()) return false;
  for (uint32_t h = 0; h < 2; ++h) {
    const CGMicroPartInfo *m = h ? B : A;
    if (!m) continue;
    const size_t nPrev = m->middleLevels.empty() ? m->topLevel.size() : m->middleLevels.back().size();
    for (const auto &o : m->lastLevel) {
      const uint32_t s = o.shuffleId & 31u;
      if (o.shuffleId < 32 || s >= nPrev) {
        consts.emplace_back(o.result, o.shuffleId < 32 ? o.shuffleId : 0u);
        continue;
      }
      uint32_t &w = out[h * 32 + s];
      if (w != 0xFFFFFFFFu) return false;
      w = o.result;
    }
  }
  uint32_t l = 0;
  for (const auto &c : consts) {
    while (l < 64 && out[l] != 0xFFFFFFFFu) ++l;
    if (l == 64) return false;
    out[l] = c.first | (c.second + 1) << 16;
  }
  std::copy(out, out + 64, row);
  return true;
}

}  // namespace

StreamImage buildStreamImage(const SimDesignInfo &d, const StimMap &stim, uint32_t nC, uint32_t groups,
                             uint32_t slots, bool forceSlots) {
  StreamImage img;
  img.nC = nC = std::max<uint32_t>(nC, 1);
  img.groups = groups = std::max<uint32_t>(groups, 1);
  img.slots = slots = std::max<uint32_t>(std::min(slots, nC), 1);
  img.n0 = d.regionPartitionIds[0].size();
  img.nTotal = d.parts.size();
  // Layout 0: one partition on all nC compute waves. Layout 1 (slots > 1):
  // nC / slots waves per partition, partitions side by side.
  const uint32_t nLayouts = slots > 1 ? 2 : 1;
  const uint32_t wavesOf[2] = {nC, nC / slots};
  auto &W = img.words;
  addRows(W, 1);  // zero row
  std::vector<uint32_t> blockOff[2], blockRows[2];
  std::vector<std::vector<std::vector<std::vector<StreamTask>>>> waveTasks[2];  // [layout][part][level][wave]
  for (uint32_t L = 0; L < nLayouts; ++L) {
    blockOff[L].resize(d.parts.size());
    blockRows[L].resize(d.parts.size());
    waveTasks[L].resize(d.parts.size());
  }
  std::vector<uint64_t> partCost[2];  // per layout: the cost the partition assignment balances
  for (auto &v : partCost) v.assign(d.parts.size(), 0);
  for (size_t pi = 0; pi < d.parts.size(); ++pi) {
    const auto &p = d.parts[pi];
    img.maxVp = std::max(img.maxVp, p.valuePoolSize);
    for (uint32_t L = 0; L < nLayouts; ++L) waveTasks[L][pi].resize(p.exec_mParts.size());
    uint64_t cost[2] = {0, 0};
    for (size_t li = 0; li < p.exec_mParts.size(); ++li) {
      std::vector<StreamTask> tasks;
      std::vector<const CGMicroPartInfo *> luts;
      std::vector<const CGMicroPartVecRead *> vr;
      std::vector<const CGMicroPartVecArithOrLogic *> vo;
      std::vector<const CGMicroPartMemRead *> mr;
      for (const auto &m : p.exec_mParts[li]) {
        if (m.isLUTPart) {
          luts.push_back(&m);
          continue;
        }
        int k = otherKind(m);
        if (k == 1) for (const auto &o : m.vecRead) vr.push_back(&o);
        if (k == 2) for (const auto &o : m.vecArithAndLogic) vo.push_back(&o);
        if (k == 3) for (const auto &o : m.memRead) mr.push_back(&o);
      }
      std::stable_sort(luts.begin(), luts.end(), [](const CGMicroPartInfo *a, const CGMicroPartInfo *b) {
        return a->middleLevels.size() > b->middleLevels.size();
      });
      for (size_t i = 0; i < luts.size();) {
        const CGMicroPartInfo *A = luts[i];
        const CGMicroPartInfo *B = nullptr;
        if (i + 1 < luts.size() && luts[i + 1]->middleLevels.size() == A->middleLevels.size()) B = luts[i + 1];
        i += B ? 2 : 1;
        const uint32_t nMid = A->middleLevels.size(), rows = 3 + nMid;
        const uint32_t off = addRows(W, rows);
        std::fill(W.begin() + off + 2 * kRow, W.begin() + off + 3 * kRow, 0xFFFFFFFFu);
        auto put = [&](const CGMicroPartInfo *m, uint32_t base) {
          for (size_t k = 0; k < m->topLevel.size(); ++k) {
            const auto &o = m->topLevel[k];
            // a pass-through op ignores op0/op1: store immediates so no lane reads an unchecked slot
            const uint32_t o0 = o.lutIndex ? o.op0 : 0u, o1 = o.lutIndex ? o.op1 : 0u;
            W[off + base + k] = uint32_t(o.lutIndex) | o0 << 16;
            W[off + kRow + base + k] = o1 | uint32_t(o.op2) << 16;
          }
          for (size_t k = 0; k < m->lastLevel.size(); ++k)
            W[off + 2 * kRow + base + k] = uint32_t(m->lastLevel[k].result) | uint32_t(m->lastLevel[k].shuffleId) << 16;
          for (uint32_t l = 0; l < nMid; ++l)
            for (size_t k = 0; k < m->middleLevels[l].size(); ++k)
              W[off + (3 + l) * kRow + base + k] = m->middleLevels[l][k].getPacked();
        };
        put(A, 0);
        if (B) put(B, 32);
        const uint32_t direct = directLastRow(A, B, &W[off + 2 * kRow]) ? kLastDirect : 0u;
        tasks.push_back({off, T_LUT2 << 4 | nMid << 8 | direct | rows << 24, 4 + 2ull * nMid});
      }
      auto vecTasks = [&](auto &ops, uint32_t kind, uint32_t rows, auto fill, auto costOf) {
        for (size_t i = 0; i < ops.size(); i += 64) {
          const uint32_t n = std::min<size_t>(64, ops.size() - i);
          const uint32_t off = addRows(W, rows);
          uint64_t c = 0;
          for (uint32_t k = 0; k < n; ++k) {
            uint32_t v[5] = {0, 0, 0, 0, 0};
            fill(*ops[i + k], v);
            for (uint32_t r = 0; r < rows; ++r) W[off + r * kRow + k] = v[r];
            c = std::max<uint64_t>(c, costOf(*ops[i + k]));
          }
          tasks.push_back({off, kind << 4 | n << 8 | rows << 24, c});
        }
      };
      vecTasks(vr, T_VECREAD, 5,
               [](const CGMicroPartVecRead &o, uint32_t *v) {
                 v[0] = uint32_t(o.vecBase) | uint32_t(o.vecLength) << 16;
                 v[1] = uint32_t(o.offset) | uint32_t(o.isConstVec ? 1 : 0) << 16;
                 v[2] = uint32_t(o.index0) | uint32_t(o.index1) << 16;
                 v[3] = uint32_t(o.index2) | uint32_t(o.index3) << 16;
                 v[4] = uint32_t(o.outRangeValue) | uint32_t(o.result) << 16;
               },
               [](const CGMicroPartVecRead &) { return uint64_t(6); });
      vecTasks(vo, T_VECOP, 3,
               [](const CGMicroPartVecArithOrLogic &o, uint32_t *v) {
                 v[0] = uint32_t(o.vec1Base) | uint32_t(o.vec2Base) << 16;
                 v[1] = uint32_t(o.vecLength) | uint32_t(o.opName) << 16 | uint32_t(o.isV1V2Const) << 24;
                 v[2] = o.result;
               },
               [](const CGMicroPartVecArithOrLogic &o) { return uint64_t(6 + o.vecLength); });
      vecTasks(mr, T_MEMREAD, 4,
               [](const CGMicroPartMemRead &o, uint32_t *v) {
                 v[0] = uint32_t(o.memBase);
                 v[1] = uint32_t(o.memBase >> 32);
                 v[2] = uint32_t(o.en) | uint32_t(o.addrVec) << 16;
                 v[3] = uint32_t(o.result) | uint32_t(o.hasMultipleWriter ? 1 : 0) << 16;
               },
               [](const CGMicroPartMemRead &) { return uint64_t(16); });
      img.tasks += tasks.size();
      std::vector<uint64_t> tc(tasks.size());
      for (size_t t = 0; t < tasks.size(); ++t) tc[t] = tasks[t].cost;
      for (uint32_t L = 0; L < nLayouts; ++L) {
        std::vector<uint32_t> binOf;
        const uint64_t span = lpt(tc, wavesOf[L], binOf);
        cost[L] += 2 + span;
        auto &lv = waveTasks[L][pi][li];
        lv.resize(wavesOf[L]);
        for (size_t t = 0; t < tasks.size(); ++t) lv[binOf[t]].push_back(tasks[t]);
      }
    }
    // partition block: header row, L0 rows, last-level lists
    std::vector<uint32_t> src, slot;
    for (const auto &o : p.ops_l0_regRead) {
      const int32_t in = stim.lookup(o.reg);
      src.push_back(in >= 0 ? uint32_t(in) | 2u << 30 : o.reg);
      slot.push_back(o.result);
    }
    for (const auto &o : p.ops_l0_exchangeRead) {
      src.push_back(o.exchange | 1u << 30);
      slot.push_back(o.result);
    }
    const uint32_t nL0 = src.size();
    const auto &ew = p.op_last_exchangeWrite;
    const auto &rw = p.op_last_regWrite;
    const bool exg = ew.count != 0;
    const uint32_t constBase = img.consts.size();
    img.consts.insert(img.consts.end(), p.constVecPool.begin(), p.constVecPool.end());
    const uint32_t lastWords = 4 * p.ops_last_memWrite.size() + p.ops_last_print.size() + p.ops_last_stop.size();
    const uint32_t lastRows = (lastWords + kRow - 1) / kRow;
    uint32_t nJ0 = 0;
    for (uint32_t L = 0; L < nLayouts; ++L) {
      const uint32_t nWaves = wavesOf[L], nT = nWaves * 64;
      const uint32_t nJ = (nL0 + nT - 1) / nT;
      if (L == 0) nJ0 = nJ;
      const uint32_t section = 1 + 2 * nJ;
      const uint32_t off = addRows(W, nWaves * section + lastRows);
      blockOff[L][pi] = off;
      blockRows[L][pi] = section;
      for (uint32_t w = 0; w < nWaves; ++w)
        for (uint32_t j = 0; j < nJ; ++j)
          for (uint32_t l = 0; l < 64; ++l) {
            const uint32_t k = j * nT + w * 64 + l;
            W[off + (w * section + 1 + 2 * j) * kRow + l] = k < nL0 ? src[k] : kL0None;
            W[off + (w * section + 2 + 2 * j) * kRow + l] = k < nL0 ? slot[k] : 0u;
          }
      uint32_t h[kRow] = {};
      h[S_VP] = p.valuePoolSize;
      h[S_NJ] = nJ;
      h[S_RW_REG] = rw.reg;
      h[S_RW_DC] = exg ? 0 : (uint32_t(rw.dat) | uint32_t(rw.count) << 16);
      h[S_EW_EXG] = ew.exchange;
      h[S_EW_DC] = uint32_t(ew.dat) | uint32_t(ew.count) << 16;
      h[S_NMW] = exg ? 0 : p.ops_last_memWrite.size();
      h[S_NPR] = exg ? 0 : p.ops_last_print.size();
      h[S_NST] = exg ? 0 : p.ops_last_stop.size();
      h[S_OFF_LAST] = nWaves * section * kRow;
      h[S_CONST_BASE] = constBase;
      h[S_PART_ID] = pi;
      for (uint32_t w = 0; w < nWaves; ++w) std::copy(h, h + kRow, W.begin() + off + w * section * kRow);
      uint32_t q = off + h[S_OFF_LAST];
      if (!exg) {
        for (const auto &o : p.ops_last_memWrite) {
          W[q++] = uint32_t(o.memBase);
          W[q++] = uint32_t(o.memBase >> 32);
          W[q++] = uint32_t(o.addrVec) | uint32_t(o.dat) << 16;
          W[q++] = uint32_t(o.en) | uint32_t(o.hasMultipleWriter ? 1 : 0) << 16;
        }
        for (const auto &o : p.ops_last_print) W[q++] = uint32_t(o.en) | uint32_t(o.msg) << 16;
        for (const auto &o : p.ops_last_stop) W[q++] = o.en;
      }
    }
    for (uint32_t L = 0; L < nLayouts; ++L) partCost[L][pi] = cost[L] + 8 + 4 * nJ0 + (p.ops_last_memWrite.empty() ? 0 : 8);
  }
  if (img.consts.empty()) img.consts.push_back(0);
  // Estimated time of a partition's steps in each layout, in LUT exec-level
  // steps (measured on the streamed kernel: a LUT level ~1 us, partition
  // begin ~2.4 us, end ~0.75 us, a level slice holding word ops ~1.9 us).
  constexpr double kBeginSteps = 2.4, kEndSteps = 0.75, kWordSteps = 1.9, kBarrierSteps = 2.0;
  // Per-step cost of a round whose partitions run side by side (measured on
  // the 10M-LUT4 design: 1.12 for two slots, 1.18 for four).
  const double splitFactor = slots == 2 ? 1.12 : 1.18;
  std::vector<double> estOf[2];
  for (uint32_t L = 0; L < nLayouts; ++L) {
    estOf[L].assign(d.parts.size(), 0.0);
    for (size_t pi = 0; pi < d.parts.size(); ++pi) {
      double e = kBeginSteps + kEndSteps;
      for (const auto &lv : waveTasks[L][pi]) {
        size_t nSub = 1;
        for (const auto &mine : lv) nSub = std::max(nSub, mine.size());
        for (size_t q = 0; q < nSub; ++q) {
          bool word = false;
          for (const auto &mine : lv) word |= q < mine.size() && ((mine[q].y >> 4) & 7u) != T_LUT2;
          e += word ? kWordSteps : 1.0;
        }
      }
      estOf[L][pi] = e;
    }
  }
  // partitions -> (slot, workgroup) bins per region, slot-major so that every
  // workgroup's first slot fills before any second slot; each region takes
  // the slot count (one or `slots`) with the shorter estimated critical path
  // (regions end in a grid barrier, so each costs its slowest workgroup)
  const size_t bins = size_t(groups) * slots;
  auto binOfSlot = [&](uint32_t g, uint32_t s) { return size_t(s) * groups + g; };
  std::vector<std::vector<uint32_t>> per(bins * 2);
  img.estCycle = 2 * kBarrierSteps;
  for (int r = 0; r < 2; ++r) {
    const auto &ids = d.regionPartitionIds[r];
    struct Option {
      uint32_t slots;
      std::vector<uint32_t> binOf;
      uint64_t maxCost;
      double time;
    };
    std::vector<Option> opts;
    for (uint32_t s : {slots, 1u}) {
      if (!opts.empty() && (s == opts[0].slots || forceSlots)) continue;
      Option o{s, {}, 0, 0.0};
      std::vector<uint64_t> c(ids.size());
      for (size_t k = 0; k < ids.size(); ++k) c[k] = partCost[s > 1 ? 1 : 0][ids[k]];
      o.maxCost = lpt(c, groups * s, o.binOf);
      std::vector<std::vector<uint32_t>> lists(size_t(groups) * s);
      for (size_t k = 0; k < ids.size(); ++k) lists[o.binOf[k]].push_back(ids[k]);
      for (uint32_t g = 0; g < groups; ++g) {
        double t = 0;
        for (size_t k = 0;; ++k) {
          double mx = 0;
          uint32_t occupied = 0, only = 0;
          for (uint32_t q = 0; q < s; ++q) {
            const auto &lst = lists[binOfSlot(g, q)];
            if (k >= lst.size()) continue;
            ++occupied;
            only = lst[k];
            mx = std::max(mx, estOf[nLayouts - 1][lst[k]]);
          }
          if (!occupied) break;
          t += occupied > 1 ? splitFactor * mx : estOf[0][only];
        }
        o.time = std::max(o.time, t);
      }
      opts.push_back(std::move(o));
    }
    const Option &best = opts.size() > 1 && opts[1].time <= opts[0].time ? opts[1] : opts[0];
    img.regionSlots[r] = best.slots;
    img.maxGroupCost += best.maxCost;
    img.estCycle += best.time;
    for (size_t k = 0; k < ids.size(); ++k) per[size_t(best.binOf[k]) * 2 + r].push_back(ids[k]);
  }
  const uint32_t rec = 1 + 2 * nC;
  // region steps: the same entry for every wave
  auto step = [&](uint32_t type, uint32_t x, uint32_t y) {
    img.steps.push_back(type);
    for (uint32_t w = 0; w < nC; ++w) {
      img.steps.push_back(x);
      img.steps.push_back(y);
    }
  };
  const uint32_t none = kKindNone << 4 | 1u << 24;
  const uint32_t noPart = ~0u;
  for (uint32_t g = 0; g < groups; ++g) {
    const uint32_t base = img.steps.size() / rec;
    for (int r = 0; r < 2; ++r) {
      size_t rounds = 0;
      for (uint32_t s = 0; s < slots; ++s) rounds = std::max(rounds, per[binOfSlot(g, s) * 2 + r].size());
      for (size_t k = 0; k < rounds; ++k) {
        // the partitions of this round: one per slot (or none); a round with
        // a single partition runs it on all the compute waves
        std::vector<uint32_t> part;
        for (uint32_t s = 0; s < slots; ++s) {
          const auto &lst = per[binOfSlot(g, s) * 2 + r];
          part.push_back(k < lst.size() ? lst[k] : noPart);
        }
        uint32_t occupied = 0;
        for (uint32_t pi : part) occupied += pi != noPart;
        const uint32_t L = occupied > 1 ? 1 : 0;
        const uint32_t ns = L ? slots : 1, nWaves = wavesOf[L];
        const uint32_t flag = L ? kSplit : 0u;
        if (!L) {  // the single partition goes to slot 0 on every wave
          uint32_t only = noPart;
          for (uint32_t pi : part)
            if (pi != noPart) only = pi;
          part.assign(1, only);
        }
        img.steps.push_back(S_PART);
        for (uint32_t w = 0; w < nC; ++w) {
          const uint32_t pi = part[w / nWaves];
          img.steps.push_back(pi == noPart ? 0u : blockOff[L][pi] + (w % nWaves) * blockRows[L][pi] * kRow);
          img.steps.push_back((pi == noPart ? none : std::min<uint32_t>(blockRows[L][pi], 255) << 24) | flag);
        }
        // each slot's level steps (a level wider than the slot's waves takes several), merged
        std::vector<std::vector<std::vector<const StreamTask *>>> subs(ns);
        size_t nSteps = 0;
        for (uint32_t s = 0; s < ns; ++s) {
          if (part[s] == noPart) continue;
          for (const auto &lv : waveTasks[L][part[s]]) {
            size_t nSub = 1;
            for (const auto &mine : lv) nSub = std::max(nSub, mine.size());
            for (size_t q = 0; q < nSub; ++q) {
              std::vector<const StreamTask *> st(nWaves, nullptr);
              for (uint32_t wi = 0; wi < nWaves; ++wi)
                if (q < lv[wi].size()) st[wi] = &lv[wi][q];
              subs[s].push_back(std::move(st));
            }
          }
          nSteps = std::max(nSteps, subs[s].size());
        }
        for (size_t i = 0; i < nSteps; ++i) {
          img.steps.push_back(S_LEVEL);
          for (uint32_t w = 0; w < nC; ++w) {
            const uint32_t s = w / nWaves;
            const StreamTask *t = i < subs[s].size() ? subs[s][i][w % nWaves] : nullptr;
            img.steps.push_back(t ? t->off : 0u);
            img.steps.push_back((t ? t->y : none) | flag);
          }
        }
        img.steps.push_back(S_PEND);
        for (uint32_t w = 0; w < nC; ++w) {
          const uint32_t pi = part[w / nWaves];
          img.steps.push_back(pi == noPart ? 0u : blockOff[L][pi]);
          img.steps.push_back((pi == noPart ? none : 1u << 24) | flag);
        }
      }
      step(S_REGION, 0, none);
    }
    img.stepBase.push_back(base);
    img.stepCount.push_back(img.steps.size() / rec - base);
    img.maxSteps = std::max(img.maxSteps, img.stepCount.back());
  }
  addRows(W, kStreamRows);  // margin past the last task
  return img;
}

}  // namespace tgs
