// Host interpreter of the stream image, for CPU tests of the image builder.
//
// It walks every workgroup's step stream exactly as the streamed kernel
// does (lanes, rows, clamped row addressing, L0 rows, last-level lists) and
// checks every index against the buffer it reads or writes, then writes the
// final register and memory pools so a test can compare them with the
// oracle. It is not linked into the runtime and never runs on the device path.
//
// Usage: stream_check <design.bin> <inputs.u32> <nc> <groups> <cycles> <seed> <out.bin> [init.bin] [slots]
// (slots < 0: up to -slots, chosen per region by the builder as the runtime does)
// init.bin, if given, holds the initial register pool followed by the memory pool.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "image.hpp"

using namespace toucanGPUSim;
using namespace tgs;

namespace {

[[noreturn]] void die(const std::string &m) {
  std::fprintf(stderr, "stream_check: %s\n", m.c_str());
  std::exit(1);
}

struct Buf {
  std::vector<uint8_t> &v;
  const char *name;
  uint8_t &at(uint64_t i) {
    if (i >= v.size()) die(std::string("out of bounds in ") + name + " at " + std::to_string(i));
    return v[i];
  }
};

struct Interp {
  const StreamImage &im;
  const SimDesignInfo &d;
  uint32_t nc, nTc;
  std::vector<uint8_t> reg, mem, exg, vp, lut;  // vp: the value pool of the slot being evaluated
  std::vector<uint64_t> draws;
  uint32_t vpBytes;

  uint32_t word(uint64_t i) const {
    if (i >= im.words.size()) die("row read past the image at " + std::to_string(i));
    return im.words[i];
  }
  uint8_t &V(uint32_t s) { return Buf{vp, "value pool"}.at(s); }
  uint32_t rdOp(uint32_t s) { return s < 16 ? s : V(s); }
  uint32_t lutAt(uint32_t pos) { return pos < kLutSize ? lut[pos] : 0u; }
  static uint32_t lutIndex(uint32_t lt, uint32_t a, uint32_t b, uint32_t c) {
    return (lt + (((a & 0xFFFFu) << 8) | (b << 4) | c)) & 0xFFFFu;
  }
  // row r of compute wave w in a step record, as the loader copies it: the
  // loader always copies rows 0..5 from x, so they must lie in the image
  uint32_t rowWord(const uint32_t *rec, uint32_t w, uint32_t r, uint32_t lane) const {
    const uint32_t ex = rec[1 + 2 * w];
    if (ex % kRow) die("entry not row aligned");
    if (uint64_t(ex) + kStreamRows * kRow > im.words.size()) die("loader rows past the image");
    if (r >= kStreamRows) die("row past the staged rows");
    return word(uint64_t(ex) + r * kRow + lane);
  }

  void lutPair(const uint32_t *rec, uint32_t w) {
    const uint32_t ey = rec[2 + 2 * w], ex = rec[1 + 2 * w];
    const uint32_t nMid = (ey >> 8) & 0xFFu;
    if ((ey >> 16) & 0x7Fu) die("LUT task entry bits 16-22 must be clear");
    uint32_t val[64], rows[64][kStreamRows];
    if (3 + nMid != (ey >> 24) && (ey >> 24) != 255) die("LUT task row count");
    for (uint32_t l = 0; l < 64; ++l)
      for (uint32_t r = 0; r < kStreamRows; ++r) rows[l][r] = r < 3 + nMid ? rowWord(rec, w, r, l) : 0xDEADBEEFu;
    for (uint32_t l = 0; l < 64; ++l) {
      const uint32_t x = rows[l][0], y = rows[l][1];
      const uint32_t lt = x & 0xFFFFu, o0 = x >> 16, o1 = y & 0xFFFFu, o2 = y >> 16;
      const uint32_t c = rdOp(o2);
      val[l] = lt == 0 ? c : lutAt(lutIndex(lt, rdOp(o0), rdOp(o1), c));
    }
    for (uint32_t m = 0; m < nMid; ++m) {
      uint32_t nv[64];
      for (uint32_t l = 0; l < 64; ++l) {
        const uint32_t wd = m < 3 ? rows[l][3 + m] : word(uint64_t(ex) + (3 + m) * kRow + l);
        const uint32_t hb = l & 32u;
        const uint32_t o2 = wd & 63u, o1 = (wd >> 6) & 63u, o0 = (wd >> 12) & 63u, lt = wd >> 18;
        const uint32_t a = o0 < 32 ? o0 : val[(o0 & 31u) | hb], b = o1 < 32 ? o1 : val[(o1 & 31u) | hb];
        const uint32_t c = o2 < 32 ? o2 : val[(o2 & 31u) | hb];
        nv[l] = lt == 0 ? c : lutAt(lutIndex(lt, a, b, c));
      }
      std::memcpy(val, nv, sizeof val);
    }
    uint32_t out[64];
    bool wr[64];
    const bool direct = (ey & kLastDirect) != 0;
    for (uint32_t l = 0; l < 64; ++l) {
      const uint32_t wl = rows[l][2], sid = (wl >> 16) & 0xFFu;
      if (direct) {
        const uint32_t k = wl >> 16;
        if (wl != 0xFFFFFFFFu && k > 32) die("direct last-level constant out of range");
        out[l] = k == 0 ? val[l] : k - 1;
      } else {
        out[l] = sid < 32 ? sid : val[(sid & 31u) | (l & 32u)];
      }
      wr[l] = wl != 0xFFFFFFFFu;
    }
    for (uint32_t l = 0; l < 64; ++l)
      if (wr[l]) V(rows[l][2] & 0xFFFFu) = uint8_t(out[l]);
  }

  void vecRead(const uint32_t *rec, uint32_t w, const uint8_t *consts, size_t nConsts) {
    const uint32_t n = (rec[2 + 2 * w] >> 8) & 0xFFFFu;
    for (uint32_t l = 0; l < n && l < 64; ++l) {
      uint32_t r[5];
      for (uint32_t k = 0; k < 5; ++k) r[k] = rowWord(rec, w, k, l);
      const uint32_t base = r[0] & 0xFFFFu, len = r[0] >> 16, offset = r[1] & 0xFFFFu, isConst = r[1] >> 16;
      const uint32_t off = ((rdOp(r[2] & 0xFFFFu) << 12) | (rdOp(r[2] >> 16) << 8) | (rdOp(r[3] & 0xFFFFu) << 4) |
                            rdOp(r[3] >> 16)) + offset;
      uint32_t v;
      if (off < len) {
        if (isConst) {
          if (base + off >= nConsts) die("const read out of range");
          v = consts[base + off];
        } else {
          v = V(base + off);
        }
      } else {
        v = rdOp(r[4] & 0xFFFFu);
      }
      V(r[4] >> 16) = uint8_t(v);
    }
  }

  void vecOp(const uint32_t *rec, uint32_t w, const uint8_t *consts, size_t nConsts) {
    const uint32_t n = (rec[2 + 2 * w] >> 8) & 0xFFFFu;
    for (uint32_t l = 0; l < n && l < 64; ++l) {
      uint32_t r[3];
      for (uint32_t k = 0; k < 3; ++k) r[k] = rowWord(rec, w, k, l);
      const uint32_t b1 = r[0] & 0xFFFFu, b2 = r[0] >> 16;
      const uint32_t len = r[1] & 0xFFFFu, name = (r[1] >> 16) & 0xFFu, cf = r[1] >> 24, res = r[2];
      unsigned __int128 v1 = 0, v2 = 0;
      for (uint32_t k = 0; k < len; ++k) {
        auto cst = [&](uint32_t i) -> unsigned __int128 {
          if (i >= nConsts) die("const read out of range");
          return consts[i];
        };
        unsigned __int128 s1 = (cf & 2) ? cst(b1 + k) : (unsigned __int128)V(b1 + k);
        unsigned __int128 s2 = (cf & 1) ? cst(b2 + k) : (unsigned __int128)V(b2 + k);
        v1 |= s1 << (4 * k);
        v2 |= s2 << (4 * k);
      }
      if (len > 0 && len < 32) {
        const uint32_t bit = 4 * len - 1;
        const unsigned __int128 ext = ~((((unsigned __int128)1) << (bit + 1)) - 1);
        if ((v1 >> bit) & 1) v1 |= ext;
        if ((v2 >> bit) & 1) v2 |= ext;
      }
      if (name <= 2) {
        unsigned __int128 rr = name == 0 ? v1 + v2 : name == 1 ? v1 - v2 : v1 * v2;
        for (uint32_t k = 0; k < len; ++k) {
          V(uint16_t(res + k)) = uint8_t(rr & 15u);
          rr >>= 4;
        }
      } else {
        const __int128 s1 = (__int128)v1, s2 = (__int128)v2;
        V(res) = uint8_t(name == 3 ? (s1 == s2) : name == 4 ? (s1 < s2) : name == 5 ? (s1 <= s2) : 0);
      }
    }
  }

  uint64_t memAddr(uint32_t addrVec, uint32_t multi, uint64_t base) {
    uint32_t addr = 0;
    for (uint32_t j = 0; j < 8; ++j) addr |= uint32_t(V(addrVec + j)) << (4 * j);
    if (multi) addr <<= 2;
    return base + addr;
  }

  void memRead(const uint32_t *rec, uint32_t w) {
    const uint32_t n = (rec[2 + 2 * w] >> 8) & 0xFFFFu;
    for (uint32_t l = 0; l < n && l < 64; ++l) {
      uint32_t r[4];
      for (uint32_t k = 0; k < 4; ++k) r[k] = rowWord(rec, w, k, l);
      if (rdOp(r[2] & 0xFFFFu) == 0) continue;
      const uint64_t at = memAddr(r[2] >> 16, r[3] >> 16, uint64_t(r[0]) | uint64_t(r[1]) << 32);
      V(r[3] & 0xFFFFu) = at < d.memPoolSize ? mem[at] : 0;
    }
  }

  // L0 of the partition on waves [w0, w0 + nw)
  void partBegin(const uint32_t *rec, uint32_t w0, uint32_t nw, uint32_t &constBase) {
    const uint32_t nJ = rowWord(rec, w0, 0, S_NJ);
    constBase = rowWord(rec, w0, 0, S_CONST_BASE);
    for (uint32_t w = w0 + 1; w < w0 + nw; ++w)
      for (uint32_t k = 0; k < S_WORDS; ++k)
        if (rowWord(rec, w, 0, k) != rowWord(rec, w0, 0, k)) die("partition header differs between waves");
    for (uint32_t t = 0; t < nw * 64; ++t) {
      const uint32_t w = w0 + t / 64, l = t % 64, ex = rec[1 + 2 * w];
      for (uint32_t j = 0; j < nJ; ++j) {
        uint32_t src, slot;
        if (j < 2) {
          src = rowWord(rec, w, 1 + 2 * j, l);
          slot = rowWord(rec, w, 2 + 2 * j, l);
        } else {
          src = word(uint64_t(ex) + (1 + 2 * j) * kRow + l);
          slot = word(uint64_t(ex) + (2 + 2 * j) * kRow + l);
        }
        const uint32_t kind = src >> 30, idx = src & 0x3FFFFFFFu;
        if (kind == 3) continue;
        uint32_t v;
        if (kind == 0) v = Buf{reg, "register pool"}.at(idx);
        else if (kind == 1) v = Buf{exg, "exchange pool"}.at(idx);
        else v = uint32_t((draws.at(idx >> 6) >> (idx & 63)) & 1u);
        V(slot) = uint8_t(v);
      }
    }
  }

  void partEnd(const uint32_t *rec, uint32_t w0, uint64_t cycle, std::vector<std::pair<uint64_t, uint32_t>> &log,
               bool &stop) {
    auto h = [&](uint32_t k) { return rowWord(rec, w0, 0, k); };
    const uint32_t ex = rec[1 + 2 * w0];
    const uint32_t ewdc = h(S_EW_DC);
    auto copyWords = [&](std::vector<uint8_t> &pool, const char *name, uint32_t dst, uint32_t dat, uint32_t cnt) {
      const uint32_t words = (cnt + 3) / 4;
      for (uint32_t k = 0; k < 4 * words; ++k)
        Buf{pool, name}.at((dst >> 2) * 4 + k) = V((dat >> 2) * 4 + k);
    };
    if (ewdc >> 16) {
      copyWords(exg, "exchange pool", h(S_EW_EXG), ewdc & 0xFFFFu, ewdc >> 16);
      return;
    }
    const uint32_t rwdc = h(S_RW_DC);
    if (rwdc >> 16) copyWords(reg, "register pool", h(S_RW_REG), rwdc & 0xFFFFu, rwdc >> 16);
    const uint32_t nmw = h(S_NMW), npr = h(S_NPR), nst = h(S_NST);
    uint64_t L = uint64_t(ex) + h(S_OFF_LAST);
    for (uint32_t k = 0; k < nmw; ++k, L += 4) {
      const uint32_t o0 = word(L), o1 = word(L + 1), o2 = word(L + 2), o3 = word(L + 3);
      if (rdOp(o3 & 0xFFFFu) == 0) continue;
      const uint32_t dat = rdOp(o2 >> 16);
      const uint64_t at = memAddr(o2 & 0xFFFFu, o3 >> 16, uint64_t(o0) | uint64_t(o1) << 32);
      if (at < d.memPoolSize) mem[at] = uint8_t(dat);
    }
    for (uint32_t k = 0; k < npr; ++k, ++L)
      if (rdOp(word(L) & 0xFFFFu) != 0) log.emplace_back(cycle, word(L) >> 16);
    for (uint32_t k = 0; k < nst; ++k, ++L)
      if (rdOp(word(L) & 0xFFFFu) != 0) stop = true;
  }
};

std::vector<uint8_t> slurp(const char *p) {
  std::ifstream f(p, std::ios::binary);
  if (!f) die(std::string("cannot open ") + p);
  return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

uint64_t xorshift64(uint64_t x) {
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  return x;
}

uint64_t splitmix(uint64_t seed) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return z ? z : 0x9E3779B97F4A7C15ull;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 8) die("usage: stream_check design.bin inputs.u32 nc groups cycles seed out.bin");
  auto bin = slurp(argv[1]);
  auto inb = slurp(argv[2]);
  const uint32_t nc = std::atoi(argv[3]), groups = std::atoi(argv[4]), cycles = std::atoi(argv[5]);
  const uint64_t seed = std::strtoull(argv[6], nullptr, 10);
  SimDesignInfo d;
  std::string err;
  if (!parseSimDesignInfo(bin.data(), bin.size(), d, &err)) die("parse: " + err);
  d.Init();
  if (validateProgram(d, &err) != TGS_OK) die("invalid: " + err);
  std::vector<uint32_t> ids(inb.size() / 4);
  std::memcpy(ids.data(), inb.data(), ids.size() * 4);
  StimMap stim;
  if (buildStimMap(d, ids.data(), ids.size(), &stim, &err) != TGS_OK) die("stimulus: " + err);
  const uint32_t slots = argc > 9 ? std::abs(std::atoi(argv[9])) : 1;
  // slots given: every region laid out for them (a negative count lets the
  // builder choose one or -slots per region, as the runtime does)
  const bool force = argc <= 9 || std::atoi(argv[9]) > 0;
  StreamImage im = buildStreamImage(d, stim, nc, groups, slots, force);
  if (im.slots != slots || nc % slots) die("slots must divide the compute waves");
  if (im.stepBase.size() != im.groups || im.stepCount.size() != im.groups) die("step tables");
  if (im.words.size() % kRow) die("image is not a whole number of rows");
  Interp it{im, d, nc, nc * 64};
  it.reg = d.regPool;
  it.reg.resize(d.regPoolSize, 0);
  it.mem = d.memPool;
  it.mem.resize(d.memPoolSize, 0);
  it.exg.assign(d.exchangePoolSize, 0);
  if (argc > 8 && std::strcmp(argv[8], "-") != 0) {
    auto init = slurp(argv[8]);
    if (init.size() != size_t(d.regPoolSize) + d.memPoolSize) die("init.bin must hold the register and memory pools");
    std::copy(init.begin(), init.begin() + d.regPoolSize, it.reg.begin());
    std::copy(init.begin() + d.regPoolSize, init.end(), it.mem.begin());
  }
  it.lut = d.lut;
  it.lut.resize(kLutSize, 0);
  it.vpBytes = std::max<uint32_t>(im.maxVp, 16);
  const uint32_t rec = 1 + 2 * nc;
  uint64_t xs = splitmix(seed);
  std::vector<std::pair<uint64_t, uint32_t>> log;
  bool stop = false;
  for (uint32_t cy = 0; cy < cycles; ++cy) {
    it.draws.assign((ids.size() + 63) / 64, 0);
    for (auto &x : it.draws) x = xs = xorshift64(xs);
    for (size_t i = 0; i < ids.size(); ++i) it.reg[ids[i]] = uint8_t((it.draws[i >> 6] >> (i & 63)) & 1u);
    // region by region; the groups of a region are independent
    for (int r = 0; r < 2; ++r) {
      for (uint32_t g = 0; g < im.groups; ++g) {
        // find this group's steps of region r: regions end with S_REGION
        int region = 0;
        std::vector<uint32_t> constBase(slots, 0);
        std::vector<std::vector<uint8_t>> vps(slots, std::vector<uint8_t>(it.vpBytes, 0xEE));
        auto kindOf = [&](const uint32_t *R, uint32_t w) { return (R[2 + 2 * w] >> 4) & 7u; };
        for (uint32_t s = 0; s < im.stepCount[g]; ++s) {
          const uint32_t *R = &im.steps.at(size_t(im.stepBase[g] + s) * rec);
          const uint32_t type = R[0];
          if (type == S_REGION) {
            ++region;
            continue;
          }
          if (region != r) continue;
          // split steps: `slots` partitions side by side on nc / slots waves
          // each; otherwise one partition on all waves (slot 0)
          const bool split = (R[2] & kSplit) != 0;
          for (uint32_t w = 1; w < nc; ++w)
            if (((R[2 + 2 * w] & kSplit) != 0) != split) die("waves disagree on the split flag");
          const uint32_t nsl = split ? slots : 1, ws = nc / nsl;
          if (split && slots == 1) die("split step with one slot");
          for (uint32_t sl = 0; sl < nsl; ++sl) {
            const uint32_t w0 = sl * ws;
            // a slot's waves agree on whether the slot has a partition in this step
            if (type != S_LEVEL)
              for (uint32_t w = w0 + 1; w < w0 + ws; ++w)
                if ((kindOf(R, w) == kKindNone) != (kindOf(R, w0) == kKindNone)) die("slot waves disagree");
            it.vp.swap(vps[sl]);
            if (type == S_PART) {
              if (kindOf(R, w0) != kKindNone) {
                it.vp.assign(it.vpBytes, 0xEE);  // uninitialized every visit
                it.partBegin(R, w0, ws, constBase[sl]);
              }
            } else if (type == S_PEND) {
              if (kindOf(R, w0) != kKindNone) it.partEnd(R, w0, cy, log, stop);
            } else if (type == S_LEVEL) {
              const uint8_t *consts = im.consts.data() + constBase[sl];
              const size_t nConsts = im.consts.size() - constBase[sl];
              for (uint32_t w = w0; w < w0 + ws; ++w) {
                const uint32_t kind = kindOf(R, w);
                if (kind == T_LUT2) it.lutPair(R, w);
                else if (kind == T_VECREAD) it.vecRead(R, w, consts, nConsts);
                else if (kind == T_VECOP) it.vecOp(R, w, consts, nConsts);
                else if (kind == T_MEMREAD) it.memRead(R, w);
                else if (kind != kKindNone) die("unknown task kind");
              }
            } else {
              die("unknown step type");
            }
            it.vp.swap(vps[sl]);
          }
        }
        if (region != 2) die("a group's stream must hold two region barriers");
      }
    }
  }
  std::ofstream o(argv[7], std::ios::binary);
  o.write(reinterpret_cast<const char *>(it.reg.data()), it.reg.size());
  o.write(reinterpret_cast<const char *>(it.mem.data()), it.mem.size());
  std::printf("groups %u maxSteps %u words %zu tasks %llu stop %d prints %zu\n", im.groups, im.maxSteps,
              im.words.size(), (unsigned long long)im.tasks, int(stop), log.size());
  return 0;
}
