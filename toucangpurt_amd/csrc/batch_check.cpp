// Host check of the batched image's LUT tasks, for CPU tests of the image
// builder.
//
// It builds the batched image twice for one instance count I per wave: with
// LUT pairs as T_LUTROWS tasks (bounds-checked lookups) and with T_LUTFAST
// tasks where the builder allows them, then, partition by partition, runs
// the LUT tasks of both images on a model of one wave's LDS (value-pool
// rows, level rows, constant rows, I instance columns, the LUT region) the
// way the batched kernel's lane rounds do, from the same random nibble
// state, and requires equal value pools after every exec level. Every offset a
// T_LUTFAST op reads or writes is checked against the wave's rows and the
// LUT region, and every operand it reads must be a nibble. Level rows start
// with different random nibbles in the two models, so a row read before the
// task wrote it shows up as a difference. Not linked into the runtime.
//
// Usage: batch_check <design.bin> <I> <seed> [rounds]
// Prints "tasks <fast> <rows>" and exits 0 when every task agrees.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "image.hpp"

using namespace toucanGPUSim;
using namespace tgs;

namespace {

[[noreturn]] void die(const std::string &m) {
  std::fprintf(stderr, "batch_check: %s\n", m.c_str());
  std::exit(1);
}

std::vector<uint8_t> slurp(const char *path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) die(std::string("cannot read ") + path);
  return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

// One wave's LDS rows: [row][RB] bytes, rows = V value-pool rows then
// kLevelRows level rows (constant rows at kConstRow hold 0..31). RB = I (a
// byte per instance) or, with nibble rows, I / 2 (instances 2j and 2j + 1 in
// the low and high nibble of byte j).
struct Area {
  uint32_t V, I;
  bool nib;
  std::vector<uint8_t> b;
  uint32_t rb() const { return nib ? I / 2 : I; }
  uint8_t &at(uint64_t off, const char *what) {
    if (off >= b.size()) die(std::string("offset out of the wave's rows: ") + what + " " + std::to_string(off));
    return b[off];
  }
  // value of instance ii in the row at byte offset `rowOff`
  uint32_t get(uint64_t rowOff, uint32_t ii, const char *what) {
    if (!nib) return at(rowOff + ii, what);
    return (at(rowOff + ii / 2, what) >> (4 * (ii & 1u))) & 15u;
  }
  void set(uint64_t rowOff, uint32_t ii, uint32_t v, const char *what) {
    if (!nib) {
      at(rowOff + ii, what) = uint8_t(v);
      return;
    }
    if (v > 15) die(std::string("value above 15 for a nibble row: ") + what);
    uint8_t &x = at(rowOff + ii / 2, what);
    x = uint8_t((ii & 1u) ? (x & 0x0Fu) | (v << 4) : (x & 0xF0u) | v);
  }
};

struct Model {
  const std::vector<uint32_t> &words;
  uint64_t base;   // first word of the task
  uint64_t limit;  // words of the partition
  std::vector<uint8_t> lut;  // LUT region: table, zeros, identity bytes
  uint32_t word(uint64_t i) const {
    if (i >= limit) die("task word past its partition at " + std::to_string(i));
    return words[base + i];
  }
};

uint32_t minOp(uint32_t q, uint32_t n) { return q < n ? q : n - 1; }

// T_LUTROWS, as lutTask: rows relative to the value pool, level rows
// relative to row V, the reference's 16-bit index and table bound.
void runRows(Model &m, Area &A) {
  if (A.nib) die("T_LUTROWS task on nibble rows");
  const uint32_t I = A.I, W = 64 / I, V = A.V;
  const uint32_t w0 = m.word(0), w1 = m.word(1);
  const uint32_t nMid = (w0 >> 8) & 0xFFFFu, nt = w1 & 0xFFFFu, nl = w1 >> 16;
  const uint32_t topOff = 4 + ((((nMid + 3) >> 2) + 1) & ~1u);
  const uint32_t midOff = m.word(2), lastOff = m.word(3);
  auto sizeOf = [&](uint32_t l) { return (m.word(4 + (l >> 2)) >> ((l & 3u) * 8)) & 0xFFu; };
  auto lutAt = [&](uint32_t lt, uint32_t a, uint32_t b, uint32_t c) -> uint32_t {
    if (lt == 0) return c;
    const uint32_t pos = (lt + (((a & 0xFFFFu) << 8) | (b << 4) | c)) & 0xFFFFu;
    return m.lut[pos < kLutSize ? pos : kLutSize];
  };
  auto rd = [&](uint32_t row, uint32_t ii) -> uint32_t { return A.at(uint64_t(row) * I + ii, "rows read"); };
  struct Wr { uint32_t row, ii, v; };
  std::vector<Wr> wr;
  auto level = [&](uint32_t n, auto &&op) {
    for (uint32_t r = 0; r < n; r += 2 * W) {
      wr.clear();
      for (uint32_t half = 0; half < 2; ++half)
        for (uint32_t sub = 0; sub < W; ++sub)
          for (uint32_t ii = 0; ii < I; ++ii) op(minOp(r + half * W + sub, n), ii);
      for (auto &x : wr) A.at(uint64_t(x.row) * I + x.ii, "rows write") = uint8_t(x.v);
    }
  };
  level(nt, [&](uint32_t q, uint32_t ii) {
    const uint32_t x = m.word(topOff + 2 * q), y = m.word(topOff + 2 * q + 1);
    const uint32_t v = lutAt(x & 0xFFFFu, rd(x >> 16, ii), rd(y & 0x1FFFu, ii), rd((y >> 13) & 0x1FFFu, ii));
    wr.push_back({V + (y >> 26), ii, v});
  });
  uint64_t o = midOff;
  for (uint32_t l = 0; l < nMid; ++l) {
    const uint32_t n = sizeOf(l);
    level(n, [&](uint32_t q, uint32_t ii) {
      const uint32_t x = m.word(o + 2 * q), y = m.word(o + 2 * q + 1);
      const uint32_t v =
          lutAt(x & 0x3FFFu, rd(V + ((x >> 14) & 0xFFu), ii), rd(V + (x >> 22), ii), rd(V + (y & 0xFFu), ii));
      wr.push_back({y >> 8, ii, v});
    });
    o += 2ull * n;
  }
  for (uint32_t r = 0; r < nl; r += W)
    for (uint32_t sub = 0; sub < W; ++sub) {
      const uint32_t w = m.word(lastOff + minOp(r + sub, nl));
      for (uint32_t ii = 0; ii < I; ++ii) A.at(uint64_t(w & 0xFFFFu) * I + ii, "last write") = uint8_t(rd(V + ((w >> 16) & 0xFFu), ii));
    }
}

// T_LUTFAST, as lutFastTask: byte offsets from the lane's column; a lane
// holds one 32-bit word of the rows (four bytes or eight nibbles), so a
// round is 256 / RB ops and every offset must be a multiple of 4.
void runFast(Model &m, Area &A) {
  const uint32_t I = A.I, RB = A.rb(), W = 256 / RB, V = A.V;
  const uint32_t w0 = m.word(0), w1 = m.word(1);
  const uint32_t nMid = (w0 >> 8) & 0xFFFFu, nt = w1 & 0xFFFFu, nl = w1 >> 16;
  const uint32_t topOff = 4 + ((((nMid + 3) >> 2) + 1) & ~1u);
  const uint32_t midOff = m.word(2), lastOff = m.word(3);
  auto sizeOf = [&](uint32_t l) { return (m.word(4 + (l >> 2)) >> ((l & 3u) * 8)) & 0xFFu; };
  auto aligned = [&](uint64_t off) {
    if (off % 4) die("T_LUTFAST offset not a multiple of 4: " + std::to_string(off));
    if (off % RB) die("T_LUTFAST offset not a whole row: " + std::to_string(off));
  };
  auto nib = [&](uint64_t off, uint32_t ii) -> uint32_t {
    const uint32_t v = A.get(off, ii, "operand");
    if (v > 15) die("operand above 15 at offset " + std::to_string(off));
    return v;
  };
  auto eval = [&](uint32_t x, uint32_t y, uint32_t ii) -> uint32_t {
    aligned(x >> 16);
    aligned(y & 0xFFFFu);
    aligned(y >> 16);
    const uint32_t pos = (((nib(x >> 16, ii) << 4) | nib(y & 0xFFFFu, ii)) << 4 | nib(y >> 16, ii)) + (x & 0xFFFFu);
    if (pos >= m.lut.size()) die("LUT position past the region: " + std::to_string(pos));
    return m.lut[pos];
  };
  struct Wr { uint64_t off; uint32_t ii, v; };
  std::vector<Wr> wr;
  uint64_t cur = topOff;
  uint32_t n = nt;
  for (uint32_t l = 0; l <= nMid; ++l) {
    const uint32_t nn = l < nMid ? sizeOf(l) : 0;
    const uint64_t next = l == 0 ? midOff : cur + 2ull * n;
    const bool lastMid = l > 0 && l == nMid;
    const uint32_t wb = (l == 0 || ((l - 1) & 1u)) ? 0u : kFastBufY;
    if (n > 64) die("task level of more than 64 ops");
    for (uint32_t r = 0; r < n; r += 2 * W) {
      wr.clear();
      for (uint32_t half = 0; half < (r + W < n ? 2u : 1u); ++half)
        for (uint32_t sub = 0; sub < W; ++sub) {
          const uint32_t lr = r + half * W + sub;  // the lane's row; past the level's end it takes the last op
          const uint32_t q = std::min(lr, n - 1);
          const uint32_t x = m.word(cur + 2 * q), y = m.word(cur + 2 * q + 1);
          uint64_t dst;
          if (lastMid) {
            const uint32_t d = m.word(cur + 2ull * n + q / 2);
            dst = (q & 1u) ? d >> 16 : d & 0xFFFFu;
          } else {
            if (lr >= kFastBufY || wb + lr >= kConstRow) die("level row past its level buffer");
            dst = uint64_t(V + wb + lr) * RB;
          }
          aligned(dst);
          for (uint32_t ii = 0; ii < I; ++ii) wr.push_back({dst, ii, eval(x, y, ii)});
        }
      for (auto &w : wr) A.set(w.off, w.ii, w.v, "write");
    }
    cur = next;
    n = nn;
  }
  for (uint32_t r = 0; r < nl; r += W)
    for (uint32_t sub = 0; sub < W; ++sub) {
      const uint32_t w = m.word(lastOff + std::min(r + sub, nl - 1));
      aligned(w & 0xFFFFu);
      aligned(w >> 16);
      for (uint32_t ii = 0; ii < I; ++ii) A.set(w & 0xFFFFu, ii, nib(w >> 16, ii), "last write");
    }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 4) die("usage: batch_check design.bin I seed [rounds] [nib]");
  auto bin = slurp(argv[1]);
  const uint32_t I = std::atoi(argv[2]);
  const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
  const uint32_t rounds = argc > 4 ? std::atoi(argv[4]) : 2;
  const bool nibRows = argc > 5 && std::string(argv[5]) == "nib";  // the fast image for nibble rows
  if (I != 8 && I != 16 && I != 32 && I != 64) die("I must be 8, 16, 32 or 64");
  SimDesignInfo d;
  std::string err;
  if (!parseSimDesignInfo(bin.data(), bin.size(), d, &err)) die("parse: " + err);
  d.Init();
  if (validateProgram(d, &err) != TGS_OK) die("invalid: " + err);
  if (!nibbleProgram(d)) die("not a nibble program");
  uint32_t maxVp = 16;
  for (const auto &p : d.parts) maxVp = std::max(maxVp, p.valuePoolSize);
  const uint32_t V = (maxVp + 15) & ~15u;
  const uint32_t RB = nibRows ? I / 2 : I;  // bytes per row of the fast image's wave area
  if ((V + kLevelRows) * RB > 65536) die("rows do not fit 16-bit offsets for this I");
  StimMap stim;
  const Wave64Image gi = buildWave64Image(d, stim, 4, V, 0);
  const Wave64Image fi = buildWave64Image(d, stim, 4, V, RB);
  if (gi.fastTasks != 0) die("general image holds T_LUTFAST tasks");
  if (gi.partOff.size() != fi.partOff.size()) die("partition counts differ");
  if (nibRows && fi.rowTasks != 0) die("nibble rows need every LUT task fast");
  std::mt19937_64 rng(seed);
  std::vector<uint8_t> lut(kLutRegionB, 0);
  for (uint32_t k = 0; k < kLutSize; ++k) lut[k] = k < d.lut.size() ? d.lut[k] : 0;
  for (uint32_t k = 0; k < 32; ++k) lut[kLutIdBase + k] = uint8_t(k);
  uint64_t nFast = 0, nRows = 0;
  for (uint32_t rep = 0; rep < rounds; ++rep) {
    if (rep > 0)  // other tables than the design's: every nibble entry
      for (uint32_t k = 0; k < kLutSize; ++k) lut[k] = uint8_t(rng() & 15u);
    for (size_t p = 0; p < gi.partOff.size(); ++p) {
      const uint32_t *GP = gi.words.data() + gi.partOff[p], *FP = fi.words.data() + fi.partOff[p];
      // the same random nibbles in the value-pool rows of both areas, other
      // random nibbles in their level rows (a row read before the task wrote
      // it shows up as a difference), constant rows 0..31 (with nibble rows
      // only 0..15 are representable: rows 16..31 stay random, so a read of
      // them shows up too)
      Area ga{V, I, false, std::vector<uint8_t>(size_t(V + kLevelRows) * I)};
      Area fa{V, I, nibRows, std::vector<uint8_t>(size_t(V + kLevelRows) * RB)};
      for (auto &x : ga.b) x = uint8_t(rng() & 15u);
      for (auto &x : fa.b) x = uint8_t(rng() & 0xFFu);
      for (uint32_t r = 0; r < V; ++r)
        for (uint32_t ii = 0; ii < I; ++ii) fa.set(uint64_t(r) * RB, ii, ga.get(uint64_t(r) * I, ii, "init"), "init");
      for (uint32_t r = V; r < V + kLevelRows; ++r)
        for (uint32_t ii = 0; ii < I; ++ii) fa.set(uint64_t(r) * RB, ii, uint32_t(rng() & 15u), "init");
      for (uint32_t r = kConstRow; r < kLevelRows; ++r)
        for (uint32_t ii = 0; ii < I; ++ii) {
          ga.set(uint64_t(V + r) * I, ii, r - kConstRow, "const");
          if (!nibRows || r - kConstRow < 16) fa.set(uint64_t(V + r) * RB, ii, r - kConstRow, "const");
        }
      const uint32_t nLev = GP[H_NLEV];
      if (FP[H_NLEV] != nLev) die("exec level counts differ");
      // the LUT tasks of one exec level are independent: run all of them in
      // each image, then compare (T_LUTFAST packs several pairs' parts)
      for (uint32_t e = 0; e < nLev; ++e) {
        const uint32_t *gl = GP + GP[H_OFF_LEV], *fl = FP + FP[H_OFF_LEV];
        uint32_t gOther = 0, fOther = 0;
        for (uint32_t k = gl[2 * e]; k < gl[2 * e + 1]; ++k) {
          const uint32_t gt = (GP + GP[H_OFF_TASK])[k];
          if ((GP[gt] & 0xFFu) != T_LUTROWS) {
            ++gOther;
            continue;
          }
          Model gm{gi.words, gi.partOff[p] + gt, gi.partWords[p] - gt, lut};
          runRows(gm, ga);
        }
        for (uint32_t k = fl[2 * e]; k < fl[2 * e + 1]; ++k) {
          const uint32_t ft = (FP + FP[H_OFF_TASK])[k];
          const uint32_t fk = FP[ft] & 0xFFu;
          Model fm{fi.words, fi.partOff[p] + ft, fi.partWords[p] - ft, lut};
          if (fk == T_LUTFAST) {
            runFast(fm, fa);
            ++nFast;
          } else if (fk == T_LUTROWS) {
            runRows(fm, fa);
            ++nRows;
          } else {
            ++fOther;
          }
        }
        if (gOther != fOther) die("other task counts differ");
        for (uint32_t r = 0; r < V; ++r)
          for (uint32_t ii = 0; ii < I; ++ii)
            if (ga.get(uint64_t(r) * I, ii, "cmp") != fa.get(uint64_t(r) * RB, ii, "cmp"))
              die("value pools differ after partition " + std::to_string(p) + " exec level " + std::to_string(e) +
                  " (slot " + std::to_string(r) + ", instance " + std::to_string(ii) + ")");
      }
    }
  }
  std::printf("tasks %llu %llu\n", (unsigned long long)nFast, (unsigned long long)nRows);
  return 0;
}
