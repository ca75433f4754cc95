// toucanGPUSim::ToucanSimulator over the C-ABI runtime.
//
// Method-for-method counterpart of src/ToucanSim.cu:102-325 and the VCD
// writer of unsupport/SimVCD.cpp:22-248 (the reference builds only the no-op
// src/SimVCD_Dummy.cpp). Host copies of design and symbols are kept in the
// public members as in the reference; the device state lives in the context.
#include <algorithm>
#include <functional>
#include <cassert>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <iterator>
#include <map>
#include <sstream>
#include <unordered_map>

#include "ToucanGPUSim.h"
#include "toucan_rt.h"

namespace toucanGPUSim {

static bool slurp(const std::string &path, std::vector<uint8_t> &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) return false;
  out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return true;
}

std::string ToucanSimulator::lastError() const {
  char buf[1024] = {0};
  if (ctx) tgs_last_error(ctx, buf, sizeof buf);
  return buf;
}

// src/ToucanSim.cu:102-204.
int ToucanSimulator::init(const int gpu_id, const std::string designBinFilename, const std::string symbolBinFilename,
                          bool shouldRandomize, bool enablePrint) {
  closeVcd();
  cycle_cnt = 0;
  printsSeen = 0;
  probeParts.clear();
  probeSlots.clear();
  probeFirst.clear();
  probeEpoch = UINT64_MAX;
  std::vector<uint8_t> d, s;
  if (!slurp(designBinFilename, d)) {
    std::cerr << "Error opening file: " << designBinFilename << std::endl;
    return -1;
  }
  if (!slurp(symbolBinFilename, s)) {
    std::cerr << "Error opening file: " << symbolBinFilename << std::endl;
    return -1;
  }
  std::string err;
  design = SimDesignInfo();
  symbols = SimDebugInfo();
  if (!parseSimDesignInfo(d.data(), d.size(), design, &err) || !parseSimDebugInfo(s.data(), s.size(), symbols, &err)) {
    std::cerr << "Malformed program: " << err << std::endl;
    return -1;
  }
  design.Init();
  if (shouldRandomize) design.Randomize(0, symbols);
  if (!ctx) {
    int rc = tgs_create(gpu_id, &ctx);
    if (rc != TGS_OK) {
      std::cerr << "No usable GPU (" << rc << ")" << std::endl;
      return -1;
    }
  }
  int rc = tgs_load(ctx, d.data(), d.size(), s.data(), s.size(), shouldRandomize, enablePrint);
  if (rc != TGS_OK) {
    std::cerr << "init failed: " << lastError() << std::endl;
    return -1;
  }
  return 0;
}

int ToucanSimulator::setKernel(int kernel) { return ctx ? tgs_set_kernel(ctx, kernel) : TGS_E_STATE; }

int ToucanSimulator::setStimulus(const std::vector<uint32_t> &inputIds, uint64_t seed) {
  return ctx ? tgs_set_stimulus(ctx, inputIds.data(), uint32_t(inputIds.size()), seed) : TGS_E_STATE;
}

// Print ops fired since the last call, in cycle order (the reference prints
// them from the device, src/SimEval.cu:696-704).
void ToucanSimulator::flushPrints() {
  uint64_t n = 0;
  if (tgs_get_print_log(ctx, 0, nullptr, nullptr, 0, &n) != TGS_OK || n <= printsSeen) return;
  std::vector<uint64_t> cyc(n);
  std::vector<uint32_t> msg(n);
  tgs_get_print_log(ctx, 0, cyc.data(), msg.data(), n, &n);
  std::vector<std::pair<uint64_t, uint32_t>> ev;
  for (uint64_t i = printsSeen; i < std::min<uint64_t>(n, cyc.size()); ++i) ev.emplace_back(cyc[i], msg[i]);
  std::stable_sort(ev.begin(), ev.end());
  for (auto &e : ev)
    if (e.second < design.printMsgs.size()) std::cout << design.printMsgs[e.second];
  printsSeen = n;
}

// src/ToucanSim.cu:206-230.
bool ToucanSimulator::eval() {
  int stopped = 0;
  int rc = tgs_eval(ctx, &stopped);
  if (rc != TGS_OK) {
    std::cerr << "eval failed: " << lastError() << std::endl;
    std::abort();
  }
  flushPrints();
  if (dumpVCD) dumpVcdWorker(cycle_cnt);
  cycle_cnt++;
  return stopped != 0;
}

// src/ToucanSim.cu:232-267.
bool ToucanSimulator::eval_free_running(uint32_t max_cycles) {
  int stopped = 0;
  int rc = tgs_eval_free_running(ctx, max_cycles, &stopped);
  if (rc != TGS_OK) {
    std::cerr << "eval_free_running failed: " << lastError() << std::endl;
    std::abort();
  }
  flushPrints();
  cycle_cnt += max_cycles;
  if (dumpVCD) dumpVcdWorker(cycle_cnt);
  return stopped != 0;
}

// src/ToucanSim.cu:271-278 + write_reg_to_gpu (src/SimEval.cu:1072-1081).
void ToucanSimulator::setRegister(const std::string &signalName, uint64_t signalValue) {
  assert(symbols.regDebugInfo.count(signalName));
  auto &locs = symbols.regDebugInfo[signalName];
  assert(locs.size() <= 16 && "Register wider than 64 bits are not supported by this function");
  std::vector<uint32_t> ids;
  std::vector<uint8_t> vals(locs.size());
  for (auto &l : locs) ids.push_back(std::get<0>(l));
  for (size_t k = locs.size(); k-- > 0;) {
    vals[k] = uint8_t(signalValue & 0xF);
    signalValue >>= 4;
  }
  assert(signalValue == 0 && "Given value is wider than register");
  if (tgs_write_regs(ctx, 0, ids.data(), uint32_t(ids.size()), vals.data()) != TGS_OK) {
    std::cerr << "setRegister failed: " << lastError() << std::endl;
    std::abort();
  }
}

// src/ToucanSim.cu:280-289 + read_reg_from_gpu (src/SimEval.cu:1058-1069).
uint64_t ToucanSimulator::readRegister(const std::string &signalName) {
  assert(symbols.regDebugInfo.count(signalName));
  auto &locs = symbols.regDebugInfo[signalName];
  assert(locs.size() <= 16 && "Register wider than 64 bits are not supported by this function");
  std::vector<uint32_t> ids;
  for (auto &l : locs) ids.push_back(std::get<0>(l));
  std::vector<uint8_t> vals(ids.size());
  if (tgs_read_regs(ctx, 0, ids.data(), uint32_t(ids.size()), vals.data()) != TGS_OK) {
    std::cerr << "readRegister failed: " << lastError() << std::endl;
    std::abort();
  }
  uint64_t r = 0;
  for (auto v : vals) r = (r << 4) | v;
  return r;
}

bool ToucanSimulator::step(const std::vector<std::pair<std::string, uint64_t>> &writes, uint32_t cycles,
                           const std::vector<std::string> &reads, std::vector<uint64_t> &out) {
  out.assign(reads.size(), 0);
  if (dumpVCD) {  // every cycle goes through eval() so the dump sees it
    for (auto &w : writes) setRegister(w.first, w.second);
    int stopped = 0;
    for (uint32_t k = 0; k < cycles; ++k) stopped = eval() ? 1 : 0;
    if (cycles == 0 && tgs_get_stop(ctx, 0, &stopped) != TGS_OK) stopped = 0;
    for (size_t k = 0; k < reads.size(); ++k) out[k] = readRegister(reads[k]);
    return stopped != 0;
  }
  std::vector<uint32_t> wids, rids, rlen;
  std::vector<uint8_t> wvals;
  for (auto &w : writes) {
    assert(symbols.regDebugInfo.count(w.first));
    auto &locs = symbols.regDebugInfo[w.first];
    assert(locs.size() <= 16 && "Register wider than 64 bits are not supported by this function");
    uint64_t v = w.second;
    const size_t base = wvals.size();
    wvals.resize(base + locs.size());
    for (auto &l : locs) wids.push_back(std::get<0>(l));
    for (size_t k = locs.size(); k-- > 0;) {
      wvals[base + k] = uint8_t(v & 0xF);
      v >>= 4;
    }
    assert(v == 0 && "Given value is wider than register");
  }
  for (auto &r : reads) {
    assert(symbols.regDebugInfo.count(r));
    auto &locs = symbols.regDebugInfo[r];
    assert(locs.size() <= 16 && "Register wider than 64 bits are not supported by this function");
    for (auto &l : locs) rids.push_back(std::get<0>(l));
    rlen.push_back(uint32_t(locs.size()));
  }
  std::vector<uint8_t> rvals(rids.size());
  int stopped = 0;
  if (tgs_step(ctx, wids.data(), uint32_t(wids.size()), wvals.data(), cycles, rids.data(), uint32_t(rids.size()),
               rvals.data(), &stopped) != TGS_OK) {
    std::cerr << "step failed: " << lastError() << std::endl;
    std::abort();
  }
  if (cycles) flushPrints();
  cycle_cnt += cycles;
  for (size_t k = 0, at = 0; k < reads.size(); at += rlen[k], ++k)
    for (uint32_t j = 0; j < rlen[k]; ++j) out[k] = (out[k] << 4) | rvals[at + j];
  return stopped != 0;
}

// Probe list management: every fragment of a probed signal is one
// (partition, value slot) probe; probeFirst maps a signal to its first probe.
uint32_t ToucanSimulator::addSignalProbes(const std::vector<std::string> &names) {
  uint32_t added = 0;
  for (auto &name : names) {
    if (probeFirst.count(name)) continue;
    auto it = symbols.signalDebugInfo.find(name);
    if (it == symbols.signalDebugInfo.end()) continue;
    probeFirst[name] = uint32_t(probeParts.size());
    for (auto &f : it->second) {
      probeParts.push_back(std::get<0>(f));
      probeSlots.push_back(std::get<1>(f));
    }
    ++added;
  }
  if (added) {
    if (tgs_set_probes(ctx, probeParts.data(), probeSlots.data(), uint32_t(probeParts.size())) != TGS_OK) {
      std::cerr << "signal probes failed: " << lastError() << std::endl;
      std::abort();
    }
    probeEpoch = cycle_cnt;  // the new list starts zeroed: values from the next cycle on
  }
  return added;
}

// src/ToucanSim.cu:291-300 is a stub that returns UINT64_MAX. Here a signal
// (value-pool slots named by SimDebugInfo::signalDebugInfo) is read from the
// signal probes: the value its partition computed in the last cycle run.
// Fragments are concatenated most significant first, `width` bits each, as
// the VCD writer prints them (unsupport/SimVCD.cpp:222-233). A signal is
// probed from the first call that names it (or from enableVCDDump) on; until
// a cycle has run after that, the call returns UINT64_MAX like the reference.
uint64_t ToucanSimulator::readSignal(const std::string &signalName) {
  assert(symbols.signalDebugInfo.count(signalName));
  auto &locs = symbols.signalDebugInfo[signalName];
  uint32_t width = 0;
  for (auto &f : locs) width += std::get<2>(f);
  assert(width <= 64 && "Signal wider than 64 bits are not supported by this function");
  if (!probeFirst.count(signalName)) addSignalProbes({signalName});
  if (probeEpoch == cycle_cnt) return UINT64_MAX;
  std::vector<uint8_t> vals(probeParts.size());
  if (tgs_read_probes(ctx, 0, vals.data(), uint32_t(vals.size())) != TGS_OK) {
    std::cerr << "readSignal failed: " << lastError() << std::endl;
    std::abort();
  }
  const uint32_t first = probeFirst[signalName];
  uint64_t r = 0;
  for (size_t k = 0; k < locs.size(); ++k) {
    const uint32_t w = std::get<2>(locs[k]);
    r = (w >= 64 ? 0 : r << w) | (vals[first + k] & ((w >= 32 ? ~0u : (1u << w) - 1u)));
  }
  return r;
}

// unsupport/SimVCD.cpp:88-97: printable identifiers from '!' in base 94.
static std::string vcdIdentifier(uint32_t id) {
  std::string s;
  do {
    s.insert(s.begin(), char('!' + id % 94));
    id /= 94;
  } while (id != 0);
  return s;
}

// unsupport/SimVCD.cpp:99-176: the signals of the symbol table, ordered by
// (partition, value slot) of their first fragment, identifiers in that
// order, widths the sum of the fragment widths; dotted names become nested
// scopes (names without a dot go in scope "top"). The registers follow as
// further variables, in name order: their values are architectural state the
// reference's writer has no access to.
void ToucanSimulator::enableVCDDump(const std::string vcdFilename) {
  closeVcd();
  dumpVCD = true;
  vcdOfs = new std::ofstream(vcdFilename, std::ios::out);
  std::vector<std::pair<uint64_t, std::string>> sigs;
  for (auto &kv : symbols.signalDebugInfo)
    if (!kv.second.empty())
      sigs.push_back({(uint64_t(std::get<0>(kv.second[0])) << 32) + std::get<1>(kv.second[0]), kv.first});
  std::sort(sigs.begin(), sigs.end());
  std::vector<std::string> sigNames;
  for (auto &x : sigs) sigNames.push_back(x.second);
  addSignalProbes(sigNames);
  std::vector<std::string> regNames;
  for (auto &kv : symbols.regDebugInfo) regNames.push_back(kv.first);
  std::sort(regNames.begin(), regNames.end());
  vcdNames.clear();
  vcdSrc.clear();
  vcdWidth.clear();
  vcdIds.clear();
  const uint32_t regBytes = uint32_t(design.regPoolSize);
  for (auto &n : sigNames) {
    std::vector<std::tuple<uint32_t, uint32_t>> src;
    uint32_t w = 0, k = probeFirst[n];
    for (auto &f : symbols.signalDebugInfo[n]) {
      src.push_back({regBytes + k++, std::get<2>(f)});
      w += std::get<2>(f);
    }
    vcdNames.push_back(n);
    vcdSrc.push_back(src);
    vcdWidth.push_back(w);
  }
  for (auto &n : regNames) {
    std::vector<std::tuple<uint32_t, uint32_t>> src;
    uint32_t w = 0;
    for (auto &f : symbols.regDebugInfo[n]) {
      src.push_back({std::get<0>(f), std::get<1>(f)});
      w += std::get<1>(f);
    }
    vcdNames.push_back(n);
    vcdSrc.push_back(src);
    vcdWidth.push_back(w);
  }
  for (uint32_t i = 0; i < vcdNames.size(); ++i) vcdIds.push_back(vcdIdentifier(i));
  vcdLast.assign(vcdNames.size(), {});
  *vcdOfs << "$date\n    " << __DATE__ << " " << __TIME__ << "\n$end\n"
          << "$version\n    VCD generator\n$end\n"
          << "$timescale\n    1ps\n$end\n";
  // scope tree
  struct Node {
    std::map<std::string, Node> kids;
    std::vector<uint32_t> sigs;
  } root;
  for (uint32_t i = 0; i < vcdNames.size(); ++i) {
    Node *n = &root;
    std::string rest = vcdNames[i];
    size_t dot;
    if (rest.find('.') == std::string::npos) n = &root.kids["top"];
    while ((dot = rest.find('.')) != std::string::npos) {
      n = &n->kids[rest.substr(0, dot)];
      rest = rest.substr(dot + 1);
    }
    n->sigs.push_back(i);
  }
  std::function<void(const std::string &, const Node &)> walk = [&](const std::string &name, const Node &n) {
    *vcdOfs << "$scope module " << name << " $end\n";
    for (auto i : n.sigs) {
      std::string leaf = vcdNames[i].substr(vcdNames[i].rfind('.') == std::string::npos ? 0 : vcdNames[i].rfind('.') + 1);
      *vcdOfs << "$var wire " << vcdWidth[i] << " " << vcdIds[i] << " " << leaf << " $end\n";
    }
    for (auto &k : n.kids) walk(k.first, k.second);
    *vcdOfs << "$upscope $end\n";
  };
  for (auto &k : root.kids) walk(k.first, k.second);
  *vcdOfs << "$enddefinitions $end\n$dumpvars\n";
}

// unsupport/SimVCD.cpp:193-241: change-only values per dumped cycle. The
// dump is pipelined: the register pool and the signal probes of this cycle
// are snapshotted asynchronously (pinned copy on a side stream) and the
// previous snapshot is written, so the copy and the formatting overlap the
// next eval.
void ToucanSimulator::dumpVcdWorker(uint64_t cycle) {
  if (!vcdOfs) return;
  uint64_t t = 0;
  if (tgs_snapshot_begin(ctx, 0, &t) != TGS_OK) {
    std::cerr << "VCD snapshot failed: " << lastError() << std::endl;
    return;
  }
  if (vcdPending) writeVcdSnapshot(vcdTicket, vcdStamp);
  vcdPending = true;
  vcdTicket = t;
  vcdStamp = cycle;
}

void ToucanSimulator::writeVcdSnapshot(uint64_t ticket, uint64_t stamp) {
  vcdPending = false;
  const uint8_t *data = nullptr;
  if (tgs_snapshot_wait(ctx, ticket, &data, nullptr) != TGS_OK) {
    std::cerr << "VCD snapshot failed: " << lastError() << std::endl;
    return;
  }
  *vcdOfs << "#" << stamp << "\n";
  for (uint32_t i = 0; i < vcdSrc.size(); ++i) {
    std::vector<uint8_t> v;
    for (auto &f : vcdSrc[i]) v.push_back(data[std::get<0>(f)]);
    if (!vcdLast[i].empty() && vcdLast[i] == v) continue;
    vcdLast[i] = v;
    if (vcdWidth[i] == 1) {
      *vcdOfs << int(v[0] & 1) << vcdIds[i] << "\n";
    } else {
      *vcdOfs << "b";
      for (size_t k = 0; k < v.size(); ++k) {
        uint32_t w = std::get<1>(vcdSrc[i][k]);
        for (int b = int(w) - 1; b >= 0; --b) *vcdOfs << char('0' + ((v[k] >> b) & 1));
      }
      *vcdOfs << " " << vcdIds[i] << "\n";
    }
  }
}

// Write the snapshot still in flight and close the VCD file, if any.
void ToucanSimulator::closeVcd() {
  if (vcdOfs) {
    if (vcdPending && ctx) writeVcdSnapshot(vcdTicket, vcdStamp);
    vcdOfs->close();
    delete vcdOfs;
    vcdOfs = nullptr;
  }
  vcdPending = false;
  dumpVCD = false;
}

void ToucanSimulator::cleanup() {
  closeVcd();
  if (ctx) {
    tgs_destroy(ctx);
    ctx = nullptr;
  }
}

// src/ToucanSim.cu:306-325: the launch profile (shape, kernel time, barrier share).
void ToucanSimulator::printProfile() {
  tgs_profile p{};
  if (!ctx || tgs_get_profile(ctx, &p) != TGS_OK) return;
  double stall = p.total_ticks ? 100.0 * double(p.barrier_ticks) / double(p.total_ticks) : 0.0;
  std::cout << "kernel " << p.kernel << " grid " << p.grid << " block " << p.block << " lds " << p.lds_bytes
            << " resident " << p.resident << " launches " << p.launches << " cycles " << p.cycles << " kernel_ms "
            << p.kernel_ms << " barrier_stall " << stall << "%\n";
}

}  // namespace toucanGPUSim
