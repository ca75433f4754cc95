// toucanGPUSim::ToucanSimulator over the C-ABI runtime.
//
// Method-for-method counterpart of src/ToucanSim.cu:102-325 and the VCD
// writer of unsupport/SimVCD.cpp:22-248 (the reference builds only the no-op
// src/SimVCD_Dummy.cpp). Host copies of design and symbols are kept in the
// public members as in the reference; the device state lives in the context.
#include <algorithm>
#include <functional>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <iterator>
#include <map>
#include <sstream>
#include <unordered_map>

#include "SimEval.h"
#include "ToucanGPUSim.h"
#include "toucan_rt.h"

namespace toucanGPUSim {

static bool slurp(const std::string &path, std::vector<uint8_t> &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) return false;
  out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return true;
}

std::string ToucanSimulator::lastError() const { return errText; }

bool ToucanSimulator::fail(int code, const std::string &what) {
  status = code;
  errText = what;
  std::cerr << what << std::endl;
  return false;
}

// The runtime's own text of its last error on this simulator's context.
static std::string runtimeError(tgs_ctx *ctx) {
  char buf[1024] = {0};
  if (ctx) tgs_last_error(ctx, buf, sizeof buf);
  return buf;
}

bool ToucanSimulator::ready(const char *call) {
  if (ctx && loadedOk) return true;
  return fail(TGS_E_STATE, std::string(call) + ": no design loaded (init has not succeeded)");
}

// The process's current simulator context, the target of the include/SimEval.h
// shim functions (the reference's process-global device state). The context
// copy_netlist_to_gpu creates lives until the process ends, as the
// reference's device globals do (src/SimEval.cu:190-215).
namespace {
tgs_ctx *g_current = nullptr;
tgs_ctx *g_shimOwned = nullptr;
}  // namespace

// src/ToucanSim.cu:102-204.
int ToucanSimulator::initWithProfile(const int gpu_id, const std::string &designBinFilename,
                                     const std::string &symbolBinFilename, bool shouldRandomize, bool enablePrint,
                                     bool profileParts) {
  closeVcd();
  status = 0;
  errText.clear();
  loadedOk = false;
  cycle_cnt = 0;
  printsSeen = 0;
  probeParts.clear();
  probeSlots.clear();
  probeFirst.clear();
  probeEpoch = UINT64_MAX;
  std::vector<uint8_t> d, s;
  if (!slurp(designBinFilename, d)) {
    fail(TGS_E_IO, "Error opening file: " + designBinFilename);
    return -1;
  }
  if (!slurp(symbolBinFilename, s)) {
    fail(TGS_E_IO, "Error opening file: " + symbolBinFilename);
    return -1;
  }
  std::string err;
  design = SimDesignInfo();
  symbols = SimDebugInfo();
  if (!parseSimDesignInfo(d.data(), d.size(), design, &err) || !parseSimDebugInfo(s.data(), s.size(), symbols, &err)) {
    fail(TGS_E_FORMAT, "Malformed program: " + err);
    return -1;
  }
  design.Init();
  if (shouldRandomize) design.Randomize(0, symbols);
  if (!ctx) {
    int rc = tgs_create(gpu_id, &ctx);
    if (rc != TGS_OK) {
      ctx = nullptr;
      fail(rc, "No usable GPU (" + std::to_string(rc) + ")");
      return -1;
    }
  }
  int rc = tgs_load(ctx, d.data(), d.size(), s.data(), s.size(), shouldRandomize, enablePrint);
  if (rc != TGS_OK) {
    fail(rc, "init failed: " + runtimeError(ctx));
    return -1;
  }
  if (profileParts && (rc = tgs_set_part_profile(ctx, 1, PROFILE_START_CYCLE, PROFILE_COLLECT_CYCLE)) != TGS_OK) {
    fail(rc, "init failed: " + runtimeError(ctx));
    return -1;
  }
  loadedOk = true;
  g_current = ctx;
  return 0;
}

int ToucanSimulator::setKernel(int kernel) { return ctx ? tgs_set_kernel(ctx, kernel) : TGS_E_STATE; }

int ToucanSimulator::setStimulus(const std::vector<uint32_t> &inputIds, uint64_t seed) {
  return ctx ? tgs_set_stimulus(ctx, inputIds.data(), uint32_t(inputIds.size()), seed) : TGS_E_STATE;
}

// Print ops fired since the last call, in cycle order (the reference prints
// them from the device, src/SimEval.cu:696-704). The runtime's log is
// consumed, so a long run never fills it.
void ToucanSimulator::flushPrints() {
  uint64_t n = 0, dropped = 0;
  if (tgs_get_print_log(ctx, 0, nullptr, nullptr, 0, &n) != TGS_OK) return;
  if (n == 0) return;
  std::vector<uint64_t> cyc(n);
  std::vector<uint32_t> msg(n);
  if (tgs_get_print_log(ctx, 0, cyc.data(), msg.data(), n, &n) != TGS_OK) return;
  tgs_clear_print_log(ctx, 0, &dropped);
  std::vector<std::pair<uint64_t, uint32_t>> ev;
  for (uint64_t i = 0; i < std::min<uint64_t>(n, cyc.size()); ++i) ev.emplace_back(cyc[i], msg[i]);
  std::stable_sort(ev.begin(), ev.end());
  for (auto &e : ev)
    if (e.second < design.printMsgs.size()) std::cout << design.printMsgs[e.second];
  if (dropped) std::cerr << dropped << " print events were dropped (more than one launch's log capacity)\n";
  printsSeen += n;
}

// src/ToucanSim.cu:206-230.
bool ToucanSimulator::eval() {
  if (!ready("eval")) return false;
  int stopped = 0;
  int rc = tgs_eval(ctx, &stopped);
  if (rc != TGS_OK) return fail(rc, "eval failed: " + runtimeError(ctx));
  status = 0;
  flushPrints();
  if (dumpVCD) dumpVcdWorker(cycle_cnt);
  cycle_cnt++;
  return stopped != 0;
}

// src/ToucanSim.cu:232-267.
bool ToucanSimulator::eval_free_running(uint32_t max_cycles) {
  if (!ready("eval_free_running")) return false;
  int stopped = 0;
  int rc = tgs_eval_free_running(ctx, max_cycles, &stopped);
  if (rc != TGS_OK) return fail(rc, "eval_free_running failed: " + runtimeError(ctx));
  status = 0;
  flushPrints();
  cycle_cnt += max_cycles;
  if (dumpVCD) dumpVcdWorker(cycle_cnt);
  return stopped != 0;
}

// Register fragment ids of a name, or an error (the reference asserts,
// src/ToucanSim.cu:271-289).
static bool regIds(const SimDebugInfo &symbols, const std::string &name, std::vector<uint32_t> &ids, std::string &err) {
  auto it = symbols.regDebugInfo.find(name);
  if (it == symbols.regDebugInfo.end()) {
    err = "unknown register " + name;
    return false;
  }
  if (it->second.size() > 16) {
    err = "Register wider than 64 bits are not supported by this function: " + name;
    return false;
  }
  ids.clear();
  for (auto &l : it->second) ids.push_back(std::get<0>(l));
  return true;
}

// src/ToucanSim.cu:271-278 + write_reg_to_gpu (src/SimEval.cu:1072-1081):
// the last fragment takes the low nibble. As there, the fragments are
// written before a value wider than the register is reported.
void ToucanSimulator::setRegister(const std::string &signalName, uint64_t signalValue) {
  if (!ready("setRegister")) return;
  std::vector<uint32_t> ids;
  std::string err;
  if (!regIds(symbols, signalName, ids, err)) {
    fail(TGS_E_ARG, "setRegister: " + err);
    return;
  }
  std::vector<uint8_t> vals(ids.size());
  for (size_t k = ids.size(); k-- > 0;) {
    vals[k] = uint8_t(signalValue & 0xF);
    signalValue >>= 4;
  }
  int rc = tgs_write_regs(ctx, 0, ids.data(), uint32_t(ids.size()), vals.data());
  if (rc != TGS_OK) {
    fail(rc, "setRegister failed: " + runtimeError(ctx));
    return;
  }
  if (signalValue != 0) {
    fail(TGS_E_ARG, "setRegister: Given value is wider than register " + signalName);
    return;
  }
  status = 0;
}

// src/ToucanSim.cu:280-289 + read_reg_from_gpu (src/SimEval.cu:1058-1069).
uint64_t ToucanSimulator::readRegister(const std::string &signalName) {
  if (!ready("readRegister")) return 0;
  std::vector<uint32_t> ids;
  std::string err;
  if (!regIds(symbols, signalName, ids, err)) {
    fail(TGS_E_ARG, "readRegister: " + err);
    return 0;
  }
  std::vector<uint8_t> vals(ids.size());
  int rc = tgs_read_regs(ctx, 0, ids.data(), uint32_t(ids.size()), vals.data());
  if (rc != TGS_OK) {
    fail(rc, "readRegister failed: " + runtimeError(ctx));
    return 0;
  }
  status = 0;
  uint64_t r = 0;
  for (auto v : vals) r = (r << 4) | v;
  return r;
}

bool ToucanSimulator::step(const std::vector<std::pair<std::string, uint64_t>> &writes, uint32_t cycles,
                           const std::vector<std::string> &reads, std::vector<uint64_t> &out) {
  out.assign(reads.size(), 0);
  if (!ready("step")) return false;
  if (dumpVCD) {  // every cycle goes through eval() so the dump sees it
    for (auto &w : writes) {
      setRegister(w.first, w.second);
      if (status != 0) return false;
    }
    int stopped = 0;
    for (uint32_t k = 0; k < cycles; ++k) {
      stopped = eval() ? 1 : 0;
      if (status != 0) return false;
    }
    if (cycles == 0 && tgs_get_stop(ctx, 0, &stopped) != TGS_OK) stopped = 0;
    for (size_t k = 0; k < reads.size(); ++k) {
      out[k] = readRegister(reads[k]);
      if (status != 0) return false;
    }
    return stopped != 0;
  }
  std::vector<uint32_t> wids, rids, rlen, ids;
  std::vector<uint8_t> wvals;
  std::string err;
  for (auto &w : writes) {
    if (!regIds(symbols, w.first, ids, err)) return fail(TGS_E_ARG, "step: " + err);
    uint64_t v = w.second;
    const size_t base = wvals.size();
    wvals.resize(base + ids.size());
    wids.insert(wids.end(), ids.begin(), ids.end());
    for (size_t k = ids.size(); k-- > 0;) {
      wvals[base + k] = uint8_t(v & 0xF);
      v >>= 4;
    }
    if (v != 0) return fail(TGS_E_ARG, "step: Given value is wider than register " + w.first);
  }
  for (auto &r : reads) {
    if (!regIds(symbols, r, ids, err)) return fail(TGS_E_ARG, "step: " + err);
    rids.insert(rids.end(), ids.begin(), ids.end());
    rlen.push_back(uint32_t(ids.size()));
  }
  std::vector<uint8_t> rvals(rids.size());
  int stopped = 0;
  int rc = tgs_step(ctx, wids.data(), uint32_t(wids.size()), wvals.data(), cycles, rids.data(),
                    uint32_t(rids.size()), rvals.data(), &stopped);
  if (rc != TGS_OK) return fail(rc, "step failed: " + runtimeError(ctx));
  status = 0;
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
  const auto oldParts = probeParts, oldSlots = probeSlots;
  const auto oldFirst = probeFirst;
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
    int rc = tgs_set_probes(ctx, probeParts.data(), probeSlots.data(), uint32_t(probeParts.size()));
    if (rc != TGS_OK) {  // the runtime keeps its old list: so does the class
      fail(rc, "signal probes failed: " + runtimeError(ctx));
      probeParts = oldParts;
      probeSlots = oldSlots;
      probeFirst = oldFirst;
      return 0;
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
  if (!ready("readSignal")) return UINT64_MAX;
  auto it = symbols.signalDebugInfo.find(signalName);
  if (it == symbols.signalDebugInfo.end()) {
    fail(TGS_E_ARG, "readSignal: unknown signal " + signalName);
    return UINT64_MAX;
  }
  const auto &locs = it->second;
  uint32_t width = 0;
  for (auto &f : locs) width += std::get<2>(f);
  if (width > 64) {
    fail(TGS_E_ARG, "readSignal: Signal wider than 64 bits are not supported by this function: " + signalName);
    return UINT64_MAX;
  }
  if (!probeFirst.count(signalName) && (addSignalProbes({signalName}), status != 0)) return UINT64_MAX;
  status = 0;
  if (probeEpoch == cycle_cnt) return UINT64_MAX;
  std::vector<uint8_t> vals(probeParts.size());
  int rc = tgs_read_probes(ctx, 0, vals.data(), uint32_t(vals.size()));
  if (rc != TGS_OK) {
    fail(rc, "readSignal failed: " + runtimeError(ctx));
    return UINT64_MAX;
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
  if (!ready("enableVCDDump")) return;
  vcdOfs = new std::ofstream(vcdFilename, std::ios::out);
  if (!vcdOfs->is_open()) {
    delete vcdOfs;
    vcdOfs = nullptr;
    fail(TGS_E_IO, "enableVCDDump: cannot open " + vcdFilename);
    return;
  }
  std::vector<std::pair<uint64_t, std::string>> sigs;
  for (auto &kv : symbols.signalDebugInfo)
    if (!kv.second.empty())
      sigs.push_back({(uint64_t(std::get<0>(kv.second[0])) << 32) + std::get<1>(kv.second[0]), kv.first});
  std::sort(sigs.begin(), sigs.end());
  std::vector<std::string> sigNames;
  for (auto &x : sigs) sigNames.push_back(x.second);
  addSignalProbes(sigNames);
  if (status != 0) {
    closeVcd();
    return;
  }
  dumpVCD = true;
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
  int rc = tgs_snapshot_begin(ctx, 0, &t);
  if (rc != TGS_OK) {
    fail(rc, "VCD snapshot failed: " + runtimeError(ctx));
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
  int rc = tgs_snapshot_wait(ctx, ticket, &data, nullptr);
  if (rc != TGS_OK) {
    fail(rc, "VCD snapshot failed: " + runtimeError(ctx));
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
    if (g_current == ctx) g_current = nullptr;
    tgs_destroy(ctx);
    ctx = nullptr;
  }
  loadedOk = false;
}

// src/ToucanSim.cu:306-325: the launch profile (shape, kernel time, the
// barrier-stall share over all workgroups), then, when the per-partition
// profile was collected, one line per partition as the reference prints it.
void ToucanSimulator::printProfile() {
  tgs_profile p{};
  if (!ctx || tgs_get_profile(ctx, &p) != TGS_OK) return;
  double stall = p.total_ticks ? 100.0 * double(p.barrier_ticks) / double(p.total_ticks) : 0.0;
  std::cout << "kernel " << p.kernel << " grid " << p.grid << " block " << p.block << " lds " << p.lds_bytes
            << " resident " << p.resident << " launches " << p.launches << " cycles " << p.cycles << " kernel_ms "
            << p.kernel_ms << " barrier_stall " << stall << "%\n";
  const size_t n = design.parts.size();
  std::vector<int64_t> useful(n), total(n);
  uint32_t collected = 0;
  if (n == 0 || tgs_get_part_profile(ctx, useful.data(), total.data(), n, &collected) != TGS_OK || collected == 0)
    return;
  for (size_t part = 0; part < n; ++part) {
    const float ratio = total[part] ? float(useful[part] * 100.0 / total[part]) : 0.0f;
    std::cout << "Part " << part << ", useful " << ratio << "%, ticks " << useful[part] << ", total ticks "
              << total[part] << "\n";
  }
}

}  // namespace toucanGPUSim

// ---- include/SimEval.h: the reference's host shim over the current context ----

using toucanGPUSim::g_current;
using toucanGPUSim::g_shimOwned;

static bool shimReady(const char *call) {
  if (g_current) return true;
  std::cerr << call << ": no design on the device (init or copy_netlist_to_gpu first)" << std::endl;
  return false;
}

static void shimError(const char *call) {
  std::cerr << call << " failed: " << toucanGPUSim::runtimeError(g_current) << std::endl;
}

uint64_t read_reg_from_gpu(const std::vector<std::tuple<uint32_t, uint32_t>> &signalLocs) {
  if (!shimReady("read_reg_from_gpu")) return 0;
  std::vector<uint32_t> ids;
  for (auto &l : signalLocs) ids.push_back(std::get<0>(l));
  std::vector<uint8_t> vals(ids.size());
  if (tgs_read_regs(g_current, 0, ids.data(), uint32_t(ids.size()), vals.data()) != TGS_OK) {
    shimError("read_reg_from_gpu");
    return 0;
  }
  uint64_t r = 0;
  for (auto v : vals) r = (r << 4) | v;
  return r;
}

void write_reg_to_gpu(const std::vector<std::tuple<uint32_t, uint32_t>> &signalLocs, uint64_t signalValue) {
  if (!shimReady("write_reg_to_gpu")) return;
  std::vector<uint32_t> ids;
  for (auto &l : signalLocs) ids.push_back(std::get<0>(l));
  std::vector<uint8_t> vals(ids.size());
  for (size_t k = ids.size(); k-- > 0;) {
    vals[k] = uint8_t(signalValue & 0xF);
    signalValue >>= 4;
  }
  if (tgs_write_regs(g_current, 0, ids.data(), uint32_t(ids.size()), vals.data()) != TGS_OK)
    shimError("write_reg_to_gpu");
  else if (signalValue != 0)
    std::cerr << "write_reg_to_gpu: Given value is wider than register" << std::endl;
}

void copy_netlist_to_gpu(toucanGPUSim::SimDesignInfo &design) {
  if (!g_current) {
    if (!g_shimOwned && tgs_create(0, &g_shimOwned) != TGS_OK) {
      g_shimOwned = nullptr;
      std::cerr << "copy_netlist_to_gpu: no usable GPU" << std::endl;
      return;
    }
    g_current = g_shimOwned;
  }
  std::ostringstream os;
  toucanGPUSim::serializeSimDesignInfo(os, design);
  const std::string bytes = os.str();
  int print = 0;
  if (tgs_load(g_current, reinterpret_cast<const uint8_t *>(bytes.data()), bytes.size(), nullptr, 0, 0, print) !=
      TGS_OK) {
    shimError("copy_netlist_to_gpu");
    return;
  }
  // the caller's pools (after its Init / Randomize), as the reference uploads them
  if ((!design.regPool.empty() &&
       tgs_write_pool(g_current, 0, TGS_POOL_REG, 0, design.regPool.size(), design.regPool.data()) != TGS_OK) ||
      (!design.memPool.empty() &&
       tgs_write_pool(g_current, 0, TGS_POOL_MEM, 0, design.memPool.size(), design.memPool.data()) != TGS_OK))
    shimError("copy_netlist_to_gpu");
}

bool get_eval_done() {
  int stopped = 0;
  if (!shimReady("get_eval_done")) return false;
  if (tgs_get_stop(g_current, 0, &stopped) != TGS_OK) shimError("get_eval_done");
  return stopped != 0;
}

uint32_t get_real_cycles() {
  uint64_t c = 0;
  if (!shimReady("get_real_cycles")) return 0;
  if (tgs_get_cycles(g_current, 0, &c) != TGS_OK) shimError("get_real_cycles");
  return uint32_t(c);
}

void setEnablePrint(bool print_en) {
  if (!shimReady("setEnablePrint")) return;
  if (tgs_set_print(g_current, print_en ? 1 : 0) != TGS_OK) shimError("setEnablePrint");
}

// Exported whatever the caller's ENABLE_SIM_PROFILE (declared under it in
// include/SimEval.h, as in the reference).
void copy_profile_data(int64_t *usefulTicks, int64_t *totalTicks, size_t count) {
  if (!shimReady("copy_profile_data")) return;
  uint32_t collected = 0;
  if (tgs_get_part_profile(g_current, usefulTicks, totalTicks, count, &collected) != TGS_OK)
    shimError("copy_profile_data");
}
