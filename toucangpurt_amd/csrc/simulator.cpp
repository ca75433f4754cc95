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
  dumpVCD = false;
  vcdOfs = nullptr;
  cycle_cnt = 0;
  printsSeen = 0;
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
  tgs_write_regs(ctx, 0, ids.data(), uint32_t(ids.size()), vals.data());
}

// src/ToucanSim.cu:280-289 + read_reg_from_gpu (src/SimEval.cu:1058-1069).
uint64_t ToucanSimulator::readRegister(const std::string &signalName) {
  assert(symbols.regDebugInfo.count(signalName));
  auto &locs = symbols.regDebugInfo[signalName];
  assert(locs.size() <= 16 && "Register wider than 64 bits are not supported by this function");
  std::vector<uint32_t> ids;
  for (auto &l : locs) ids.push_back(std::get<0>(l));
  std::vector<uint8_t> vals(ids.size());
  tgs_read_regs(ctx, 0, ids.data(), uint32_t(ids.size()), vals.data());
  uint64_t r = 0;
  for (auto v : vals) r = (r << 4) | v;
  return r;
}

// src/ToucanSim.cu:291-300: value-pool signals are not observable after a cycle.
uint64_t ToucanSimulator::readSignal(const std::string &) {
  std::cerr << "Not implemented! This might not be necessary" << std::endl;
  return UINT64_MAX;
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

// VCD of every register in the symbol table (unsupport/SimVCD.cpp:99-176):
// dotted names become nested scopes, names without a dot go in scope "top".
void ToucanSimulator::enableVCDDump(const std::string vcdFilename) {
  dumpVCD = true;
  vcdOfs = new std::ofstream(vcdFilename, std::ios::out);
  std::vector<std::string> names;
  for (auto &kv : symbols.regDebugInfo) names.push_back(kv.first);
  std::sort(names.begin(), names.end());
  vcdNames = names;
  vcdRegs.clear();
  vcdWidth.clear();
  vcdIds.clear();
  for (uint32_t i = 0; i < names.size(); ++i) {
    vcdRegs.push_back(symbols.regDebugInfo[names[i]]);
    uint32_t w = 0;
    for (auto &f : vcdRegs.back()) w += std::get<1>(f);
    vcdWidth.push_back(w);
    vcdIds.push_back(vcdIdentifier(i));
  }
  vcdLast.assign(names.size(), {});
  *vcdOfs << "$date\n    " << __DATE__ << " " << __TIME__ << "\n$end\n"
          << "$version\n    VCD generator\n$end\n"
          << "$timescale\n    1ps\n$end\n";
  // scope tree
  struct Node {
    std::map<std::string, Node> kids;
    std::vector<uint32_t> sigs;
  } root;
  for (uint32_t i = 0; i < names.size(); ++i) {
    Node *n = &root;
    std::string rest = names[i];
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
// dump is pipelined: this cycle's register pool is snapshotted
// asynchronously (pinned copy on a side stream) and the previous snapshot is
// written, so the copy and the formatting overlap the next eval.
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
  const uint8_t *pool = nullptr;
  if (tgs_snapshot_wait(ctx, ticket, &pool, nullptr) != TGS_OK) {
    std::cerr << "VCD snapshot failed: " << lastError() << std::endl;
    return;
  }
  *vcdOfs << "#" << stamp << "\n";
  for (uint32_t i = 0; i < vcdRegs.size(); ++i) {
    std::vector<uint8_t> v;
    for (auto &f : vcdRegs[i]) v.push_back(pool[std::get<0>(f)]);
    if (!vcdLast[i].empty() && vcdLast[i] == v) continue;
    vcdLast[i] = v;
    if (vcdWidth[i] == 1) {
      *vcdOfs << int(v[0] & 1) << vcdIds[i] << "\n";
    } else {
      *vcdOfs << "b";
      for (size_t k = 0; k < v.size(); ++k) {
        uint32_t w = std::get<1>(vcdRegs[i][k]);
        for (int b = int(w) - 1; b >= 0; --b) *vcdOfs << char('0' + ((v[k] >> b) & 1));
      }
      *vcdOfs << " " << vcdIds[i] << "\n";
    }
  }
}

void ToucanSimulator::cleanup() {
  if (vcdOfs) {
    if (vcdPending && ctx) writeVcdSnapshot(vcdTicket, vcdStamp);
    vcdOfs->close();
    delete vcdOfs;
    vcdOfs = nullptr;
  }
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
