// Driver for the reference's host code (test infrastructure only).
//
// Compiled together with the reference's own src/Serializer.cpp and
// src/SimInit.cpp (see Makefile, target `ref`) so tests can check this
// project's format writer and Randomize against the reference's functions
// byte for byte. Only those two reference files are used; nothing here
// replaces any part of the reference.
#include "ToucanGPUGenDataTypes.h"

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>

using namespace toucanGPUSim;

static int to_heap(const std::string &s, uint8_t **out, size_t *n) {
  *n = s.size();
  *out = static_cast<uint8_t *>(std::malloc(s.size() ? s.size() : 1));
  if (!*out) return -1;
  std::memcpy(*out, s.data(), s.size());
  return 0;
}

extern "C" {

// deserializeSimDesignInfo (Serializer.cpp:226) then serializeSimDesignInfo (:199).
int ref_roundtrip_design(const uint8_t *in, size_t n, uint8_t **out, size_t *outn) {
  SimDesignInfo d;
  std::istringstream is(std::string(reinterpret_cast<const char *>(in), n));
  deserializeSimDesignInfo(is, d);
  if (!is) return -1;
  std::ostringstream os;
  serializeSimDesignInfo(os, d);
  return to_heap(os.str(), out, outn);
}

// deserializeSimDebugInfo (Serializer.cpp:303) then serializeSimDebugInfo (:264).
int ref_roundtrip_symbols(const uint8_t *in, size_t n, uint8_t **out, size_t *outn) {
  SimDebugInfo d;
  std::istringstream is(std::string(reinterpret_cast<const char *>(in), n));
  deserializeSimDebugInfo(is, d);
  if (!is) return -1;
  std::ostringstream os;
  serializeSimDebugInfo(os, d);
  return to_heap(os.str(), out, outn);
}

// Init (SimInit.cpp:16) then Randomize(seed) (SimInit.cpp:28); copies the pools out.
int ref_randomize(const uint8_t *design, size_t dn, const uint8_t *symbols, size_t sn, uint32_t seed,
                  uint8_t *reg_out, size_t reg_n, uint8_t *mem_out, size_t mem_n) {
  SimDesignInfo d;
  SimDebugInfo s;
  std::istringstream id(std::string(reinterpret_cast<const char *>(design), dn));
  deserializeSimDesignInfo(id, d);
  std::istringstream is(std::string(reinterpret_cast<const char *>(symbols), sn));
  deserializeSimDebugInfo(is, s);
  if (!id || !is) return -1;
  d.Init();
  d.Randomize(seed, s);
  if (reg_n != d.regPool.size() || mem_n != d.memPool.size()) return -2;
  if (reg_n) std::memcpy(reg_out, d.regPool.data(), reg_n);
  if (mem_n) std::memcpy(mem_out, d.memPool.data(), mem_n);
  return 0;
}

void ref_free(uint8_t *p) { std::free(p); }

}  // extern "C"
