// Kernel argument blocks and host launchers of the gfx950 cycle kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.h"

namespace tgs {

struct PlainArgs {
  CommonArgs c;
  const uint8_t *net;
  const PlainPart *parts;
  const uint32_t *levels;
  const uint32_t *mparts;
  uint32_t n0, nTotal;
};

struct W64Args {
  CommonArgs c;
  const uint32_t *words;     // wave64 image
  const uint32_t *partOff;   // word offset per partition
  const uint32_t *partWords; // words per partition
  const uint32_t *schedOff;  // groups * 2 + 1
  const uint32_t *schedPart;
  const uint32_t *schedLds;
  uint32_t vpBytes;          // value-pool bytes in LDS (16-byte multiple)
  uint32_t barrierBase;      // barrier index offset (counter is zeroed per launch)
};

hipError_t launchPlain(const PlainArgs &args, uint32_t grid, uint32_t block, uint32_t lds, hipStream_t stream);
hipError_t plainMaxBlocksPerCU(int *blocks, uint32_t block, uint32_t lds);
hipError_t plainSetLds(uint32_t lds);

hipError_t launchWave64(const W64Args &args, bool resident, uint32_t grid, uint32_t block, uint32_t lds,
                        hipStream_t stream);
hipError_t wave64MaxBlocksPerCU(int *blocks, bool resident, uint32_t block, uint32_t lds);
hipError_t wave64SetLds(uint32_t lds);
uint32_t wave64LutBytes();

}  // namespace tgs
