// Device constants: drop-in for the reference's include/GPUConsts.h:1-3.
//
// GPUMemPaddingSize is defined there and read nowhere in the reference
// sources; it is kept so code that includes this header or names the macro
// still compiles. The MI355X runtime pads its own device buffers (16-byte
// aligned program chunks, 4-nibble register and exchange words) where its
// kernels need it.
#pragma once

#define GPUMemPaddingSize 128
