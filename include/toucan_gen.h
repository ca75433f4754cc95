/* Seeded synthetic design generator: the stand-in for the external Toucan
 * compiler ("ToucanGPUGen", include/ToucanGPUGenDataTypes.h:285-286 of the
 * reference, not shipped with it). It emits the same serialized program
 * format the runtime loads (GPUSimDesign.bin / GPUSimIOSymbols.bin,
 * src/Serializer.cpp:199-300), plus a logical netlist file that the
 * netlist-level test simulator evaluates independently of the program.
 *
 * No counterpart in the reference's runtime: it replaces the compiler side.
 * All functions return 0 on success and a negative TGS_E_* code on error.
 */
#ifndef TOUCAN_GEN_H
#define TOUCAN_GEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tgs_gen_params {
  uint64_t seed;
  uint32_t n_luts;          /* LUT4 count (1-bit, 4-input, uint16 truth table) */
  uint32_t depth;           /* LUT levels per cycle, split over the two regions */
  uint32_t n_regs;          /* 1-bit state registers */
  uint32_t n_inputs;        /* 1-bit primary inputs (stimulus-driven) */
  uint32_t n_vecops;        /* word-level ops (add, sub, mux, eq/lt, shift) */
  uint32_t word_bits_mix;   /* percent of word registers that are 64-bit (rest 32-bit) */
  uint32_t n_srams;         /* SRAM arrays, 1 read + 1 write port each */
  uint32_t sram_depth;      /* words per SRAM, power of two, <= 65536 */
  uint32_t sram_width;      /* bits per word, multiple of 4, <= 128 */
  uint32_t part_luts;       /* target LUT4s per partition */
  uint32_t cluster_luts;    /* LUT4s per logic cluster (partition-independent atom) */
  uint32_t sources_a;       /* register/input sources per region-0 cluster */
  uint32_t sources_b;       /* exchange sources per region-1 cluster */
  uint32_t word_part_ops;   /* word ops per word partition */
  uint32_t n_prints;        /* print ops (test designs) */
  uint32_t n_stops;         /* stop ops (test designs) */
  uint32_t stop_rarity;     /* stop enables are AND of this many random bits */
  uint32_t reserved[8];
} tgs_gen_params;

typedef struct tgs_gen_stats {
  uint64_t n_parts[2];       /* partitions per region */
  uint64_t n_luts;           /* LUT4s emitted */
  uint64_t n_vecops;         /* word ops emitted */
  uint64_t n_lut_ops;        /* nibble LUT ops (top + middle) */
  uint64_t n_microparts;
  uint64_t reg_pool;         /* nibbles */
  uint64_t mem_pool;
  uint64_t exchange_pool;
  uint64_t max_value_pool;
  uint64_t design_bytes;
  uint64_t symbol_bytes;
  uint64_t netlist_bytes;
  uint64_t n_inputs;
  uint64_t reserved[8];
} tgs_gen_stats;

typedef struct tgs_gen_result tgs_gen_result;

/* Defaults: the 20,000-LUT4 design (depth 32, 256 registers, 64 inputs). */
int tgs_gen_default_params(tgs_gen_params *p);

/* Build a design in memory. */
int tgs_gen_build(const tgs_gen_params *p, tgs_gen_result **out);

/* which: 0 = design .bin, 1 = symbols .bin, 2 = netlist, 3 = input register
 * ids (uint32 array, stimulus order). Pointers stay valid until free. */
int tgs_gen_get(const tgs_gen_result *r, int which, const uint8_t **data, size_t *size);
int tgs_gen_get_stats(const tgs_gen_result *r, tgs_gen_stats *s);
int tgs_gen_error(const tgs_gen_result *r, char *buf, size_t len);
void tgs_gen_free(tgs_gen_result *r);

#ifdef __cplusplus
}
#endif

#endif /* TOUCAN_GEN_H */
