/* CPU oracle of the Toucan cycle-simulation path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker: tests/, the smoke check and
 * bench.py's CPU-baseline leg may use it; the product (toucangpurt_amd) never
 * links, loads or calls it.
 *
 * It is a plain-C restatement of the reference evaluator
 * (src/SimEval.cu:304-761 micropart/L0/last level, :765-864 partition,
 * :889-995 cycle loop) over the serialized program (src/Serializer.cpp:199-259),
 * with its own parser, so it shares no code with the runtime it checks.
 * It also records what the reference leaves undefined: reads of value-pool
 * slots not written in the current partition visit, and memory addresses
 * touched by two partitions of one region.
 */
#ifndef TOUCAN_ORACLE_H
#define TOUCAN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_prog orc_prog;
typedef struct orc_state orc_state;

/* Parse a design image. Returns 0, or -1 with a message in err. */
int orc_prog_load(const uint8_t *buf, size_t size, orc_prog **out, char *err, size_t errlen);
void orc_prog_free(orc_prog *p);
uint32_t orc_pool_size(const orc_prog *p, int which); /* 0 reg, 1 mem, 2 exchange */
uint32_t orc_num_parts(const orc_prog *p, int region);

/* New state: zeroed pools (src/SimInit.cpp:16-26). */
int orc_state_new(const orc_prog *p, orc_state **out);
void orc_state_free(orc_state *s);
int orc_state_copy_pool(orc_state *s, int which, int to_state, uint8_t *buf, size_t n);

/* Stimulus (no reference counterpart; the synthetic workload driver):
 * before each cycle s = xorshift64(s) once per 64 inputs, input i takes bit
 * i%64 of draw i/64; s starts at splitmix64(seed + instance). */
int orc_set_stimulus(orc_state *s, const uint32_t *ids, uint32_t n, uint64_t seed, uint32_t instance);

/* Run `cycles` cycles. Returns the sticky stop flag. */
int orc_run(orc_state *s, uint32_t cycles, int enable_print);
/* The same with the partitions of each region split over `threads` threads
 * and a barrier between regions; results equal orc_run's. */
int orc_run_threads(orc_state *s, uint32_t cycles, int enable_print, int threads);
uint64_t orc_cycles(const orc_state *s);
int orc_stopped(const orc_state *s);
uint64_t orc_stop_cycle(const orc_state *s);

/* Undefined-behaviour counters: [0] reads of unwritten value-pool slots,
 * [1] memory addresses shared by partitions of one region with a write,
 * [2] out-of-range memory accesses, [3] out-of-range LUT positions. */
void orc_violations(const orc_state *s, uint64_t out[4]);

/* Signal probes (value-pool slot values a partition leaves behind in its
 * last visit, SimDebugInfo::signalDebugInfo's (part, valId) locations):
 * set the list, read the n values in list order after a run. */
int orc_set_probes(orc_state *s, const uint32_t *parts, const uint32_t *slots, uint32_t n);
void orc_read_probes(const orc_state *s, uint8_t *out);
/* Slots every visit of `part` writes; returns the count (stores up to cap). */
uint32_t orc_written_slots(const orc_prog *p, uint32_t part, uint32_t *out, uint32_t cap);

/* Print log: (cycle, msg) pairs. Returns the number stored in total. */
uint64_t orc_print_log(const orc_state *s, uint64_t *cycles, uint32_t *msgs, uint64_t cap);

/* Reference utilities restated for tests. */
uint64_t orc_xorshift64(uint64_t x);
uint64_t orc_splitmix64(uint64_t x);

#ifdef __cplusplus
}
#endif

#endif
