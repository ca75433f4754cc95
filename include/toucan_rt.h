/* C-ABI of the MI355X Toucan cycle-simulation runtime (libtoucan_rt.so).
 *
 * This is the lower boundary under toucanGPUSim::ToucanSimulator. It replaces
 * the reference's C++-linkage kernel shim (include/SimEval.h:21-39,
 * src/SimEval.cu:1058-1343) and the launch code of src/ToucanSim.cu:16-267,
 * with plain pointers and sizes, an opaque per-design context instead of
 * process globals, and negative status codes instead of exit()/assert().
 * Every entry point names the reference function it replaces.
 *
 * Threading: a context is used by one host thread at a time. All device work
 * of a context is ordered on the context's own stream.
 */
#ifndef TOUCAN_RT_H
#define TOUCAN_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TGS_OK 0
#define TGS_E_ARG (-1)       /* bad argument */
#define TGS_E_IO (-2)        /* file missing or unreadable (ToucanSim.cu:108-122 returns -1) */
#define TGS_E_FORMAT (-3)    /* truncated / malformed program (reference: undefined) */
#define TGS_E_LIMIT (-4)     /* value pool or LDS over the limit (ToucanSim.cu:153,194-197) */
#define TGS_E_HIP (-5)       /* HIP runtime error (reference: gpuErrchk exit, SimEval.h:11-19) */
#define TGS_E_NODEV (-6)     /* no device / no cooperative launch (ToucanSim.cu:16-43) */
#define TGS_E_BUDGET (-7)    /* HBM budget exhausted (serving admission) */
#define TGS_E_STATE (-8)     /* call out of order (e.g. eval before load) */
#define TGS_E_INVALID (-9)   /* program violates a format invariant (SimEval.cu:92-97,1183,1292-1311) */
#define TGS_E_UNSUPPORTED (-10)

/* kernel families */
#define TGS_KERNEL_PLAIN 0   /* reference execution model: block per partition, 32-lane micropart per wave */
#define TGS_KERNEL_WAVE64 1  /* persistent wave64 kernel, program evaluated from LDS */
#define TGS_KERNEL_BATCH 2   /* instance-per-lane batched kernel */
#define TGS_KERNEL_STREAM 3  /* persistent kernel, program rows streamed into LDS by a loader wave */

/* pools */
#define TGS_POOL_REG 0
#define TGS_POOL_MEM 1
#define TGS_POOL_EXCHANGE 2

typedef struct tgs_ctx tgs_ctx;

typedef struct tgs_program_stats {
  uint64_t n_parts[2];
  uint64_t n_top, n_mid, n_last;          /* LUT op records */
  uint64_t n_vecread, n_vecop, n_memread; /* word-level records */
  uint64_t n_regread, n_exgread, n_regwrite, n_exgwrite, n_memwrite, n_print, n_stop;
  uint64_t n_microparts, n_exec_levels;
  uint64_t b_prog;       /* F1 record bytes read per cycle (shared by instances) */
  uint64_t b_state;      /* state nibbles moved per cycle per instance, memory ops counted enabled */
  uint64_t max_value_pool;
  uint64_t reg_pool, mem_pool, exchange_pool;
} tgs_program_stats;

typedef struct tgs_profile {
  double kernel_ms;       /* sum of kernel durations measured with HIP events on the context stream */
  uint64_t launches;      /* kernel launches measured */
  uint64_t cycles;        /* simulated cycles covered by those launches (summed over instances) */
  uint32_t grid, block;   /* launch shape of the last launch */
  uint32_t lds_bytes;     /* dynamic LDS of the last launch */
  uint32_t kernel;        /* kernel family of the last launch */
  uint32_t resident;      /* program source: 0 global, 1 staged in LDS per visit, 2 LDS-resident, 3 streamed */
  uint64_t prog_bytes;    /* device program image bytes */
  uint64_t barrier_ticks; /* wave64 / stream: real-time ticks spent waiting at grid barriers, summed over
                             every workgroup of every launch */
  uint64_t total_ticks;   /* wave64 / stream: real-time ticks of the cycle loop, summed the same way */
  uint32_t batch_group;   /* batched: instances per wave (each wave evaluates 64 / batch_group ops per round) */
  uint32_t batch_fast_tasks; /* batched: LUT micropartition pairs evaluated without index bounds checks
                                (every value a nibble; 0 when the program or the state does not allow it) */
  uint32_t batch_nibble_rows; /* batched: 1 when the rows hold a nibble per instance (every LUT task
                                 evaluated without bounds checks), 0 when a byte */
  uint32_t reserved;
} tgs_profile;

/* ---- program inspection (host only, no device needed) ---- */

/* Parse and validate a design image; fill counts. The byte model is the one
 * the roofline uses (record sizes of include/ToucanGPUGenDataTypes.h). */
int tgs_program_stats_from_bytes(const uint8_t *design, size_t size, tgs_program_stats *out,
                                 char *err, size_t errlen);

/* Init (src/SimInit.cpp:16-26) and optionally Randomize(seed) (:28-64) of a
 * design, host only: writes the initial register and memory pools. */
int tgs_initial_state(const uint8_t *design, size_t design_size, const uint8_t *symbols, size_t symbols_size,
                      int randomize, uint32_t seed, uint8_t *reg, size_t reg_size, uint8_t *mem, size_t mem_size);

/* ---- context ---- */

/* setupGPU (src/ToucanSim.cu:16-100): select device, check cooperative launch. */
int tgs_create(int device, tgs_ctx **out);
/* The same on the caller's stream (a hipStream_t of `device`, passed as a
 * pointer): every copy and launch of the context is ordered on it, so the
 * simulator interleaves with the caller's own work (the reference runs on the
 * default stream, src/ToucanSim.cu:206-267). The stream stays the caller's:
 * tgs_destroy does not destroy it, and it must outlive the context. */
int tgs_create_on_stream(int device, void *stream, tgs_ctx **out);
/* cleanup (src/ToucanSim.cu:302-304, a no-op there): frees all device memory. */
void tgs_destroy(tgs_ctx *ctx);
/* Text of the last error on this context (empty if none). */
int tgs_last_error(const tgs_ctx *ctx, char *buf, size_t len);

/* init (src/ToucanSim.cu:102-204) + copy_netlist_to_gpu (src/SimEval.cu:1114-1343):
 * parse, Init, optional Randomize(0), validate, build device images, upload.
 * symbols may be NULL (then randomize must be 0). */
int tgs_load(tgs_ctx *ctx, const uint8_t *design, size_t design_size, const uint8_t *symbols,
             size_t symbols_size, int randomize, int enable_print);
int tgs_load_files(tgs_ctx *ctx, const char *design_path, const char *symbol_path, int randomize,
                   int enable_print);

/* Choose the kernel family used by eval / eval_free_running. Until a
 * family is chosen the streamed kernel runs, and a design whose value pool
 * or step stream does not fit its LDS budget runs on the wave64 kernel
 * instead; a chosen family that cannot run fails with TGS_E_LIMIT. */
int tgs_set_kernel(tgs_ctx *ctx, int kernel);
/* Launch shape of a kernel family: threads per workgroup (multiple of 64,
 * <= 1024) and workgroups (0 = automatic). The autotuner's knob; the
 * reference fixes 1024 threads and one block per SM (ToucanSim.cu:56-80). */
int tgs_set_launch(tgs_ctx *ctx, int kernel, uint32_t block, uint32_t groups);
/* setEnablePrint (src/SimEval.cu:1095-1097). */
int tgs_set_print(tgs_ctx *ctx, int enable);

/* Number of independent stimulus instances held resident (default 1).
 * Resets every instance to the loaded initial state. */
int tgs_set_instances(tgs_ctx *ctx, uint32_t n);

/* Primary-input stimulus: before every cycle, per instance, s = xorshift64(s)
 * once per 64 inputs; input i takes bit (i mod 64) of draw i/64. Instance k
 * starts from splitmix64(seed + k). n == 0 disables stimulus. The reference
 * has no stimulus path; this is the synthetic workload driver. */
int tgs_set_stimulus(tgs_ctx *ctx, const uint32_t *input_ids, uint32_t n, uint64_t seed);

/* eval (src/ToucanSim.cu:206-230): one cycle of every instance. *stopped gets
 * the sticky stop flag of instance 0 (get_eval_done, src/SimEval.cu:1083-1087). */
int tgs_eval(tgs_ctx *ctx, int *stopped);
/* eval_free_running (src/ToucanSim.cu:232-267): `cycles` cycles of every instance. */
int tgs_eval_free_running(tgs_ctx *ctx, uint32_t cycles, int *stopped);
/* Batched run: instance k advances cycles[k] cycles (uneven counts). */
int tgs_run_batch(tgs_ctx *ctx, const uint32_t *cycles, uint32_t n, int *any_stopped);

/* No counterpart in the reference (the serving form of batching): tgs_run_batch
 * of n batched contexts of one device (up to 64, distinct, typically one per
 * design) in shared persistent launches, so that small batches of different
 * designs occupy the GPU together instead of one after another. cycles[k]
 * holds one count per instance of ctxs[k]; any_stopped (null or n entries)
 * gets each context's flag as tgs_run_batch would. Every launch cycle runs
 * region 0 of every context, then region 1; results are those of
 * tgs_run_batch on each context alone. Contexts whose images were built for
 * different instances per wave run in separate launches. The launches use
 * the first context's stream after the other contexts' streams have
 * drained; each context's profile counts the shared launch time. */
int tgs_run_batch_multi(tgs_ctx *const *ctxs, uint32_t n, const uint32_t *const *cycles, int *any_stopped);

/* No counterpart in the reference: device memory (bytes) the context holds
 * now (program images, instance state, logs, probes, staging), for callers
 * that budget HBM across many resident contexts. */
int tgs_device_bytes(tgs_ctx *ctx, uint64_t *bytes);

/* No counterpart in the reference: tgs_run_batch for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as tgs_run_batch on a batched context; the
 * per-launch scratch of the batched kernel (slot cycle counts and bases, the
 * active-group table, the grid-barrier tree, the launch arguments) lives
 * in a static array of 320,256 bytes per GPU, held until the process ends, instead of context
 * allocations, and no hipMalloc / hipFree or other allocating call is made.
 * At most 4,096 instances (TGS_E_LIMIT beyond); the batched program image
 * must be built, which tgs_set_kernel(TGS_KERNEL_BATCH), tgs_set_instances,
 * tgs_set_stimulus and tgs_set_launch do (TGS_E_STATE otherwise). Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (from
 * several threads or contexts) race and must be serialized by the caller. */
int tgs_run_batch_noalloc(tgs_ctx *ctx, const uint32_t *cycles, uint32_t n, int *any_stopped);

/* read_reg_from_gpu / write_reg_to_gpu (src/SimEval.cu:1058-1081), batched:
 * one nibble per id, no per-nibble memcpy. */
int tgs_read_regs(tgs_ctx *ctx, uint32_t instance, const uint32_t *ids, uint32_t n, uint8_t *out);
int tgs_write_regs(tgs_ctx *ctx, uint32_t instance, const uint32_t *ids, uint32_t n, const uint8_t *in);
/* Testbench step of every instance in one call: the reference's
 * setRegister -> eval -> readRegister loop (src/ToucanSim.cu:206-289) makes a
 * synchronous copy per nibble for each register and a synchronizing launch
 * per cycle. Writes register write_ids[k] of instance i to
 * write_vals[i * n_write + k] (ids distinct), runs `cycles` cycles of every
 * instance (0: none), then reads register read_ids[k] of instance i into
 * read_vals[i * n_read + k]. The ids and values travel through one pinned
 * mailbox. One instance on the streamed or wave64 kernel with cycles > 0:
 * one launch, which applies the writes before its first cycle and takes the
 * reads after its last from and to the mailbox, and one synchronization.
 * Otherwise one host-to-device copy and one scatter kernel before the run,
 * one gather kernel and one device-to-host copy after it, one
 * synchronization. *any_stopped (may be null): some instance's sticky stop
 * flag is set. */
int tgs_step(tgs_ctx *ctx, const uint32_t *write_ids, uint32_t n_write, const uint8_t *write_vals, uint32_t cycles,
             const uint32_t *read_ids, uint32_t n_read, uint8_t *read_vals, int *any_stopped);
/* Whole-pool access (checkpoint/restore of the architectural state). */
int tgs_read_pool(tgs_ctx *ctx, uint32_t instance, int pool, uint64_t offset, uint64_t n, uint8_t *out);
int tgs_write_pool(tgs_ctx *ctx, uint32_t instance, int pool, uint64_t offset, uint64_t n, const uint8_t *in);
int tgs_pool_size(const tgs_ctx *ctx, int pool, uint64_t *size);

/* Restart one instance from the loaded initial state (Init/Randomize pools,
 * zero exchange pool, cycle 0, stop cleared) with the stimulus generator of
 * instance `index` of seed `seed` (splitmix64(seed + index)). The serving
 * loop reuses a slot for a new job with it; the reference has one global
 * instance and no counterpart (ToucanSim.cu:102-204 reloads everything). */
int tgs_reset_instance(tgs_ctx *ctx, uint32_t instance, uint64_t seed, uint32_t index);

/* SimVCD (unsupport/SimVCD.cpp:193-241) dump source: start an asynchronous
 * snapshot of one instance's register pool. The copy is ordered after the
 * cycles already run; it goes to a device staging buffer on the context
 * stream and to pinned host memory on a side stream, so it overlaps the next
 * eval. Up to 4 snapshots are in flight; a snapshot's bytes stay valid until
 * the fourth snapshot after it is started. */
int tgs_snapshot_begin(tgs_ctx *ctx, uint32_t instance, uint64_t *ticket);
/* Wait for a snapshot; *data points to pinned memory holding the register
 * pool (regPoolSize bytes) followed by the signal-probe values in probe-list
 * order (tgs_set_probes; none when no probes are set); *cycle (optional) is
 * the instance's cycle count when it was taken. */
int tgs_snapshot_wait(tgs_ctx *ctx, uint64_t ticket, const uint8_t **data, uint64_t *cycle);

/* Signal probes: the values of value-pool slots (SimDebugInfo signal
 * locations, (partition, valId) pairs, include/ToucanGPUGenDataTypes.h:232)
 * as each partition leaves them after its exec levels in the last cycle run.
 * The reference's readSignal is a stub (src/ToucanSim.cu:291-300) and its
 * SimVCD reads host copies (unsupport/SimVCD.cpp:193-241); here every kernel
 * copies the listed slots out at the end of the partition's visit.
 * tgs_set_probes replaces the list (n = 0 clears it; a slot outside the
 * partition's value pool is TGS_E_ARG) and zeroes the values;
 * tgs_read_probes returns the n values in list order for one instance. */
int tgs_set_probes(tgs_ctx *ctx, const uint32_t *parts, const uint32_t *slots, uint32_t n);
int tgs_read_probes(tgs_ctx *ctx, uint32_t instance, uint8_t *out, uint32_t n);

/* State checkpoint of one instance (no reference counterpart: the reference
 * keeps one global instance and re-runs init, src/ToucanSim.cu:102-204).
 * The image is a 64-byte header (magic "TGSCKPT1", FNV-1a 64 of the
 * serialized design, pool sizes, stop flag, cycle count, stop cycle,
 * stimulus generator state) followed by the register, memory and exchange
 * pools, nibble per byte. It is the same in every kernel layout, so an
 * instance saved from a batched context restores into a single-instance
 * context of the same design and the other way round. The print log is
 * output, not state, and is not part of it; the stimulus input list is
 * configuration (tgs_set_stimulus, which resets every generator, goes
 * before a restore). Restoring an image of another design fails with
 * TGS_E_ARG; a damaged one with TGS_E_FORMAT. */
int tgs_checkpoint_size(tgs_ctx *ctx, uint64_t *bytes);
int tgs_checkpoint_save(tgs_ctx *ctx, uint32_t instance, uint8_t *buf, uint64_t bytes);
int tgs_checkpoint_restore(tgs_ctx *ctx, uint32_t instance, const uint8_t *buf, uint64_t bytes);

/* get_eval_done (src/SimEval.cu:1083-1087) and get_real_cycles (:1089-1093). */
int tgs_get_stop(tgs_ctx *ctx, uint32_t instance, int *stopped);
int tgs_get_cycles(tgs_ctx *ctx, uint32_t instance, uint64_t *cycles);
/* Cycle at which the stop flag was first raised (UINT64_MAX if never). */
int tgs_get_stop_cycle(tgs_ctx *ctx, uint32_t instance, uint64_t *cycle);

/* Print log of fired print ops: (cycle, message id) pairs in firing order of
 * cycles, every event since the instance was loaded, reset or restored or
 * since the last tgs_clear_print_log. *count gets the number of entries
 * held; up to `cap` are copied. The reference printf()s from the device
 * (src/SimEval.cu:696-704). The device log holds 65,536 events per launch;
 * the runtime moves it to host memory after every launch that may have
 * printed, so a long run loses nothing unless one launch fires more. */
int tgs_get_print_log(tgs_ctx *ctx, uint32_t instance, uint64_t *cycles, uint32_t *msgs, uint64_t cap,
                      uint64_t *count);
/* Drop the instance's held print events (the caller has consumed them).
 * *dropped (may be null) gets the number of events lost since the last
 * clear because a launch fired more than the device log holds. */
int tgs_clear_print_log(tgs_ctx *ctx, uint32_t instance, uint64_t *dropped);

/* printProfile (src/ToucanSim.cu:306-325): launch shape and measured kernel
 * time; barrier_ticks / total_ticks cover every workgroup of every launch. */
int tgs_get_profile(tgs_ctx *ctx, tgs_profile *out);
int tgs_reset_profile(tgs_ctx *ctx);

/* Per-partition profile (ENABLE_SIM_PROFILE of src/SimEval.cu:894-983).
 * When enabled, the single-instance kernels time launch-local cycles
 * start .. start + count - 1 of every launch with the shader clock
 * (s_memtime, the counterpart of clock64): for each partition, the ticks
 * from the start of its region on its workgroup (after the grid barrier or
 * the launch start) to the end of its evaluation ("useful"), and to the end
 * of the region's grid barrier ("total"). The batched kernel is not timed.
 * The reference's defaults are PROFILE_START_CYCLE = PROFILE_COLLECT_CYCLE
 * = 100 (include/SimConfigs.h:3-6). Disabling it costs the kernels nothing. */
int tgs_set_part_profile(tgs_ctx *ctx, int enable, uint32_t start_cycle, uint32_t collect_cycles);
/* copy_profile_data (src/SimEval.cu:1100-1104): per-partition averages over
 * the cycles collected by the last launch that reached the window (the
 * reference's running means restart with every launch); `count` entries
 * (partition order of the design). *collected (may be null) gets the number
 * of cycles averaged, 0 when no launch has reached the window. */
int tgs_get_part_profile(tgs_ctx *ctx, int64_t *useful, int64_t *total, size_t count, uint32_t *collected);

/* Program statistics of the loaded design. */
int tgs_get_program_stats(tgs_ctx *ctx, tgs_program_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* TOUCAN_RT_H */
