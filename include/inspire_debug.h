/* Inspection and stage-level entry points of the server library, used by
 * the parity tests and the benchmark. No counterpart in the reference (its
 * tests link the C++ internals directly); they expose intermediate device
 * tensors through host buffers. All return 0 on success, nonzero on bad
 * arguments or device failure, and synchronise before returning.
 */
#ifndef INSPIRE_DEBUG_H
#define INSPIRE_DEBUG_H
#include <stddef.h>
#include <stdint.h>

#include "capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One group's precomp in the reference's layout (converted back from the
 * resident dot-product form): u32 words [a: 2N][D+: 2*(N/2-1)*2*N]
 * [D-: same][D_final: 2*2*N], each in [limb][..][coeff] order. */
size_t ipir_dbg_precomp_group_words(void);
int ipir_dbg_precomp_group(const ipir_server* srv, size_t group, uint32_t* out);
/* rows [row0, row0+nrows) of the encoded u16 database, returned row-major */
int ipir_dbg_db_rows(const ipir_server* srv, size_t row0, size_t nrows, uint16_t* out);
/* The block of rows [row0, row0+nrows) and columns [col0, col0+ncols) of
 * the encoded u16 database, returned row-major (nrows x ncols). */
int ipir_dbg_db_block(const ipir_server* srv, size_t row0, size_t nrows, size_t col0, size_t ncols,
                      uint16_t* out);
/* Database pass alone over nq (<= max_batch) query vectors b [nq][2][db_rows]
 * (residues mod q0 | q1, u64): interm [nq][2][db_cols]. path -1 = the tuned
 * plans of the answer pipeline, 0 = the fp64 pass per query, 1 = the int8
 * MFMA pass. */
int ipir_dbg_db_pass(ipir_server* srv, const uint64_t* b, size_t nq, int path, uint32_t* interm);
/* Device time of the full-width database pass alone (the pass kernel of
 * the plan answer_batch uses for nq <= 32 queries, and its slice reduction
 * if the plan has row slices, on the server stream) over the first nq
 * query slots as last staged, averaged over iters passes; the query digits
 * are prepared once, outside the timed launches. */
int ipir_dbg_time_dbpass(ipir_server* srv, size_t nq, int iters, double* ms_per_pass);
/* The database-pass kernel (profiler name) and row-slice count the server
 * uses for a batch of nq queries, NUL-terminated into out[cap]. */
int ipir_dbg_dbpass_kernel(const ipir_server* srv, size_t nq, char* out, size_t cap);
/* The setup autotuner's choices as one line of text (database-pass plans
 * at batch 1 and at the largest batch, their form and workgroup target,
 * the pipelined chunks' plans, the collapse variant, the pipeline shapes),
 * NUL-terminated into out[cap]. */
int ipir_dbg_tuning(const ipir_server* srv, char* out, size_t cap);
/* Answer one host flat query and return the intermediate tensors:
 * interm [2][db_cols] (DB pass, coefficient form), packed_b
 * [n_packed][2][N] (after the collapse), resp (flat response). Any output
 * pointer may be NULL. */
int ipir_dbg_answer_taps(ipir_server* srv, const uint64_t* query, uint32_t* interm, uint32_t* packed_b,
                         uint64_t* resp);
/* Standalone kernels on the current device. ntt: npoly RNS polynomials
 * [npoly][2][N] u32, forward (inverse = 0) or inverse, in place on a copy.
 * horner: cts [L][a|b][limb][N] NTT form, rgsw flat (4*d_eff*2*N u64,
 * [part][digit][limb][coeff]); out [a|b][limb][N]. */
int ipir_dbg_ntt(const uint32_t* in, uint32_t* out, size_t npoly, int inverse);
/* Device time of one launch transforming npoly RNS polynomials (both limbs)
 * resident in HBM, averaged over iters launches. */
int ipir_dbg_ntt_bench(size_t npoly, int inverse, int iters, double* ms_per_call);
/* Time a plain device copy of `bytes` (a multiple of 16) between two fresh
 * buffers: the known-size reference that calibrates memory-traffic
 * counters (FETCH_SIZE, WRITE_SIZE) of the other kernels. */
int ipir_dbg_copy_bench(size_t bytes, int iters, double* ms_per_copy);
int ipir_dbg_horner(const uint64_t* cts, size_t L, const uint64_t* rgsw, uint64_t* out);
/* horner with a chosen kernel: variant 0 = one workgroup per chain, 1 = a
 * workgroup pair (one per limb). */
int ipir_dbg_horner_variant(const uint64_t* cts, size_t L, const uint64_t* rgsw, uint64_t* out, int variant);
/* ipir_dbg_horner over DEVICE buffers, allocating nothing: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops). No counterpart in the reference (its gpu_horner_eval,
 * src/gpu_online.cu:474-540, allocates per call). d_cts [L][a|b][limb][N]
 * u64 (NTT form), d_rgsw 4*d_eff*2*N u64, d_out [a|b][limb][N] u64; enqueued
 * on hip_stream (NULL: the null stream), asynchronous. Its scratch is one
 * static array per GPU (8.3 MiB, held until the process ends), set up at
 * the first call on a device; calls that overlap on one GPU share it and
 * race, so callers order them. 0 ok; 1 null pointer or L == 0; 2 L above
 * 256 (the longest chain of any geometry, the array's limit), nothing
 * launched; 4 launch failure. Outputs equal ipir_dbg_horner's. */
int ipir_horner_noalloc(const uint64_t* d_cts, size_t L, const uint64_t* d_rgsw, uint64_t* d_out,
                        void* hip_stream);
/* Set the groups per chunk of the pipelined answer (a divisor of n_packed;
 * 1 if the partial-sum buffer cannot hold the resulting plans). */
int ipir_dbg_set_group_chunk(ipir_server* srv, size_t groups);
/* Split chunks of the pipelined answer where the chunk size divides by 4.
 * Bit 0 of `on` splits the lowest chunk into 1/2, 1/4 and 1/4 of its
 * groups, so fewer Horner steps are left after the last pass. Bit 1 splits
 * the highest chunk into 1/4, 1/4 and 1/2, so the Horner chains start after
 * a quarter chunk's passes. The default is set at setup (INSPIRE_TAIL_SPLIT
 * overrides it). 1 for values outside 0..3. */
int ipir_dbg_set_tail_split(ipir_server* srv, int on);
/* Select the Horner kernel a server's answers use: mode as variant above,
 * or -1 (the default) to use the pair only for serial answers with at most
 * two chains. Whatever the mode, the asynchronous device entry points
 * (ipir_answer_batch_device, ipir_answer_batch_packed_device) never use the
 * pair: its bounded wait reports through a status word that only the
 * blocking entry points and the serving loop read. */
int ipir_dbg_set_horner(ipir_server* srv, int mode);
/* Bound of the pair kernel's partner wait, in sleeps (default 2^26, ~2 s).
 * 0 makes every pair chain report the bound exceeded at its first
 * exchange and stop (all kernels still end normally): the call that ran it
 * then returns 4, and later calls are unaffected. 5 while a loop runs. */
int ipir_dbg_set_horner_spin_limit(ipir_server* srv, uint32_t limit);
/* The same bound for Horner chains waiting for a streamed database pass's
 * groups (default 2^23; 0: give up at once and report, as above; the pass
 * itself runs to its end). */
int ipir_dbg_set_stream_spin_limit(ipir_server* srv, uint32_t limit);
/* Keep a serving loop's partitions busy after their device work finished
 * (on = 1) until released (on = 0): requests admitted meanwhile go to the
 * other partitions or wait. For tests of admission while batches are in
 * flight. ipir_loop_stop releases a hold. */
int ipir_dbg_loop_hold(ipir_loop* loop, int on);
/* Select the pack kernel of pipelined answers' chunks: 1 one 1024-lane
 * workgroup per query, 0 (the default, or INSPIRE_PACK_WG at setup) a grid
 * of one 256-lane workgroup per (group, limb, query), -1 one workgroup for
 * one-query calls, 2 no kernel of its own: each one-workgroup Horner chain
 * packs its own groups before its steps (the two-workgroup chain keeps a
 * pack kernel). Outputs are equal. 1 for other values. */
int ipir_dbg_set_pack(ipir_server* srv, int mode);
/* Answer through captured hipGraphs (1) or launch by launch (0, the default
 * unless INSPIRE_GRAPH=1 at setup). With 1, the first answer of each
 * (input buffer, output buffer, count) on a stream that is not the null
 * stream is captured (its kernels, cross-stream events and waits) and later
 * answers of the same key replay it; a call made while the caller's stream
 * is capturing is recorded into the caller's graph instead. The
 * ipir_dbg_set_* calls above drop the captured graphs. No counterpart in
 * the reference (whose answer launches kernel by kernel). 1 for other
 * values. */
int ipir_dbg_set_graph(ipir_server* srv, int on);
/* Set the register-ring depth of the batched database passes a server uses
 * and their form: 0 the waves of a 4-wave block share the digit tile (depth
 * 3..5), 1 (17..32 queries) they share the DB tile (3..4), 2 the waves of an
 * 8-wave block share the digit tile (3..7), 3 and 4 (17..32 queries) the
 * 8- and 4-wave forms with 7 digit rows per query (3..7, 3..5), 5 (17..32
 * queries) those digits accumulated by power of 256 (3..5, also for 9..16
 * queries); other 9..16-query plans cap the depth at 4 and use form 0 for
 * 1 and 4 and 2 for 3. The setting applies to every 9..32-query batch
 * (also where setup timed 9..16 queries apart). 1 for other values. */
int ipir_dbg_set_mfma_depth(ipir_server* srv, uint32_t depth, int form);
/* Row slices of the pipelined answer's database-pass chunks for batches
 * planned like an nq-query batch, or for one-query answers when nq is 1
 * (0: back to the narrowing rule; a power of two up to 64, used when the
 * plan reaches it and the partial-sum buffer holds it). */
int ipir_dbg_set_chunk_slices(ipir_server* srv, size_t nq, uint32_t slices);
/* Row slices of the pipelined answer's database-pass chunks from chunk
 * index from_chunk on (chunks go from the highest groups down), for batches
 * planned like an nq-query batch or for one-query answers when nq is 1:
 * the chunks that run after the collapse beside them has finished can use
 * the whole device. 0 slices: as the earlier chunks (the default unless
 * the setup autotuner picked otherwise). 1 for bad values. */
int ipir_dbg_set_late_slices(ipir_server* srv, size_t nq, size_t from_chunk, uint32_t slices);
/* Form of the 17..32-query collapse pass (0..5; 1 otherwise), and the one
 * in use (the setup-time autotuner's pick unless set). */
int ipir_dbg_set_collapse_variant(ipir_server* srv, int variant);
int ipir_dbg_get_collapse_variant(const ipir_server* srv);
/* Set the pipeline shape of one-query and many-query answers: 0 serial, 2
 * Horner chains on a second stream per group chunk, 3 the collapse on a
 * third stream as well, 4 the collapse of all groups first and then the
 * database-pass chunks with the Horner chains beside them, 5 as 3 with the
 * collapse stream a chunk ahead of the database pass; streamed: one
 * full-width database pass that hands each group over as its columns are
 * final, the Horner chains waiting for the groups inside the kernel (a
 * bounded wait, reported like the pair kernel's), 6 after the collapse of
 * all groups, 7 with the collapse beside it, 8 with the collapse inside the
 * pass's launch, each group's collapse and pass work items together (up to
 * 8 queries; 6 for more). The setup autotuner picks 2, 3, 6 or 8 (4, 5
 * and 7 are tested alternatives that never won at 16 GB).
 * Answers under stream capture, and every answer of a server that is not
 * the device's first live one, take the chunked shape the autotuner picked
 * instead of a streamed one. 1 for other values. */
int ipir_dbg_set_pipeline(ipir_server* srv, int single, int batched);
/* Row slices (a power of two <= 16) of the streamed pass of one-query
 * (nq_class 1) or many-query (other values) answers; raised to what the
 * exactness bounds need, and ignored (chunked shape) where the partials
 * would not fit their buffer. */
int ipir_dbg_set_stream_slices(ipir_server* srv, size_t nq_class, uint32_t slices);
/* Split the streamed pass of one-query (nq_class 1) or many-query answers
 * into two launches: the top `eighths` eighths of the groups first (0..7;
 * 0: one launch), the Horner chains starting only after that launch (they
 * hold their CUs from then on). */
int ipir_dbg_set_stream_split(ipir_server* srv, size_t nq_class, uint32_t eighths);
/* 1 if this server may take the streamed shapes (the device's first live
 * server), else 0. */
int ipir_dbg_stream_ok(const ipir_server* srv);
/* Make srv the device's streaming server; the one that was (if any) then
 * answers with its chunked shapes. Its own mutex is taken: do not call
 * this while that server is inside another call on another thread that
 * itself waits for srv. */
int ipir_dbg_take_streaming(ipir_server* srv);
/* Timings (ms, mean of 5 after a warm-up) of the streamed pieces for nq
 * zero queries at `slices` row slices: out[0] the streamed database pass
 * alone, out[1] the same plan on its 2-D grid with the slice reduction,
 * out[2] the streamed pack alone with every group handed over, out[3] the
 * streamed Horner chains alone with every group published, out[4] the
 * plain chains over every group (5 doubles). */
int ipir_dbg_stream_parts(ipir_server* srv, size_t nq, uint32_t slices, double* out_ms);
/* Streamed 32-query sub-batches this server has enqueued (a test counter). */
size_t ipir_dbg_streamed_count(const ipir_server* srv);
/* Database-pass time inside answers: on = 1 starts recording HIP events
 * around every sub-batch's database-pass launches (the streamed launch, or
 * each chunk's pass) on the stream they run on; on = 0 stops, waits for
 * them and returns the summed pass time and the number of sub-batch passes
 * (each over all columns). 1 if more than 4096 launches were marked (the
 * rest went untimed). */
int ipir_dbg_pass_timing(ipir_server* srv, int on, double* total_ms, size_t* passes);
/* Per-stage device time (ms) of the last ipir_dbg_profile_batch run:
 * [unmarshal+key NTT, DB pass, collapse pass, pack finalize, Horner]. */
int ipir_dbg_profile_batch(ipir_server* srv, const uint64_t* d_queries, size_t count, uint64_t* d_out,
                           double* stage_ms);

/* Device memory helpers for callers without their own HIP binding (tests,
 * benchmark): allocate, copy (kind 1 host->device, 2 device->host,
 * 3 device->device), free, and synchronise the device. */
void* ipir_dev_alloc(size_t bytes);
int ipir_dev_free(void* p);
int ipir_dev_copy(void* dst, const void* src, size_t bytes, int kind);
int ipir_dev_sync(void);
/* A HIP stream (default priority) for the asynchronous entry points, its
 * synchronisation and destruction; free/total device memory now. */
void* ipir_dev_stream_create(void);
int ipir_dev_stream_sync(void* stream);
int ipir_dev_stream_destroy(void* stream);
int ipir_dev_mem_info(size_t* free_bytes, size_t* total_bytes);

#ifdef __cplusplus
}
#endif
#endif /* INSPIRE_DEBUG_H */
