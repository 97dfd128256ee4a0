/* C ABI of the InsPIRe PIR library (server half: HIP, one MI355X; client
 * half: CPU only).
 *
 * Drop-in for the reference's src/capi.h: every declaration of that header
 * (src/capi.h:27-182) is reproduced with the same name, signature, layout
 * and return codes, so an FFI consumer of the reference links against this
 * library unchanged. Flat layouts:
 *   query    : [lwe_b_limb0 : db_rows][lwe_b_limb1 : db_rows]
 *              [ksk5_b : blk][kskneg1_b : blk][rgsw : 4*blk], blk = d_eff*2*ring_n,
 *              ksk [digit][limb][coeff] coefficient form,
 *              rgsw [top_a|top_b|bot_a|bot_b][digit][limb][coeff] NTT form
 *   response : [query][ct][a|b][limb][coeff], NTT form
 * All values are residues in u64 words. Non-canonical residues (>= q) are
 * reduced mod q by the server.
 *
 * Thread-safety: one ipir_server is externally serialized for
 * ipir_answer_batch / ipir_answer_batch_device; the serving loop
 * (ipir_loop_*) is the thread-safe, concurrent front end. While a loop runs
 * on a server, the server's other answer entry points return 5 (busy)
 * instead of blocking. Asynchronous calls on different HIP streams are
 * ordered by the library: each call's work waits for the previous call's.
 */
#ifndef INSPIRE_CAPI_H
#define INSPIRE_CAPI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ipir_server ipir_server;

typedef struct {
    size_t n_entries;
    size_t entry_bytes;
    size_t db_rows;
    size_t db_cols;
    size_t n_packed;
    size_t interp_d;
    size_t num_cts;
    size_t ring_n;
    size_t d_eff;
    uint8_t seed[64];
} ipir_params;

typedef struct {
    size_t max_batch;
    size_t resident_bytes;
    size_t device_free_bytes;
    size_t num_cts;
} ipir_caps;

typedef struct {
    const uint64_t* lwe_b_limb0;
    const uint64_t* lwe_b_limb1;
    const uint64_t* ksk5_b;
    const uint64_t* kskneg1_b;
    const uint64_t* rgsw;
} ipir_query_view;

/* ---- server half (replaces src/capi.h:76-91) ---- */
/* NULL on error (bad geometry, null db/out_params, max_batch == 0, device
 * failure). db_rows_or_zero == 0 picks the default geometry (32768 rows). */
ipir_server* ipir_server_create(size_t n_entries, size_t entry_bytes, size_t db_rows_or_zero,
                                size_t max_batch, const uint16_t* db,
                                const uint8_t* crs_seed_or_null, ipir_params* out_params);
void ipir_server_caps(const ipir_server* srv, ipir_caps* out);
/* 0 ok; 1 null argument; 2 count > max_batch; 3 null view field;
 * 4 device failure; 5 busy (a serving loop owns the server; the reference
 * has no loop and no such code). count == 0 returns 0 and writes nothing. */
int ipir_answer_batch(ipir_server* srv, const ipir_query_view* queries, size_t count,
                      uint64_t* out_resp);
void ipir_server_destroy(ipir_server* srv);

/* ---- client half (replaces src/capi.h:101-182; CPU only) ---- */
typedef struct ipir_client_query ipir_client_query;
size_t ipir_query_u64s(const ipir_params* params);
size_t ipir_response_u64s(const ipir_params* params);
size_t ipir_entry_slots(const ipir_params* params);
int ipir_query_view_from_flat(const ipir_params* params, const uint64_t* flat,
                              ipir_query_view* out_view);
ipir_client_query* ipir_query_build(const ipir_params* params, uint64_t idx, uint64_t* out_query);
int ipir_extract(const ipir_params* params, const ipir_client_query* q, const uint64_t* resp,
                 uint16_t* out_slots);
void ipir_client_query_destroy(ipir_client_query* q);
size_t ipir_query_packed_bytes(const ipir_params* params);
size_t ipir_response_packed_bytes(const ipir_params* params);
int ipir_query_pack(const ipir_params* params, const uint64_t* flat, uint8_t* out);
int ipir_response_pack(const ipir_params* params, const uint64_t* flat, uint8_t* out);
int ipir_query_unpack(const ipir_params* params, const uint8_t* in, uint64_t* out_flat);
int ipir_response_unpack(const ipir_params* params, const uint8_t* in, uint64_t* out_flat);
size_t ipir_response_compressed_bytes(const ipir_params* params);
int ipir_response_compress(const ipir_params* params, const uint64_t* flat, uint8_t* out);
int ipir_extract_compressed(const ipir_params* params, const ipir_client_query* q,
                            const uint8_t* in, uint16_t* out_slots);

/* ---- additions without a counterpart in the reference ---- */

/* The geometry setup() computes for a database (params with a zero seed),
 * for callers that lay out the row-major db_rows x db_cols u16 database
 * before ipir_server_create. 0 on success, 1 if setup() rejects it. */
int ipir_geometry(size_t n_entries, size_t entry_bytes, size_t db_rows_or_zero, ipir_params* out);

/* Client query under a caller-seeded generator (reproducible tests and
 * benchmarks); otherwise identical to ipir_query_build. */
ipir_client_query* ipir_query_build_seeded(const ipir_params* params, uint64_t idx,
                                           uint64_t rng_seed, uint64_t* out_query);

/* Answer `count` queries whose flat layouts sit back to back in DEVICE
 * memory (count * ipir_query_u64s u64s), writing the flat responses to
 * DEVICE memory, on the given HIP stream (NULL = the server's own stream).
 * Asynchronous: returns after enqueueing. Return codes as
 * ipir_answer_batch (4 only for launch failures: the answer's kernels
 * report nothing afterwards). The work waits for the server's previous
 * call if that one was on another stream; inside a caller's stream capture
 * the caller orders the captured work itself. For callers that keep
 * queries resident in HBM. Answers of the device's streaming server may
 * run their Horner chains beside a database pass they wait for inside the
 * kernel; such a wait is bounded (about a second or more), and a chain
 * that gives up writes all-ones words as its response (no residue has
 * that value; it can happen only if the pass gets no CU for that long)
 * and is reported by ipir_device_status. */
int ipir_answer_batch_device(ipir_server* srv, const uint64_t* d_queries, size_t count,
                             uint64_t* d_out_resp, void* hip_stream);

/* Answer `count` queries given in the packed wire form (ipir_query_pack:
 * count * ipir_query_packed_bytes bytes back to back) and return the
 * responses compressed (ipir_response_compress: count * num_cts * 12288
 * bytes, query-major), with the unpacking and the compression run on the
 * device: the host-device copies carry the packed sizes (the reference
 * client's codecs, src/capi_client.cpp:246-385, run on the host only).
 * Blocking; return codes as ipir_answer_batch. */
int ipir_answer_batch_packed(ipir_server* srv, const uint8_t* packed_queries, size_t count,
                             uint8_t* out_compressed);
/* The same over DEVICE buffers, asynchronous on the given HIP stream (NULL
 * = the server's own stream), like ipir_answer_batch_device (a chain that
 * gives up its bounded wait leaves its compressed response undefined). */
int ipir_answer_batch_packed_device(ipir_server* srv, const uint8_t* d_packed_queries, size_t count,
                                    uint8_t* d_out_compressed, void* hip_stream);

/* For callers of the asynchronous entry points (no counterpart in the
 * reference): waits for the server's last call and returns 4 if a Horner
 * chain of an asynchronous call gave up a bounded wait since the last
 * blocking call or status query (those responses are invalid), else 0;
 * then clears the report. 5 while a serving loop owns the server. */
int ipir_device_status(ipir_server* srv);

/* Device memory of a server of this geometry (setup()'s rules; 1 if it
 * rejects the geometry): resident_fixed (DB, precomp, fixed scratch) plus
 * per_slot for each of max_batch answer slots is what ipir_server_caps
 * reports as resident_bytes; preprocess_peak is the most gpu_preprocess
 * holds at once (DB, precomp and its transient buffers). For sizing
 * max_batch against free device memory before ipir_server_create, as the
 * reference's caps are meant for (src/gpu_protocol.h:26-33). */
typedef struct {
    size_t resident_fixed;
    size_t per_slot;
    size_t preprocess_peak;
} ipir_footprint;
int ipir_server_footprint(size_t n_entries, size_t entry_bytes, size_t db_rows_or_zero, ipir_footprint* out);
/* The largest max_batch (at most cap) whose server fits the current
 * device's free memory with reserve_bytes left over, both while
 * preprocessing and once set up; 0 if none fits or the geometry is
 * rejected. */
size_t ipir_fit_max_batch(size_t n_entries, size_t entry_bytes, size_t db_rows_or_zero, size_t reserve_bytes,
                          size_t cap);

/* Continuous-batching serving loop over one server. Requests are admitted
 * into free slots as they arrive, batches of different requests overlap on
 * the device, and each request completes independently.
 *   max_inflight: concurrent device batches (1..4), each taking
 *   max_batch / max_inflight slots; 0 sizes them from the server's slots:
 *   partitions of 32 (one full batched database pass; all slots if fewer),
 *   as many as max_batch holds, up to 4.
 * One loop per server: NULL if one runs already (or on bad arguments).
 * While it runs, the server's other answer entry points return 5.
 * ipir_loop_submit copies the query, returns a ticket (> 0) or 0 on error.
 * ipir_loop_wait blocks until the ticket's response is written to the
 * out_resp given at submit; returns 0 and the request latency (submit to
 * completion) in *latency_us, 4 on a device failure of its batch.
 * ipir_loop_stop lets admitted work finish, then returns the server to
 * its other entry points. */
typedef struct ipir_loop ipir_loop;
ipir_loop* ipir_loop_start(ipir_server* srv, size_t max_inflight);
uint64_t ipir_loop_submit(ipir_loop* loop, const ipir_query_view* q, uint64_t* out_resp);
int ipir_loop_wait(ipir_loop* loop, uint64_t ticket, double* latency_us);
void ipir_loop_stop(ipir_loop* loop);
/* Counters of a running loop (a snapshot; 0 ok, 1 null argument). */
typedef struct {
    uint64_t partitions, slots_per_partition;
    uint64_t submitted, completed, failed;
    uint64_t batches;              /* device batches launched */
    uint64_t max_batch_size;       /* most requests admitted into one batch */
    uint64_t max_busy_partitions;  /* most partitions holding admitted requests at once */
    uint64_t admitted_while_busy;  /* batches admitted while another partition held requests */
    uint64_t device_overlaps;      /* batches whose device work (first to last command, by GPU timestamps)
                                      overlapped another partition's batch */
    uint64_t pending;              /* requests waiting for a free partition now */
    uint64_t busy_partitions;      /* partitions holding admitted requests now */
} ipir_loop_stats;
int ipir_loop_get_stats(ipir_loop* loop, ipir_loop_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* INSPIRE_CAPI_H */
