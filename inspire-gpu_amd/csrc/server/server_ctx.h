// Internal definition of the opaque GpuServerCtx (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <atomic>
#include <mutex>
#include <utility>
#include <vector>

#include "gpu_protocol.h"
#include "kernels.h"

namespace inspire {

void hip_check(hipError_t e, const char* what);
inline void dfree(void* p) {
    if (p) (void)hipFree(p);
}

struct GpuServerCtx {
    PublicParams pp{};
    GpuServerConfig cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    size_t G = 0, Da = 0, T = 0;

    // resident data
    // encoded DB, byte planes in 16-row blocks: plane p (low/high byte) of
    // (r, c) at byte ((r/16)*2 + p)*C*16 + c*16 + r%16
    uint16_t* d_db = nullptr;
    const uint8_t* db_planes() const { return reinterpret_cast<const uint8_t*>(d_db); }
    uint32_t* d_precomp = nullptr;  // [G][PC_GROUP] arena
    uint32_t* d_tw = nullptr;       // twiddles
    ipx::NttConst nc{};
    ipx::CrtConst cc{};

    // per-query slot pools, max_batch each
    ipx::SlotLayout lay{};
    ipx::DbPlan plan1{};            // database pass of a single query over all columns (autotuned)
    size_t plan1_target = 2048;     // its workgroup target
    size_t mfma_target = 1024;      // workgroups the batched pass aims for
    uint32_t mfma_depth = 4;        // register-ring depth of the 9..32-query passes
    int mfma_form = 0;              // 9..32 queries: form of the pass (dbplan_form)
    // 9..16 queries when timed apart from a larger largest batch (their
    // best form differs): target, depth and form
    bool tuned16 = false;
    size_t mfma16_target = 1024;
    uint32_t mfma16_depth = 4;
    int mfma16_form = 0;
    // row slices of the batched chunk plan, timed at setup for plans with
    // chunk_param digit tiles (0: the narrowing rule)
    uint32_t chunk_slices = 0, chunk_param = 0;
    uint32_t chunk_slices1 = 0;     // the same for one-query answers (plan1's kind)
    // chunks from index late_from on (the collapse beside them has finished)
    // take late_slices row slices instead (0: as the chunks before them);
    // one-query answers use the *1 pair
    uint32_t late_from = 0, late_slices = 0, late_from1 = 0, late_slices1 = 0;
    // The plan of a chunk of ng groups at index k of the pipelined answer.
    ipx::DbPlan chunk_plan_at(size_t nq, size_t k, size_t ng, const ipx::DbPlan& whole) const;
    int collapse_variant = 0;       // 17..32-query collapse pass form (launch_collapse)
    size_t gchunk = 0;              // groups per chunk of the pipelined answer
    // the pipelined answer's chunks [lo, hi), highest groups first: chunks of
    // gchunk groups, the lowest one split into gchunk/2, /4, /4 when
    // tail_split (its Horner steps are the ones left after every pass), the
    // highest one into gchunk/4, /4, /2 when head_split (the chains start
    // after a quarter chunk's passes)
    bool tail_split = true, head_split = false;
    std::vector<std::pair<size_t, size_t>> chunks;
    void set_chunks();
    // pipeline shape of one-query and many-query calls: 0 serial on the
    // caller's stream, 2 Horner chains beside the passes, 3 collapse beside
    // too, 4 the collapse of all groups before the database-pass chunks, 5
    // collapse beside and a chunk ahead; streamed (one full-width database
    // pass handing groups over as it runs, the chains waiting for them in
    // the kernel): 6 after the collapse of all groups, 7 with the collapse
    // beside, 8 with the collapse in the pass's own launch, group by group
    // (up to 8 queries) (2, 3, 6 and 8 timed at setup)
    int pipe_single = 3, pipe_batched = 2;
    int pipe_single_chunked = 3, pipe_batched_chunked = 3;  // where a streamed shape cannot run
    bool pipe_forced = false;       // INSPIRE_PIPELINE set the shapes (not autotuned)
    // row slices of the streamed pass (one query, many), and whether this
    // server may stream at all: the chains of a streamed answer hold their
    // CUs while they wait, so only one server per device streams (at most
    // kLanes x 32 waiting chains, half the CUs at most)
    uint32_t stream_slices1 = 4, stream_slices = 4;
    // eighths of the groups (from the top) in a streamed pass's first
    // launch, the chains starting after it (0: one launch, chains at once)
    uint32_t stream_split1 = 0, stream_split = 0;
    bool stream_ok = false;
    // The streamed pass's plan for nq queries (the tuned kernel, the streamed
    // slice count, at least the slices exactness needs); kind F64 if none fits.
    // One- and two-query passes stream with 128-column tiles per wave (fewer,
    // longer workgroups per group), except inside the collapse's launch
    // (fused: the 64-column tiles its work items share).
    ipx::DbPlan stream_plan(size_t nq, uint32_t slices, bool fused = false) const;
    // Whether an answer on s may take a streamed shape now (not under capture).
    bool can_stream(hipStream_t s) const;
    // A streamed sub-batch of nq queries holds nq*T chain workgroups and
    // ceil(nq/pack_width(nq)) pack workgroups (one CU each) while they wait:
    // at most kStreamMaxWaiters, so 4 lanes leave 64 of 256 CUs to passes.
    static constexpr size_t kStreamMaxWaiters = 48;
    bool stream_fits(size_t nq) const;
    // Queries per streamed pack workgroup: 2 up to 16 queries (one round of
    // its 4 teams per group, so the pack keeps pace with the faster passes
    // of small batches), kPackQueries above (two rounds, fewer CUs held
    // beside the slower 17..32-query pass).
    static uint32_t pack_width(size_t nq) { return nq <= 16 ? 2u : ipx::kPackQueries; }
    // Make this the device's streaming server (the previous one, if any,
    // then answers with chunked shapes). The caller holds this server's mu.
    void take_streaming();
    size_t qlen = 0, resp_u64s = 0;
    size_t part_stride = 0, part2_stride = 0, keys_stride = 0, rgsw_stride = 0, packed_stride = 0;
    uint64_t* d_q = nullptr;        // flat queries (device staging)
    uint64_t* d_resp = nullptr;     // flat responses (device staging)
    double* d_bq = nullptr;
    uint32_t* d_keys = nullptr;
    uint32_t* d_rgsw = nullptr;
    uint32_t* d_part = nullptr;     // DB-pass row-slice partials [B][S][2][C]
    uint32_t* d_interm = nullptr;   // DB-pass result [B][2][C]
    uint32_t* d_part2 = nullptr;    // collapse sums [B][G][2][N]
    // scratch of the batched database pass, one set per concurrent lane
    static constexpr int kLanes = 4;
    int8_t* d_dig[kLanes] = {};     // query digits [R/16][<=256][16]
    uint32_t* d_S[kLanes] = {};     // digit-offset corrections [64]
    // scratch of the collapse pass: key digits [2][2N/16][<=128][16], key sums
    static constexpr size_t kKdigBytes = 2 * (2 * ipx::N / 16) * 128 * 16;
    int8_t* d_kdig[kLanes] = {};
    uint32_t* d_KS[kLanes] = {};
    // per lane, one set per 32-query sub-batch of a call: the streamed
    // pass's per-group completion counts [G], then the streamed pack's
    // per-(workgroup, group) publication flags [32 / kPackQueries][G]
    // (pack_width gives at most that many workgroups: 16 queries at 2)
    static constexpr size_t kStreamSetWords = 1 + ipx::kMfmaMaxQueries / ipx::kPackQueries;
    static_assert(16 / 2 <= ipx::kMfmaMaxQueries / ipx::kPackQueries, "pack workgroups of a 16-query sub-batch");
    uint32_t* d_sflags[kLanes] = {};
    size_t stream_sets = 0;
    // per lane: the streams the collapse and the Horner chains run on beside
    // the database pass, and the events that hand each group chunk over
    static constexpr int kChunkEvents = 16;
    hipStream_t hstream[kLanes] = {};
    hipStream_t cstream[kLanes] = {};
    hipEvent_t chunk_ev[kLanes][kChunkEvents] = {};  // database pass of a chunk done
    hipEvent_t coll_ev[kLanes][kChunkEvents] = {};   // collapse of a chunk done
    hipEvent_t prep_ev[kLanes] = {};                 // keys of a batch in place
    hipEvent_t pass_ev[kLanes] = {};                 // a split streamed pass's first launch done
    hipEvent_t horner_done[kLanes] = {};
    uint32_t* d_packed_b = nullptr;
    // Horner: 0 = one workgroup per chain, 1 = a workgroup pair (one per
    // limb), -1 = pair when a call has at most kPairMaxChains chains and its
    // answer is serial (beside the passes the pair's exchange is slower)
    int horner_mode = -1;
    static constexpr size_t kPairMaxChains = 2;
    // pack kernel of pipelined chunks: 1 one workgroup per query, 0 a grid
    // over (group, limb), -1 one workgroup for one-query calls, 2 fused into
    // the Horner chains' own workgroups (where the one-workgroup Horner runs)
    int pack_wg = 0;
    uint32_t* d_hxbuf = nullptr;    // pair exchange buffers [B*T chains]
    uint32_t* d_hflags = nullptr;   // pair step flags [B*T][2]
    // per scratch lane: nonzero once a pair chain of that lane's calls gave
    // up its bounded wait (read and cleared by check_status)
    uint32_t* d_status = nullptr;   // [kLanes]
    uint32_t horner_spin_limit = 0; // the pair kernel's wait bound in sleeps (0: report at once)
    uint32_t stream_spin_limit = 0; // the same for chains waiting for a streamed pass's groups
    uint32_t* d_hstate = nullptr;   // Horner chain state between chunk launches [B*T][4][N]
    uint64_t* h_q = nullptr;        // pinned host staging
    uint64_t* h_resp = nullptr;

    // packed-wire path (53-bit CRT-packed queries in, compressed responses out)
    static constexpr size_t kCompressedCtBytes = ipx::N * (28 + 20) / 8;  // 12288
    size_t qpack_bytes = 0;
    uint8_t* d_qpack = nullptr;     // packed query staging [B][qpack_bytes]
    uint32_t* d_cmp = nullptr;      // coefficient-form responses [B*T][4][N]
    uint8_t* d_rcmp = nullptr;      // compressed response staging [B*T][12288]

    size_t resident_bytes = 0;
    // mu serialises every entry point that touches the slots, scratch lanes
    // or plans; loop_active (under mu) marks a serving loop that owns them
    // until it stops, and those entry points then return "busy".
    std::mutex mu;
    bool loop_active = false;
    // Calls on lane 0 and slots [0, count) from different streams are
    // ordered: each waits for the previous one's work (order_ev, recorded on
    // order_stream) before enqueueing its own.
    hipEvent_t order_ev = nullptr;
    hipStream_t order_stream = nullptr;
    bool order_stream_set = false;
    void order_before(hipStream_t s);
    void order_after(hipStream_t s);

    // Captured answer steps: with graph_mode on, the first answer of a given
    // (input, output, count, slots, lane, form) is captured from its stream
    // into a hipGraph and every later one replays it. Kernel arguments are
    // frozen at capture, so anything that changes a plan drops the cache.
    struct StepGraph {
        const void* in;
        void* out;
        size_t count, slot0;
        int lane, packed;
        hipGraphExec_t exec;
    };
    static constexpr size_t kMaxGraphs = 64;
    int graph_mode = 0;
    std::vector<StepGraph> graphs;
    std::mutex graph_mu;  // the cache (device-side calls take no other lock)
    // fn enqueues the step on s; captured or replayed as above, run as is when
    // graph_mode is off or s is the null stream or already capturing.
    void run_step(const void* in, void* out, size_t count, size_t slot0, int lane, int packed, hipStream_t s,
                  const std::function<void()>& fn);
    void drop_graphs();

    void alloc_slots();
    // Pick the fastest database-pass variant on this device and geometry.
    void autotune_dbpass();
    // Pipeline stages over slots [slot0, slot0+count) on stream s.
    void stage_prep(const uint64_t* d_in, size_t count, hipStream_t s, size_t slot0);
    // The tuned database-pass plan for nq queries over ncols columns.
    ipx::DbPlan plan_for(size_t nq, size_t ncols) const;
    // The tuned full-width plan narrowed to ncols columns at its workgroup count.
    ipx::DbPlan chunk_plan(size_t nq, size_t ncols) const;
    // Digits of queries [q0, q0+nq) for an MFMA plan (no-op for fp64).
    void dbpass_prep(size_t q0, size_t nq, const ipx::DbPlan& p, hipStream_t s, int lane);
    // Database pass of queries [q0, q0+nq) over the columns of groups [g_lo, g_hi) into interm.
    void dbpass_range(size_t q0, size_t nq, size_t g_lo, size_t g_hi, const ipx::DbPlan& p, hipStream_t s, int lane);
    // force: -1 the tuned plans, 0 the fp64 pass per query, 1 the MFMA pass
    void stage_dbpass(size_t count, hipStream_t s, size_t slot0, int lane = 0, int force = -1);
    // The pair kernel only where the caller reads the lane's status after
    // the call (status_read): its bounded wait reports through it.
    bool use_pair(size_t chains, bool beside_passes, bool status_read) const;
    // Horner chains of `count` queries over groups [g_lo, g_hi).
    // fp: the chains compute their groups' packed b-parts themselves first
    void horner_range(size_t count, uint64_t* d_out, size_t g_lo, size_t g_hi, hipStream_t s, size_t slot0,
                      int lane, bool status_read, bool beside_passes = false, const ipx::FusedPack* fp = nullptr);
    void stage_collapse(size_t count, hipStream_t s, size_t slot0, int lane = 0);
    void stage_finalize(size_t count, hipStream_t s, size_t slot0);
    void stage_horner(size_t count, uint64_t* d_out, hipStream_t s, size_t slot0, bool status_read);
    // Throw (and clear) if a pair chain of a call on `lane` reported a fault
    // through its status word; s is the stream the call ran on.
    void check_status(hipStream_t s, int lane);
    // Zero the lane's status word on s (before a call whose status is read,
    // so an earlier asynchronous call's report is not taken for its own).
    void clear_status(hipStream_t s, int lane);
    std::atomic<size_t> n_streamed{0};  // streamed sub-batches enqueued (a test counter)
    // row slices whose partials the last answer's first sub-batch left in
    // d_part instead of a sum in d_interm (a streamed pass; 0: summed)
    uint32_t last_part_slices = 0;
    // Database-pass timing inside answers (a measurement aid): while
    // time_pass is on, the launches of each sub-batch's database pass (one
    // streamed launch, or its chunks' launches) are bracketed on their
    // stream by events from pass_marks (start, end, start, end, ...).
    static constexpr size_t kMaxPassMarks = 8192;
    bool time_pass = false, time_pass_overflow = false;
    std::vector<hipEvent_t> pass_marks;
    size_t pass_marks_used = 0, passes_timed = 0;
    void mark_pass(hipStream_t s, bool first);
    // Seconds of sleeps: far above any partner skew or any streamed pass's
    // time to reach a group, bounded so a chain whose partner or pass never
    // runs still ends.
    static constexpr uint32_t kHornerSpinLimit = 1u << 26;
    // The streamed chains' bound: ~1 s or more of sleeps, far above a
    // database pass's time even queued behind other batches on the device.
    static constexpr uint32_t kStreamSpinLimit = 1u << 23;
    // Enqueue the answer of `count` flat queries at d_in (device) into d_out
    // (device) on stream s, using slots [slot0, slot0 + count) and scratch
    // set `lane` (work that may overlap on the device uses distinct lanes).
    // status_read: the caller reads the lane's status word after the call
    // (check_status), so the pair Horner kernel may be used.
    void enqueue(const uint64_t* d_in, size_t count, uint64_t* d_out, hipStream_t s, size_t slot0 = 0,
                 int lane = 0, bool status_read = false);
    // Host-staged round trip through the pinned staging of slots
    // [slot0, slot0+count): h_q -> device -> h_resp, then copy to out.
    void answer_host(size_t count, uint64_t* out, hipStream_t s = nullptr, size_t slot0 = 0);
    // enqueue() behind the device wire codecs: count packed queries at d_in
    // (device, qpack_bytes each) -> count*T compressed cts at d_out (device).
    void enqueue_packed(const uint8_t* d_in, size_t count, uint8_t* d_out, hipStream_t s, size_t slot0 = 0,
                        int lane = 0, bool status_read = false);
    // Host round trip of the packed forms through the pinned staging.
    void answer_packed_host(const uint8_t* in, size_t count, uint8_t* out);
};

// Device bytes of a server: the resident DB and precomp with the fixed
// scratch (twiddles, lanes), what each of max_batch answer slots adds, and
// the peak of gpu_preprocess (DB, precomp and its transient buffers).
// resident_bytes of a server with max_batch B is fixed + B * per_slot.
struct ServerFootprint {
    size_t fixed = 0, per_slot = 0, preprocess_peak = 0;
    size_t part_words = 0;  // row-slice partials per slot (u32 words)
};
ServerFootprint server_footprint(const PublicParams& pp, size_t gchunk);
// Groups per chunk of the pipelined answer by default (about 8 chunks).
size_t default_group_chunk(size_t G);

}  // namespace inspire

// C ABI handle (capi.h): one database generation.
struct ipir_server {
    inspire::PublicParams pp;
    inspire::GpuServerCtx* ctx;
};

