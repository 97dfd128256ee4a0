// Server orchestration on one MI355X: preprocess, context setup, the
// batched answer pipeline and capability reporting.
//
// Lifecycle and ownership follow the reference (src/gpu_protocol.cu:535-1584):
// gpu_preprocess returns device buffers that gpu_setup_server adopts; every
// per-query buffer is allocated at setup for max_batch queries, so the
// request path never allocates. Differences by design: the precomp of all
// groups lives in one arena (one allocation, fixed group stride), the
// answer pipeline is a fixed sequence of batched kernels on one stream (no
// per-group launches), and errors surface as exceptions instead of exit().
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpu_protocol.h"
#include "kernels.h"
#include "server_ctx.h"
#include "trace.h"

using namespace ipx;

namespace inspire {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPX(x) hip_check((x), #x)

namespace {

template <class T> T* dalloc(size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    hip_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    return static_cast<T*>(p);
}

// Twiddles + Shoup companions for both limbs in the device layout of
// NttConst, built from the host NTT tables (the same tables the client uses).
uint32_t* upload_tables(NttConst& nc, CrtConst& cc) {
    std::vector<uint32_t> h(2 * 4 * (size_t)ipx::N);
    for (int l = 0; l < 2; l++) {
        const NttTables& t = ntt_tables(l);
        for (size_t i = 0; i < ipx::N; i++) {
            h[(l * 4 + 0) * ipx::N + i] = (uint32_t)t.fwd_twiddles[i];
            h[(l * 4 + 1) * ipx::N + i] = (uint32_t)t.fwd_primes[i];
            h[(l * 4 + 2) * ipx::N + i] = (uint32_t)t.inv_twiddles[i];
            h[(l * 4 + 3) * ipx::N + i] = (uint32_t)t.inv_primes[i];
        }
        nc.inv_n[l] = (uint32_t)t.inv_n;
        nc.inv_n_p[l] = (uint32_t)t.inv_n_prime;
    }
    uint32_t* d = dalloc<uint32_t>(h.size());
    hip_check(hipMemcpy(d, h.data(), h.size() * 4, hipMemcpyHostToDevice), "upload twiddles");
    nc.tw = d;
    cc.q1inv_q0 = (uint32_t)mod_inv(Q1 % Q0, Q0);
    cc.q0inv_q1 = (uint32_t)mod_inv(Q0 % Q1, Q1);
    return d;
}

void poly_to_u32(const RnsPoly& p, uint32_t* dst) {
    for (size_t i = 0; i < N; i++) { dst[i] = (uint32_t)p.limbs[0][i]; dst[N + i] = (uint32_t)p.limbs[1][i]; }
}

}  // namespace

// ---------------------------------------------------------------- preprocess
PreprocessData gpu_preprocess(const PublicParams& pp, const uint16_t* db) {
    if (!db) throw std::invalid_argument("gpu_preprocess: null db");
    const size_t R = pp.db_rows, C = pp.db_cols, G = pp.n_packed;
    if (R == 0 || R % N || C == 0 || C % N || G != C / N)
        throw std::invalid_argument("gpu_preprocess: inconsistent geometry");
    const size_t Da = pp.D < G ? pp.D : G;
    const size_t nr = R / N;
    hipStream_t s = nullptr;
    NttConst nc{};
    CrtConst cc{};
    uint32_t* d_tw = nullptr;
    uint16_t* d_db = nullptr;
    uint16_t* d_stage = nullptr;
    uint32_t* arena = nullptr;
    uint32_t *d_crs = nullptr, *d_k5 = nullptr, *d_kn = nullptr, *d_a32 = nullptr, *d_states = nullptr;
    uint32_t* d_dtmp = nullptr;
    auto cleanup = [&](bool all) {
        dfree(d_tw); dfree(d_crs); dfree(d_k5); dfree(d_kn); dfree(d_a32); dfree(d_states); dfree(d_stage);
        dfree(d_dtmp);
        d_dtmp = nullptr;
        d_stage = nullptr;
        if (all) { dfree(d_db); dfree(arena); }
        if (s) (void)hipStreamDestroy(s);
    };
    TraceRange tr_all("gpu_preprocess %zu x %zu, %zu groups", R, C, G);
    try {
        HIPX(hipStreamCreate(&s));
        trace_name_stream(s, "preprocess");
        d_tw = upload_tables(nc, cc);
        // The host DB is row-major. It goes up in slabs of whole rows
        // (~256 MiB): copy, encode in place (row-local), then scatter into
        // the resident layout (byte planes in 16-row blocks).
        d_db = dalloc<uint16_t>(R * C);
        size_t slab = ((size_t)1 << 27) / C / 16 * 16;
        if (slab < 16) slab = 16;
        if (slab > R) slab = R;
        d_stage = dalloc<uint16_t>(slab * C);
        TraceRange tr_up("preprocess: upload, encode, byte planes");
        for (size_t r0 = 0; r0 < R; r0 += slab) {
            const size_t nrows = R - r0 < slab ? R - r0 : slab;
            HIPX(hipMemcpyAsync(d_stage, db + r0 * C, nrows * C * sizeof(uint16_t), hipMemcpyHostToDevice, s));
            HIPX(launch_encode(d_stage, nrows, C, Da, s));
            HIPX(launch_to_planes(d_stage, reinterpret_cast<uint8_t*>(d_db), C, r0 / 16, nrows / 16, s));
        }
        HIPX(hipStreamSynchronize(s));
        dfree(d_stage);
        d_stage = nullptr;

        // CRS rows A_r = NTT(tau_{2N-1}(expand_seed("rlwe", r)))  (src/gpu_protocol.cu:567-583)
        std::vector<uint32_t> h_crs(nr * 2 * N);
        for (size_t r = 0; r < nr; r++) {
            RnsPoly a = expand_seed(pp.seed.data(), "rlwe", r);
            RnsPoly rev = automorphism(a, 2 * N - 1);
            rev.to_ntt();
            poly_to_u32(rev, &h_crs[r * 2 * N]);
        }
        d_crs = dalloc<uint32_t>(h_crs.size());
        HIPX(hipMemcpy(d_crs, h_crs.data(), h_crs.size() * 4, hipMemcpyHostToDevice));
        // Server-side key a-parts, NTT form, [digit][limb][N]  (src/gpu_protocol.cu:405-456)
        std::vector<uint32_t> h_k5(D_EFF * 2 * N), h_kn(D_EFF * 2 * N);
        for (int j = 0; j < D_EFF; j++) {
            RnsPoly a5 = expand_seed(pp.seed.data(), "ksk", (uint64_t)AUTO_GEN * 100 + j + 1);
            RnsPoly an = expand_seed(pp.seed.data(), "ksk", (uint64_t)(2 * N - 1) * 100 + j + 1);
            a5.to_ntt();
            an.to_ntt();
            poly_to_u32(a5, &h_k5[j * 2 * N]);
            poly_to_u32(an, &h_kn[j * 2 * N]);
        }
        d_k5 = dalloc<uint32_t>(h_k5.size());
        d_kn = dalloc<uint32_t>(h_kn.size());
        HIPX(hipMemcpy(d_k5, h_k5.data(), h_k5.size() * 4, hipMemcpyHostToDevice));
        HIPX(hipMemcpy(d_kn, h_kn.data(), h_kn.size() * 4, hipMemcpyHostToDevice));

        arena = dalloc<uint32_t>(G * PC_GROUP);
        // Groups go in chunks of up to 128 (two chains per group fill the
        // CUs), bounded by the transient ring-embed table (32 MiB per group)
        // and the chains' D+/D- in the reference order (64 MiB per group),
        // which the relayout then turns into the resident dot-product form.
        const size_t chunk = G < 128 ? G : 128;
        d_a32 = dalloc<uint32_t>(chunk * 2 * (size_t)N * N);
        d_states = dalloc<uint32_t>(chunk * 4 * (size_t)N);
        d_dtmp = dalloc<uint32_t>(chunk * 2 * RD_SZ);
        for (size_t g0 = 0; g0 < G; g0 += chunk) {
            const size_t ng = G - g0 < chunk ? G - g0 : chunk;
            TraceRange tr_g("preprocess: groups %zu..%zu (a-side, ring embed, collapse chains, relayout)", g0,
                            g0 + ng - 1);
            uint32_t* pc = arena + g0 * PC_GROUP;
            HIPX(launch_aside(reinterpret_cast<const uint8_t*>(d_db), R, C, g0, ng, d_crs, d_a32, nc, s));
            HIPX(launch_ring_embed(d_a32, ng, nc, s));
            HIPX(launch_collapse_chain(d_a32, ng, d_k5, d_dtmp, d_states, nc, cc, s));
            HIPX(launch_collapse_final(d_states, ng, d_kn, pc, nc, cc, s));
            HIPX(launch_dd_relayout(d_dtmp, d_dtmp + RD_SZ, 2 * RD_SZ, ng, pc, s));
        }
        HIPX(hipStreamSynchronize(s));
        cleanup(false);
    } catch (...) {
        cleanup(true);
        throw;
    }
    PreprocessData pd;
    pd.d_db_col = d_db;
    pd.d_precomp_arena = arena;
    pd.d_precomp_a.resize(G);
    pd.d_precomp_D_plus.resize(G);
    pd.d_precomp_D_minus.resize(G);
    pd.d_precomp_D_final.resize(G);
    for (size_t g = 0; g < G; g++) {
        uint32_t* b = arena + g * PC_GROUP;
        pd.d_precomp_a[g] = b + PC_A;
        pd.d_precomp_D_plus[g] = b + PC_DD;
        pd.d_precomp_D_minus[g] = nullptr;
        pd.d_precomp_D_final[g] = b + PC_DF;
    }
    return pd;
}

// ---------------------------------------------------------------- setup
GpuServerCtx* gpu_setup_server(const PublicParams& pp, const PreprocessData& precomp) {
    return gpu_setup_server(pp, precomp, GpuServerConfig{});
}

GpuServerCtx* gpu_setup_server(const PublicParams& pp, const PreprocessData& precomp, const GpuServerConfig& cfg) {
    if (cfg.max_batch == 0) throw std::invalid_argument("gpu_setup_server: max_batch must be >= 1");
    const size_t G = pp.n_packed;
    if (!precomp.d_db_col || precomp.d_precomp_a.size() != G || precomp.d_precomp_D_plus.size() != G ||
        precomp.d_precomp_D_minus.size() != G || precomp.d_precomp_D_final.size() != G)
        throw std::invalid_argument("gpu_setup_server: precomp does not match the geometry");
    const bool per_group = !precomp.d_precomp_arena;
    if (per_group)
        for (size_t g = 0; g < G; g++)
            if (!precomp.d_precomp_a[g] || !precomp.d_precomp_D_plus[g] || !precomp.d_precomp_D_minus[g] ||
                !precomp.d_precomp_D_final[g])
                throw std::invalid_argument("gpu_setup_server: missing per-group precomp buffer");
    auto* ctx = new GpuServerCtx;
    ctx->pp = pp;
    ctx->cfg = cfg;
    ctx->G = G;
    ctx->Da = pp.D < G ? pp.D : G;
    ctx->T = pp.num_cts;
    // The caller's buffers stay the caller's until everything below has
    // succeeded: on failure the context frees only what it allocated.
    ctx->d_db = precomp.d_db_col;
    try {
        HIPX(hipGetDevice(&ctx->device));
        HIPX(hipStreamCreate(&ctx->stream));
        ctx->d_tw = upload_tables(ctx->nc, ctx->cc);
        if (!per_group) {
            ctx->d_precomp = static_cast<uint32_t*>(precomp.d_precomp_arena);
        } else {
            // per-group buffers in the reference layout: gather a / D_final
            // into one arena and convert D+/D- into the dot-product form
            ctx->d_precomp = dalloc<uint32_t>(G * PC_GROUP);
            for (size_t g = 0; g < G; g++) {
                uint32_t* b = ctx->d_precomp + g * PC_GROUP;
                HIPX(hipMemcpyAsync(b + PC_A, precomp.d_precomp_a[g], 2 * N * 4, hipMemcpyDeviceToDevice,
                                    ctx->stream));
                HIPX(hipMemcpyAsync(b + PC_DF, precomp.d_precomp_D_final[g], 2 * D_EFF * N * 4,
                                    hipMemcpyDeviceToDevice, ctx->stream));
                HIPX(launch_dd_relayout(precomp.d_precomp_D_plus[g], precomp.d_precomp_D_minus[g], 0, 1, b,
                                        ctx->stream));
            }
            HIPX(hipStreamSynchronize(ctx->stream));
        }
        ctx->alloc_slots();
    } catch (...) {
        ctx->d_db = nullptr;                        // still the caller's
        if (!per_group) ctx->d_precomp = nullptr;  // likewise
        gpu_free_server(ctx);
        throw;
    }
    // ownership transfer (src/gpu_protocol.cu:856-885); converted per-group
    // buffers are released now that the arena holds them
    if (per_group)
        for (size_t g = 0; g < G; g++) {
            dfree(precomp.d_precomp_a[g]); dfree(precomp.d_precomp_D_plus[g]);
            dfree(precomp.d_precomp_D_minus[g]); dfree(precomp.d_precomp_D_final[g]);
        }
    auto& pd = const_cast<PreprocessData&>(precomp);
    pd.d_db_col = nullptr;
    pd.d_precomp_arena = nullptr;
    pd.d_precomp_a.clear(); pd.d_precomp_D_plus.clear(); pd.d_precomp_D_minus.clear(); pd.d_precomp_D_final.clear();
    return ctx;
}

namespace {
// The server of each device that may take the streamed shapes (at most one).
std::mutex g_stream_mu;
constexpr size_t kMaxDevices = 64;
GpuServerCtx* g_streaming[kMaxDevices] = {};

// Row slices of a streamed pass over C columns: up to 16, with partials of
// at most 32 MiB per query.
constexpr uint32_t kStreamSlicesMax = 16;
constexpr size_t kStreamPartWords = (size_t)1 << 23;
uint32_t stream_slices_cap(size_t C) {
    uint32_t s = 1;
    while (s < kStreamSlicesMax && (size_t)s * 2 * 2 * C <= kStreamPartWords) s *= 2;
    return s;
}

// Single-query database-pass candidates: (MFMA?, fp64 unroll, target workgroups).
struct DbCand {
    bool mfma;
    int unroll;
    size_t target;
};
const DbCand kDbCands[] = {{false, 4, 2048}, {false, 8, 2048}, {false, 2, 2048}, {false, 4, 4096},
                           {false, 8, 1024}, {true, 0, 512},   {true, 0, 1024},  {true, 0, 2048}};
const size_t kMfmaTargets[] = {256, 512, 1024, 2048};

DbPlan cand_plan(const DbCand& c, size_t R, size_t ncols) {
    return c.mfma ? dbplan_mfma(R, ncols, 1, c.target) : dbplan_f64(R, ncols, c.unroll, c.target);
}

// The plan of `full` (over all C columns) narrowed to ncols columns: the
// same kernel, with row slices multiplied so the grid keeps the full
// plan's workgroup count up to kChunkWgs (a narrow grid leaves most CUs
// idle; past that, more slices only add partial sums to write and reduce).
constexpr size_t kChunkWgs = 2048;
DbPlan narrow_plan(const DbPlan& full, size_t target, size_t R, size_t C, size_t ncols, size_t nq) {
    if (ncols == C) return full;
    size_t wgs = C / dbplan_cols(full) * full.slices;
    if (wgs > kChunkWgs) wgs = kChunkWgs;
    const size_t t = wgs > target ? wgs : target;
    return full.kind == DbPlan::F64 ? dbplan_f64(R, ncols, (int)full.param, t)
                                     : dbplan_mfma(R, ncols, nq, t, full.depth, full.dbtile, full.waves, full.crt,
                                                   full.pow);
}
}  // namespace

size_t default_group_chunk(size_t G) {
    // group chunks of the pipelined answer: about 8, of at least 8 groups
    return G / 8 >= 8 ? G / 8 : (G < 8 ? G : 8);
}

ServerFootprint server_footprint(const PublicParams& pp, size_t gchunk) {
    const size_t R = pp.db_rows, C = pp.db_cols, G = pp.n_packed, T = pp.num_cts;
    ServerFootprint f;
    // row-slice partials of one query: [slices][2][ncols] words for the
    // widest plan any candidate (full width or narrowed to a chunk) needs
    size_t part_words = 2 * C;
    auto need = [&](const DbPlan& x, size_t ncols) {
        if (x.slices > 1 && x.slices * 2 * ncols > part_words) part_words = x.slices * 2 * ncols;
    };
    const size_t ccols = gchunk * N;
    for (const DbCand& c : kDbCands) {
        const DbPlan full = cand_plan(c, R, C);
        need(full, C);
        need(narrow_plan(full, c.target, R, C, ccols, 1), ccols);
    }
    for (size_t t : kMfmaTargets)
        for (size_t nq = 1; nq <= kMfmaMaxQueries; nq++)
            for (int form = 0; form < kMfmaForms; form++) {  // the slice count depends on the block's columns
                const DbPlan full = dbplan_form(R, C, nq, t, 3, form);
                need(full, C);
                need(narrow_plan(full, t, R, C, ccols, nq), ccols);
            }
    // streamed passes' partials: the whole width at their largest slice count
    const uint32_t scap = stream_slices_cap(C);
    if (scap > 1 && (size_t)scap * 2 * C > part_words) part_words = (size_t)scap * 2 * C;
    f.part_words = part_words;
    SlotLayout lay{R, C, G};
    const size_t qlen = lay.query_u64s(), resp = T * 4 * N;
    f.per_slot = qlen * 8 + resp * 8 + 2 * R * 8 +
                 (lay.keys_words() + lay.rgsw_words() + part_words + 2 * C + 2 * G * 2 * N) * 4 +
                 (horner_pair_xbuf_words(T) + 2 * T + T * 4 * N) * 4 +
                 query_packed_bytes(R) + T * 4 * N * 4 + T * GpuServerCtx::kCompressedCtBytes;
    f.fixed = R * C * 2 + G * PC_GROUP * 4 + 8 * N * 4 +
              GpuServerCtx::kLanes * (R * 256 + GpuServerCtx::kKdigBytes + 2 * 64 * 4 + 4);
    // streamed answers' flags: one set per lane and 32 slots (rounded up)
    const size_t set_bytes = GpuServerCtx::kStreamSetWords * G * 4;
    f.fixed += GpuServerCtx::kLanes * set_bytes;
    f.per_slot += (GpuServerCtx::kLanes * set_bytes + kMfmaMaxQueries - 1) / kMfmaMaxQueries;
    // gpu_preprocess: the DB with its upload slab, then the DB, the arena and
    // one chunk's ring-embed table, chain states and D+/D- in reference order
    size_t slab = ((size_t)1 << 27) / C / 16 * 16;
    if (slab < 16) slab = 16;
    if (slab > R) slab = R;
    const size_t chunk = G < 128 ? G : 128;
    const size_t trans = (chunk * (2 * (size_t)N * N + 4 * (size_t)N + 2 * RD_SZ) + (R / N) * 2 * N +
                          4 * (size_t)D_EFF * N + 8 * (size_t)N) * 4;
    const size_t stage = slab * C * 2, after = G * PC_GROUP * 4 + trans;
    f.preprocess_peak = R * C * 2 + (stage > after ? stage : after);
    return f;
}

void GpuServerCtx::alloc_slots() {
    const size_t R = pp.db_rows, C = pp.db_cols, B = cfg.max_batch;
    plan1 = dbplan_f64(R, C, 4, 2048);
    plan1_target = 2048;
    gchunk = default_group_chunk(G);
    if (const char* gc = std::getenv("INSPIRE_GROUP_CHUNK")) {
        const long v = std::atol(gc);
        if (v > 0 && (size_t)v <= G && G % (size_t)v == 0) gchunk = (size_t)v;
    }
    // INSPIRE_TAIL_SPLIT=<0..3>: bit 0 splits the lowest chunk, bit 1 the highest
    if (const char* ts = std::getenv("INSPIRE_TAIL_SPLIT")) {
        const int v = std::atoi(ts);
        if (v >= 0 && v <= 3) {
            tail_split = v & 1;
            head_split = (v >> 1) & 1;
        }
    }
    set_chunks();
    const ServerFootprint fp = server_footprint(pp, gchunk);
    lay.R = R; lay.C = C; lay.G = G;
    qlen = lay.query_u64s();
    part_stride = fp.part_words;
    part2_stride = G * 2 * N;
    keys_stride = lay.keys_words();
    rgsw_stride = lay.rgsw_words();
    packed_stride = G * 2 * N;
    resp_u64s = T * 4 * N;
    d_q = dalloc<uint64_t>(B * qlen);
    d_resp = dalloc<uint64_t>(B * resp_u64s);
    d_bq = dalloc<double>(B * 2 * R);
    d_keys = dalloc<uint32_t>(B * keys_stride);
    d_rgsw = dalloc<uint32_t>(B * rgsw_stride);
    d_part = dalloc<uint32_t>(B * part_stride);
    d_interm = dalloc<uint32_t>(B * 2 * C);
    d_part2 = dalloc<uint32_t>(B * part2_stride);
    d_packed_b = dalloc<uint32_t>(B * packed_stride);
    for (int l = 0; l < kLanes; l++) {
        d_dig[l] = dalloc<int8_t>(R * 256);
        d_S[l] = dalloc<uint32_t>(64);
        d_kdig[l] = dalloc<int8_t>(kKdigBytes);
        d_KS[l] = dalloc<uint32_t>(64);
    }
    d_hstate = dalloc<uint32_t>(B * T * 4 * N);
    trace_name_stream(stream, "server");
    for (int l = 0; l < kLanes; l++) {
        HIPX(hipStreamCreate(&hstream[l]));
        HIPX(hipStreamCreate(&cstream[l]));
        static const char* hn[kLanes] = {"horner lane 0", "horner lane 1", "horner lane 2", "horner lane 3"};
        static const char* cn[kLanes] = {"collapse lane 0", "collapse lane 1", "collapse lane 2", "collapse lane 3"};
        trace_name_stream(hstream[l], hn[l]);
        trace_name_stream(cstream[l], cn[l]);
        for (auto& e : chunk_ev[l]) HIPX(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (auto& e : coll_ev[l]) HIPX(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIPX(hipEventCreateWithFlags(&prep_ev[l], hipEventDisableTiming));
        HIPX(hipEventCreateWithFlags(&pass_ev[l], hipEventDisableTiming));
        HIPX(hipEventCreateWithFlags(&horner_done[l], hipEventDisableTiming));
    }
    stream_sets = (B + kMfmaMaxQueries - 1) / kMfmaMaxQueries;
    for (int l = 0; l < kLanes; l++) {
        const size_t words = stream_sets * kStreamSetWords * G;
        d_sflags[l] = dalloc<uint32_t>(words);
        HIPX(hipMemset(d_sflags[l], 0, words * 4));
    }
    // INSPIRE_STREAM=0: no streamed shapes; otherwise the device's first
    // live server takes them (see stream_ok)
    const char* se = std::getenv("INSPIRE_STREAM");
    if (!(se && se[0] == '0')) {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        stream_ok = (size_t)device < kMaxDevices && !g_streaming[device];
        if (stream_ok) g_streaming[device] = this;
    }
    d_hxbuf = dalloc<uint32_t>(horner_pair_xbuf_words(B * T));
    d_hflags = dalloc<uint32_t>(2 * B * T);
    d_status = dalloc<uint32_t>(kLanes);
    HIPX(hipMemset(d_status, 0, kLanes * 4));
    horner_spin_limit = kHornerSpinLimit;
    stream_spin_limit = kStreamSpinLimit;
    HIPX(hipEventCreateWithFlags(&order_ev, hipEventDisableTiming));
    // INSPIRE_PACK_WG=0/1/-1: the pipelined chunks' pack kernel, a grid over
    // (group, limb), one workgroup per query, or one workgroup for one query
    if (const char* pw = std::getenv("INSPIRE_PACK_WG"))
        pack_wg = pw[0] == '1' ? 1 : pw[0] == '0' ? 0 : pw[0] == '2' ? 2 : -1;
    // INSPIRE_GRAPH=1: answers replay captured hipGraphs
    if (const char* gm = std::getenv("INSPIRE_GRAPH")) graph_mode = gm[0] == '1' ? 1 : 0;
    if (const char* h = std::getenv("INSPIRE_HORNER"))
        horner_mode = std::strcmp(h, "pair") == 0 ? 1 : std::strcmp(h, "single") == 0 ? 0 : -1;
    // INSPIRE_PIPELINE=<single>,<batched>: the pipeline shape (0 serial, 2
    // two streams, 3 three streams, 4 the collapse first, 5 three streams
    // with the collapse a chunk ahead, 6 and 7 streamed after or beside the
    // collapse, 8 streamed with the collapse in the pass's launch) of
    // one-query and many-query calls
    if (const char* pl = std::getenv("INSPIRE_PIPELINE")) {
        int a1 = -1, a2 = -1;
        auto ok = [](int v) { return v == 0 || (v >= 2 && v <= 8); };
        if (std::sscanf(pl, "%d,%d", &a1, &a2) == 2 && ok(a1) && ok(a2)) {
            pipe_single = a1;
            pipe_batched = a2;
            pipe_forced = true;
        }
    }
    // packed-wire path: packed queries, coefficient-form scratch, compressed responses
    qpack_bytes = query_packed_bytes(R);
    d_qpack = dalloc<uint8_t>(B * qpack_bytes);
    d_cmp = dalloc<uint32_t>(B * T * 4 * N);
    d_rcmp = dalloc<uint8_t>(B * T * kCompressedCtBytes);
    HIPX(hipMemset(d_bq, 0, B * 2 * R * sizeof(double)));
    HIPX(hipHostMalloc((void**)&h_q, B * qlen * 8, hipHostMallocDefault));
    HIPX(hipHostMalloc((void**)&h_resp, B * resp_u64s * 8, hipHostMallocDefault));
    resident_bytes = fp.fixed + B * fp.per_slot;
    autotune_dbpass();
}

void GpuServerCtx::autotune_dbpass() {
    // INSPIRE_DBPASS forces the single-query plan: "f64,<unroll>,<target>" or
    // "mfma,<target>" (the batched forms are still tuned); INSPIRE_MFMA_TARGET
    // sets the batched pass's target; INSPIRE_AUTOTUNE=0 keeps the defaults.
    const size_t R = pp.db_rows, C = pp.db_cols;
    if (const char* t = std::getenv("INSPIRE_MFMA_TARGET")) {
        const long v = std::atol(t);
        if (v > 0) mfma_target = (size_t)v;
    }
    // INSPIRE_MFMA_DEPTH=<depth>[,<form>] forces the batched passes' ring
    // depth and form (dbplan_form; 0, shared digits in 4-wave blocks, by default)
    bool depth_forced = false;
    if (const char* t = std::getenv("INSPIRE_MFMA_DEPTH")) {
        const long v = std::atol(t);
        for (uint32_t d : kMfmaDepths)
            if (v == (long)d) { mfma_depth = d; depth_forced = true; }
        const char* f = std::strchr(t, ',');
        const long form = f ? std::atol(f + 1) : 0;
        mfma_form = depth_forced && form >= 0 && form < kMfmaForms ? (int)form : 0;
    }
    bool plan1_forced = false;
    if (const char* force = std::getenv("INSPIRE_DBPASS")) {
        int u = 4, t = 2048;
        if (std::sscanf(force, "f64,%d,%d", &u, &t) == 2 && (u == 2 || u == 4 || u == 8) && t > 0) {
            plan1 = dbplan_f64(R, C, u, (size_t)t);
            plan1_target = (size_t)t;
            plan1_forced = true;
        } else if (std::sscanf(force, "mfma,%d", &t) == 1 && t > 0) {
            plan1 = dbplan_mfma(R, C, 1, (size_t)t);
            plan1_target = (size_t)t;
            plan1_forced = true;
        }
    }
    const char* at = std::getenv("INSPIRE_AUTOTUNE");
    if (at && at[0] == '0') return;
    TraceRange tr("setup: autotune");
    hipEvent_t e0, e1;
    HIPX(hipEventCreate(&e0));
    HIPX(hipEventCreate(&e1));
    // Candidates are timed round-robin: every round runs each candidate once
    // (after its own preparation, untimed) and the first round only warms
    // up. Under sustained int8-MFMA work the power limit lowers the clock
    // within tens of milliseconds, so timing the candidates one after
    // another would favour whichever ran first.
    auto fastest = [&](size_t n, auto&& prep, auto&& run, int rounds = 4) {
        std::vector<float> tot(n, 0.f);
        for (int round = 0; round < rounds; round++)
            for (size_t i = 0; i < n; i++) {
                prep(i);
                HIPX(hipEventRecord(e0, stream));
                run(i);
                HIPX(hipEventRecord(e1, stream));
                HIPX(hipEventSynchronize(e1));
                float ms = 0;
                HIPX(hipEventElapsedTime(&ms, e0, e1));
                if (round) tot[i] += ms;
            }
        size_t b = 0;
        for (size_t i = 1; i < n; i++)
            if (tot[i] < tot[b]) b = i;
        return b;
    };
    // one full-width database pass of plan x over nq staged queries (zeroed)
    auto prep_plan = [&](const DbPlan& x, size_t nq) {
        if (x.kind == DbPlan::MFMA) HIPX(launch_mfma_digits(d_bq, nq, R, x, d_dig[0], d_S[0], stream));
    };
    auto run_plan = [&](const DbPlan& x, size_t nq) {
        if (x.kind == DbPlan::F64)
            HIPX(launch_dbpass_f64(db_planes(), d_bq, d_part, C, R, C, 0, C, x, stream));
        else
            HIPX(launch_dbpass_mfma(db_planes(), d_dig[0], d_S[0], nq, d_part, part_stride, C, R, C, 0, C, x, stream));
    };
    // a plan whose row-slice partials would not fit their buffer is not run
    auto fits = [&](const DbPlan& x) { return x.slices == 1 || x.slices * 2 * C <= part_stride; };
    if (!plan1_forced) {
        std::vector<DbPlan> xs;
        std::vector<size_t> ts;
        for (const DbCand& c : kDbCands) {
            const DbPlan x = cand_plan(c, R, C);
            if (!fits(x)) continue;
            xs.push_back(x);
            ts.push_back(c.target);
        }
        if (!xs.empty()) {
            const size_t b = fastest(
                xs.size(), [&](size_t i) { prep_plan(xs[i], 1); }, [&](size_t i) { run_plan(xs[i], 1); });
            plan1 = xs[b];
            plan1_target = ts[b];
        }
    }
    // batched pass: target, depth and form at the largest batch this context
    // serves, and apart for 9..16 queries when that batch is larger (their
    // best form differs)
    const size_t nq = cfg.max_batch < kMfmaMaxQueries ? cfg.max_batch : kMfmaMaxQueries;
    auto tune_batched = [&](size_t nqc, size_t& target_out, uint32_t& depth_out, int& form_out) {
        const uint32_t param = dbplan_mfma(R, C, nqc, target_out).param;
        // forms (dbplan_form): the shared DB tile and the 7-digit forms
        // regrouped through LDS exist for 17..32 queries only, the power rows
        // for 9..32; the LDS-regrouped forms (3, 4) never beat the power rows
        // (5) and are timed only when forced
        const int forced_form = form_out;
        struct Cand {
            DbPlan x;
            size_t target;
            uint32_t depth;
            int form;
        };
        std::vector<Cand> cands;  // targets that give the same grid are timed once
        auto seen = [&](const DbPlan& x) {
            for (const Cand& y : cands)
                if (y.x.param == x.param && y.x.slices == x.slices && y.x.depth == x.depth &&
                    y.x.dbtile == x.dbtile && y.x.waves == x.waves && y.x.crt == x.crt && y.x.pow == x.pow)
                    return true;
            return false;
        };
        for (size_t t : kMfmaTargets)
            for (int form = 0; form < kMfmaForms; form++) {
                if (depth_forced ? form != forced_form
                                 : (form == 1 && param != 16) || (form > 0 && param < 8) || form == 3 || form == 4)
                    continue;
                for (uint32_t d : kMfmaDepths) {
                    if ((depth_forced || param < 8) && d != depth_out) continue;
                    const DbPlan x = dbplan_form(R, C, nqc, t, d, form);
                    if (x.depth != d && param >= 8) continue;  // a depth this kernel is not built with
                    if (fits(x) && !seen(x)) cands.push_back({x, t, d, form});
                }
            }
        if (!cands.empty()) {  // empty: a forced depth the forced form is not built with keeps the defaults
            const size_t b = fastest(
                cands.size(), [&](size_t i) { prep_plan(cands[i].x, nqc); },
                [&](size_t i) { run_plan(cands[i].x, nqc); });
            target_out = cands[b].target;
            depth_out = cands[b].depth;
            form_out = cands[b].form;
        }
    };
    if (nq > 1) {
        tune_batched(nq, mfma_target, mfma_depth, mfma_form);
        if (nq > 16 && !depth_forced) {
            mfma16_target = mfma_target;
            mfma16_depth = mfma_depth;
            tune_batched(16, mfma16_target, mfma16_depth, mfma16_form);
            tuned16 = true;
        }
        // the 17..32-query collapse pass form, over all groups (zeroed keys)
        if (nq > 16) {
            HIPX(hipMemset(d_keys, 0, nq * keys_stride * 4));
            HIPX(launch_collapse_keys(d_keys, keys_stride, nq, d_kdig[0], d_KS[0], stream));
            collapse_variant = (int)fastest(
                (size_t)kCollapseVariants, [](size_t) {},
                [&](size_t v) {
                    HIPX(launch_collapse(d_precomp, G, d_kdig[0], d_KS[0], nq, d_part2, part2_stride, stream, (int)v));
                });
        }
    }
    // The pipelined answer's chunk plans and pipeline shapes, timed together
    // on whole answers of a synthetic (zero) batch, one query and the
    // largest batch, with the Horner chains beside the passes as when
    // serving. A chunk's database pass runs the full-width plan's kernel over
    // the chunk's columns with a row-slice count (0: the narrowing rule); the
    // shape puts the collapse beside it (3), or after it (2); the forms with
    // the collapse a chunk ahead (5) or all of it first (4) are timed only
    // when forced. The two interact:
    // a one-slice grid needs no reduction kernel and leaves CUs that only a
    // collapse beside it uses; more slices fill the device alone but pay
    // partials and their reduction. The best of these and the serial answer
    // is then timed against the streamed shapes
    // at each row-slice count their full-width pass can take.
    const bool chunked_ok = G / gchunk > 1;
    if (chunked_ok || can_stream(stream)) {
        const size_t ccols = gchunk * N;
        HIPX(hipMemset(d_q, 0, nq * qlen * 8));
        for (const size_t count : {(size_t)1, nq}) {
            const DbPlan full = plan_for(count, C);
            const size_t blocks = ccols / dbplan_cols(full);
            struct Pick {
                uint32_t slices;
                int shape;
            };
            std::vector<Pick> cands;
            const int forced = count == 1 ? pipe_single : pipe_batched;
            // the serial answer: with many chunks it can still win where a
            // sub-batch's Horner chains hold many of the CUs the chunked
            // passes would run beside them
            if (!pipe_forced || forced == 0 || (!chunked_ok && forced < 6 && forced != 3)) cands.push_back({0, 0});
            // one chunk: the collapse beside the whole pass (the two branches'
            // small kernels and launch gaps overlap)
            if (!chunked_ok && (!pipe_forced || forced == 3)) cands.push_back({0, 3});
            for (uint32_t sl = 0; chunked_ok && sl <= 8; sl = sl ? 2 * sl : 1) {
                if (sl) {
                    const size_t t = blocks * sl;
                    const DbPlan x = count > 1 ? dbplan_form(R, ccols, count, t, mfma_depth, mfma_form)
                                     : full.kind == DbPlan::F64 ? dbplan_f64(R, ccols, (int)full.param, t)
                                                                : dbplan_mfma(R, ccols, 1, t);
                    if (x.slices != sl || (sl > 1 && sl * 2 * ccols > part_stride)) continue;
                }
                // shapes 4 and 5 never won at 16 GB and take setup time: forced only
                for (const int shape : {2, 3, 4, 5}) {
                    if (!pipe_forced && shape > 3) continue;
                    if (pipe_forced && shape != (count == 1 ? pipe_single : pipe_batched)) continue;
                    cands.push_back({sl, shape});
                }
            }
            // streamed shapes, at each row-slice count a streamed plan takes
            // (the slices field is then the streamed pass's) from 2: with one
            // slice every workgroup streams its tile's whole column stripe, so
            // groups complete in a few bursts and the chains wait, then trail
            // the pass (at 16 GB 0.2-0.5 ms slower than 2..8 slices in every
            // measurement; a setup timing that favours it is noise)
            const uint32_t sl0 = stream_slices_cap(C) >= 2 ? 2 : 1;
            if (can_stream(stream) && stream_fits(count))
                for (uint32_t sl = sl0; sl <= stream_slices_cap(C); sl *= 2) {
                    const DbPlan x = stream_plan(count, sl);
                    if (x.kind != DbPlan::MFMA || x.slices != sl) continue;
                    // shape 7 (the collapse beside) never won at 16 GB: forced only;
                    // shape 8 (the collapse in the pass's launch) up to 8 queries
                    for (const int shape : {6, 7, 8})
                        if (pipe_forced ? shape == forced : shape == 6 || (shape == 8 && count <= 8 && x.param <= 4))
                            cands.push_back({sl, shape});
                }
            auto use = [&](const Pick& p) {
                const bool st = p.shape >= 6;
                if (count == 1) {
                    if (st) stream_slices1 = p.slices; else chunk_slices1 = p.slices;
                    if (!pipe_forced) pipe_single = p.shape;
                } else {
                    if (st) {
                        stream_slices = p.slices;
                    } else {
                        chunk_slices = p.slices;
                        chunk_param = p.slices ? full.param : 0;
                    }
                    if (!pipe_forced) pipe_batched = p.shape;
                }
            };
            // the chunked shapes first (the fallback where an answer cannot
            // stream), then the best of them against the streamed ones
            std::vector<Pick> chunked, streamed;
            for (const Pick& c : cands) (c.shape >= 6 ? streamed : chunked).push_back(c);
            auto time_picks = [&](std::vector<Pick>& v) {
                const size_t b = fastest(
                    v.size(), [&](size_t i) { use(v[i]); }, [&](size_t) { enqueue(d_q, count, d_resp, stream); }, 6);
                use(v[b]);
                return v[b];
            };
            if (!chunked.empty()) {
                const Pick c = time_picks(chunked);
                (count == 1 ? pipe_single_chunked : pipe_batched_chunked) = c.shape;
                if (!streamed.empty()) streamed.insert(streamed.begin(), c);
            }
            // (a split streamed pass, the chains released after its first
            // launch, never won at 16 GB: set by hand only)
            if (!streamed.empty()) time_picks(streamed);
            if (nq == 1) break;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

void GpuServerCtx::stage_prep(const uint64_t* d_in, size_t count, hipStream_t s, size_t slot0) {
    HIPX(launch_unmarshal(d_in, count, pp.db_rows, d_bq + slot0 * 2 * pp.db_rows, d_keys + slot0 * keys_stride,
                          d_rgsw + slot0 * rgsw_stride, s));
    HIPX(launch_ntt_rns(d_keys + slot0 * keys_stride, count * 2 * D_EFF, nc, false, s));
}

DbPlan GpuServerCtx::plan_for(size_t nq, size_t ncols) const {
    const size_t R = pp.db_rows;
    if (nq == 1)
        return plan1.kind == DbPlan::F64 ? dbplan_f64(R, ncols, (int)plan1.param, plan1_target)
                                         : dbplan_mfma(R, ncols, 1, plan1_target);
    if (tuned16 && nq >= 9 && nq <= 16) return dbplan_form(R, ncols, nq, mfma16_target, mfma16_depth, mfma16_form);
    return dbplan_form(R, ncols, nq, mfma_target, mfma_depth, mfma_form);
}

void GpuServerCtx::set_chunks() {
    chunks.clear();
    const size_t n = G / gchunk, h = gchunk / 2, q = gchunk / 4;
    const bool quarters = n > 1 && gchunk % 4 == 0;
    for (size_t k = n; k-- > 0;) {
        const size_t lo = k * gchunk, hi = lo + gchunk;
        if (quarters && head_split && k == n - 1) {
            chunks.push_back({hi - q, hi});
            chunks.push_back({hi - h, hi - q});
            chunks.push_back({lo, hi - h});
        } else if (quarters && tail_split && k == 0) {
            chunks.push_back({lo + h, hi});
            chunks.push_back({lo + q, lo + h});
            chunks.push_back({lo, lo + q});
        } else {
            chunks.push_back({lo, hi});
        }
    }
}

// The chunk plan over ncols (<= gchunk N) columns: a timed slice count sets
// the workgroup count of a whole chunk, which narrower chunks keep by
// cutting their rows into more slices.
DbPlan GpuServerCtx::chunk_plan(size_t nq, size_t ncols) const {
    const DbPlan full = plan_for(nq, pp.db_cols);
    const size_t whole = gchunk * N;
    auto fits = [&](const DbPlan& x) { return x.slices == 1 || x.slices * 2 * ncols <= part_stride; };
    if (nq == 1 && chunk_slices1) {
        const size_t t = whole / dbplan_cols(full) * chunk_slices1;
        const DbPlan x = full.kind == DbPlan::F64 ? dbplan_f64(pp.db_rows, ncols, (int)full.param, t)
                                                  : dbplan_mfma(pp.db_rows, ncols, 1, t);
        if (fits(x)) return x;
    }
    if (nq > 1 && chunk_slices && full.kind == DbPlan::MFMA && full.param == chunk_param) {
        const DbPlan x =
            dbplan_form(pp.db_rows, ncols, nq, whole / dbplan_cols(full) * chunk_slices, mfma_depth, mfma_form);
        if (fits(x)) return x;
    }
    return narrow_plan(full, nq == 1 ? plan1_target : mfma_target, pp.db_rows, pp.db_cols, ncols, nq);
}

DbPlan GpuServerCtx::chunk_plan_at(size_t nq, size_t k, size_t ng, const DbPlan& whole) const {
    const uint32_t from = nq == 1 ? late_from1 : late_from, sl = nq == 1 ? late_slices1 : late_slices;
    if (sl && k >= from) {
        // the chunk's columns at sl row slices per column block (a narrower
        // chunk multiplies them to keep a whole chunk's workgroup count)
        const size_t ncols = ng * N, t = gchunk * N / dbplan_cols(whole) * sl;
        const DbPlan x = whole.kind == DbPlan::F64
                             ? dbplan_f64(pp.db_rows, ncols, (int)whole.param, t)
                             : dbplan_mfma(pp.db_rows, ncols, nq, t, whole.depth, whole.dbtile, whole.waves, whole.crt,
                                           whole.pow);
        if (x.param == whole.param && (x.slices == 1 || x.slices * 2 * ncols <= part_stride)) return x;
    }
    return ng == gchunk ? whole : chunk_plan(nq, ng * N);
}

DbPlan GpuServerCtx::stream_plan(size_t nq, uint32_t slices, bool fused) const {
    const size_t R = pp.db_rows, C = pp.db_cols;
    DbPlan p = plan_for(nq, C);
    if (p.kind != DbPlan::MFMA || p.dbtile)  // a kernel that cannot stream: the direct or 8-wave form
        p = nq == 1 ? dbplan_mfma(R, C, 1, plan1_target) : dbplan_form(R, C, nq, mfma_target, mfma_depth, 2);
    if (p.param == 1 && !fused) p.nt = 8;
    // at least the slices the exactness bounds need (2^15 rows for power rows,
    // else 2^16), at most kStreamSlicesMax (the pack's 32-bit partial sums)
    uint32_t sl = slices ? slices : 1;
    while (R / sl > (p.pow ? 32768u : 65536u)) sl *= 2;
    p.slices = sl;
    const size_t kbps = R / 16 / sl;
    if (!dbplan_streams(p, C) || sl > kStreamSlicesMax || kbps < 8 || kbps % 8 ||
        (sl > 1 && (size_t)sl * 2 * C > part_stride))
        p.kind = DbPlan::F64;  // no streamed plan
    return p;
}

void GpuServerCtx::take_streaming() {
    std::lock_guard<std::mutex> lk(g_stream_mu);
    if ((size_t)device >= kMaxDevices) return;
    GpuServerCtx* prev = g_streaming[device];
    if (prev == this) return;
    if (prev) {
        std::lock_guard<std::mutex> pl(prev->mu);  // its calls see the change between answers
        prev->stream_ok = false;
    }
    g_streaming[device] = this;
    stream_ok = true;
}

void GpuServerCtx::mark_pass(hipStream_t s, bool first) {
    if (!time_pass) return;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (s && hipStreamIsCapturing(s, &cap) == hipSuccess && cap != hipStreamCaptureStatusNone) return;
    if (pass_marks_used + 1 >= kMaxPassMarks) {  // pool full: the rest go untimed
        time_pass_overflow = true;
        return;
    }
    while (pass_marks.size() <= pass_marks_used) {
        hipEvent_t e = nullptr;
        HIPX(hipEventCreate(&e));
        pass_marks.push_back(e);
    }
    // marks alternate start / end; a whole answer's database pass (one
    // launch, or the launches of its chunks) counts once
    if (first && pass_marks_used % 2 == 0) passes_timed++;
    HIPX(hipEventRecord(pass_marks[pass_marks_used++], s));
}

bool GpuServerCtx::stream_fits(size_t nq) const {
    return nq * T + (nq + pack_width(nq) - 1) / pack_width(nq) <= kStreamMaxWaiters;
}

bool GpuServerCtx::can_stream(hipStream_t s) const {
    if (!stream_ok) return false;
    // a graph replays its branches one after another, so chains captured
    // beside a streamed pass could wait for a pass that runs after them
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (s && hipStreamIsCapturing(s, &cap) != hipSuccess) return false;
    return cap == hipStreamCaptureStatusNone;
}

void GpuServerCtx::dbpass_prep(size_t q0, size_t nq, const DbPlan& p, hipStream_t s, int lane) {
    if (p.kind == DbPlan::MFMA)
        HIPX(launch_mfma_digits(d_bq + q0 * 2 * pp.db_rows, nq, pp.db_rows, p, d_dig[lane], d_S[lane], s));
}

void GpuServerCtx::dbpass_range(size_t q0, size_t nq, size_t g_lo, size_t g_hi, const DbPlan& p, hipStream_t s,
                                int lane) {
    const size_t R = pp.db_rows, C = pp.db_cols, col0 = g_lo * N, ncols = (g_hi - g_lo) * N;
    // one slice: straight into interm; else row-slice partials [slice][2][ncols]
    // relative to col0, reduced below
    const bool direct = p.slices == 1;
    if (!direct && p.slices * 2 * ncols > part_stride)
        throw std::logic_error("database pass: row-slice partials exceed their buffer");
    uint32_t* out = direct ? d_interm + q0 * 2 * C + col0 : d_part + q0 * part_stride;
    const size_t qstride = direct ? 2 * C : part_stride, ostride = direct ? C : ncols;
    if (p.kind == DbPlan::F64) {
        for (size_t q = 0; q < nq; q++)
            HIPX(launch_dbpass_f64(db_planes(), d_bq + (q0 + q) * 2 * R, out + q * qstride, ostride, R, C, col0, ncols,
                                   p, s));
    } else {
        HIPX(launch_dbpass_mfma(db_planes(), d_dig[lane], d_S[lane], nq, out, qstride, ostride, R, C, col0, ncols, p,
                                s));
    }
    if (p.slices > 1)
        HIPX(launch_reduce_slices(d_part + q0 * part_stride, part_stride, p.slices, C, d_interm + q0 * 2 * C, nq, col0,
                                  ncols, s));
}

void GpuServerCtx::stage_dbpass(size_t count, hipStream_t s, size_t slot0, int lane, int force) {
    const size_t R = pp.db_rows, C = pp.db_cols;
    const size_t step = force == 0 ? 1 : kMfmaMaxQueries;
    for (size_t q0 = slot0; q0 < slot0 + count; q0 += step) {
        const size_t nq = slot0 + count - q0 < step ? slot0 + count - q0 : step;
        DbPlan p = plan_for(nq, C);
        if (force == 0 && p.kind != DbPlan::F64) p = dbplan_f64(R, C, 4, 2048);
        if (force == 1 && p.kind == DbPlan::F64)
            p = dbplan_form(R, C, nq, mfma_target, mfma_depth, mfma_form);
        dbpass_prep(q0, nq, p, s, lane);
        dbpass_range(q0, nq, 0, G, p, s, lane);
    }
}

void GpuServerCtx::stage_collapse(size_t count, hipStream_t s, size_t slot0, int lane) {
    for (size_t q0 = slot0; q0 < slot0 + count; q0 += kCollapseMaxQueries) {
        const size_t nq = slot0 + count - q0 < kCollapseMaxQueries ? slot0 + count - q0 : kCollapseMaxQueries;
        HIPX(launch_collapse_keys(d_keys + q0 * keys_stride, keys_stride, nq, d_kdig[lane], d_KS[lane], s));
        HIPX(launch_collapse(d_precomp, G, d_kdig[lane], d_KS[lane], nq, d_part2 + q0 * part2_stride, part2_stride,
                             s, collapse_variant));
    }
}

void GpuServerCtx::stage_finalize(size_t count, hipStream_t s, size_t slot0) {
    HIPX(launch_pack_finalize(d_interm + slot0 * 2 * pp.db_cols, pp.db_cols, d_part2 + slot0 * part2_stride,
                              part2_stride, d_precomp, 0, G, d_keys + slot0 * keys_stride, keys_stride,
                              d_packed_b + slot0 * packed_stride, packed_stride, count, nc, s));
}

bool GpuServerCtx::use_pair(size_t chains, bool beside_passes, bool status_read) const {
    // never where nobody reads the status its bounded wait reports through;
    // auto: the pair kernel shortens a lone chain run by itself (its per-step
    // exchange costs more than it saves once many chains share the L2s, or
    // while database passes stream through them)
    if (!status_read) return false;
    return horner_mode == 1 || (horner_mode < 0 && chains <= kPairMaxChains && !beside_passes);
}

void GpuServerCtx::horner_range(size_t count, uint64_t* d_out, size_t g_lo, size_t g_hi, hipStream_t s,
                                size_t slot0, int lane, bool status_read, bool beside_passes, const FusedPack* fp) {
    if (use_pair(count * T, beside_passes, status_read))
        HIPX(launch_horner_pair(d_precomp + PC_A, PC_GROUP, G, Da, T, d_packed_b + slot0 * packed_stride,
                                packed_stride, d_rgsw + slot0 * rgsw_stride, rgsw_stride, d_out,
                                d_hstate + slot0 * T * 4 * N, g_lo, g_hi, count,
                                d_hxbuf + horner_pair_xbuf_words(slot0 * T), d_hflags + 2 * slot0 * T,
                                d_status + lane, horner_spin_limit, nc, cc, s));
    else
        HIPX(launch_horner(d_precomp + PC_A, PC_GROUP, G, Da, T, d_packed_b + slot0 * packed_stride, packed_stride,
                           d_rgsw + slot0 * rgsw_stride, rgsw_stride, d_out, d_hstate + slot0 * T * 4 * N, g_lo, g_hi,
                           count, nc, cc, s, fp ? *fp : FusedPack{}));
}

void GpuServerCtx::stage_horner(size_t count, uint64_t* d_out, hipStream_t s, size_t slot0, bool status_read) {
    horner_range(count, d_out, 0, G, s, slot0, 0, status_read);
}

void GpuServerCtx::enqueue(const uint64_t* d_in, size_t count, uint64_t* d_out, hipStream_t s, size_t slot0,
                           int lane, bool status_read) {
    if (lane < 0 || lane >= kLanes) throw std::invalid_argument("enqueue: scratch lane out of range");
    TraceRange tr("answer: %zu queries, slots %zu.., lane %d", count, slot0, lane);
    stage_prep(d_in, count, s, slot0);
    const size_t C = pp.db_cols;
    const size_t nchunks = G / gchunk;
    const hipStream_t hs = hstream[lane];
    // Pipelined over group chunks, highest first (the order every Horner
    // chain consumes groups). The database pass (on s) and the collapse (on
    // s, or on the lane's collapse stream) need nothing from each other; the
    // pack and the Horner chains of a chunk (on the lane's Horner stream)
    // wait for both halves of that chunk. s finally waits for the chains,
    // and through their waits for every collapse.
    bool joined = false;
    const bool streamable = can_stream(s);
    last_part_slices = 0;  // the first sub-batch's result lands in interm unless it streams with slices
    // Serial sub-batches (shape 0) leave their Horner chains to one launch
    // after the last of them: a sub-batch's nq x T chains each hold a CU for
    // the whole chain, and the chains of all of them together fill the
    // device (at 128 GB, 4 x 128 chains in two rounds over the 256 CUs
    // instead of four rounds over 128). Contiguous slots [defer_lo, +defer_n).
    size_t defer_lo = 0, defer_n = 0;
    auto flush_deferred = [&]() {
        if (defer_n)
            horner_range(defer_n, d_out + (defer_lo - slot0) * resp_u64s, 0, G, s, defer_lo, lane, status_read);
        defer_n = 0;
    };
    for (size_t q0 = slot0; q0 < slot0 + count; q0 += kMfmaMaxQueries) {
        const size_t nq = slot0 + count - q0 < kMfmaMaxQueries ? slot0 + count - q0 : kMfmaMaxQueries;
        int shape = nq == 1 ? pipe_single : pipe_batched;
        // one chunk: nothing to pipeline but the collapse beside the pass (3)
        if (shape < 6 && shape != 3 && nchunks == 1) shape = 0;
        uint32_t* p2 = d_part2 + q0 * part2_stride;
        const uint32_t* keys = d_keys + q0 * keys_stride;
        uint32_t* packed = d_packed_b + q0 * packed_stride;
        uint64_t* out = d_out + (q0 - slot0) * resp_u64s;
        DbPlan sp;
        // shape 8 (the collapse inside the streamed pass's launch) takes the
        // direct pass kernel: up to 8 queries, else shape 6
        if (shape == 8 && (nq > 8 || stream_plan(nq, nq == 1 ? stream_slices1 : stream_slices).param > 4)) shape = 6;
        if (shape >= 6) {
            sp = stream_plan(nq, nq == 1 ? stream_slices1 : stream_slices, shape == 8);
            // a forced pair Horner cannot wait for a streamed pass, and a
            // sub-batch whose waiting workgroups exceed kStreamMaxWaiters
            // CUs would leave too few to the passes of 4 lanes: the chunked
            // shape then
            if (!streamable || sp.kind != DbPlan::MFMA || horner_mode == 1 || !stream_fits(nq))
                shape = nchunks == 1 ? 0 : nq == 1 ? pipe_single_chunked : pipe_batched_chunked;
        }
        if (shape != 0) flush_deferred();
        if (shape >= 6) {
            // Streamed: one full-width database pass on s hands each group
            // over as soon as its columns are in place (highest first); the
            // pack kernel on the lane's collapse stream packs each group as
            // it arrives and publishes it to the Horner chains on the Horner
            // stream. The collapse runs on s before the pass (6) or on the
            // collapse stream beside it (7); the pack and the chains start
            // after it. Every consumer is enqueued after its producer, so
            // wherever two of these streams share a hardware queue the
            // producer runs first.
            TraceRange trs("streamed pass: %zu queries, %u row slices, shape %d", nq, sp.slices, shape);
            n_streamed.fetch_add(1, std::memory_order_relaxed);
            if (q0 == slot0) last_part_slices = sp.slices;
            const bool fused = shape == 8;
            const hipStream_t ps_s = cstream[lane], cs = shape == 7 ? ps_s : s;
            joined = true;
            uint32_t* pass_done = d_sflags[lane] + ((q0 - slot0) / kMfmaMaxQueries) * kStreamSetWords * G;
            uint32_t* packed_done = pass_done + G;
            const uint32_t pq = pack_width(nq);
            const size_t nblk = (nq + pq - 1) / pq;
            HIPX(launch_fill_u32(pass_done, (1 + nblk) * G, 0, s));  // this set's last readers are done
            if (cs != s) {
                HIPX(hipEventRecord(prep_ev[lane], s));
                HIPX(hipStreamWaitEvent(cs, prep_ev[lane], 0));
            }
            HIPX(launch_collapse_keys(keys, keys_stride, nq, d_kdig[lane], d_KS[lane], cs));
            if (!fused)
                HIPX(launch_collapse(d_precomp, G, d_kdig[lane], d_KS[lane], nq, p2, part2_stride, cs,
                                     collapse_variant));
            // The pack and the chains are released here (fused: the flags and
            // key digits in place), before the query digits: each of their
            // workgroups needs a whole CU, which they get while those small
            // kernels run; released together with the pass they would wait
            // for a CU the pass's workgroups leave wholly free, often its end.
            HIPX(hipEventRecord(coll_ev[lane][0], cs));
            dbpass_prep(q0, nq, sp, s, lane);
            uint32_t* interm = d_interm + q0 * 2 * C;
            const bool direct = sp.slices == 1;
            uint32_t* part = direct ? interm : d_part + q0 * part_stride;
            const size_t part_qstride = direct ? 2 * C : part_stride;
            // split: the top eighths of the groups in a first launch, the
            // chains released after it (they hold their CUs only from then)
            const uint32_t split = fused ? 0 : nq == 1 ? stream_split1 : stream_split;
            const size_t g_split = split ? G - G * split / 8 : 0;
            PassStream pst;
            pst.group_done = pass_done;
            mark_pass(s, true);
            if (fused) {
                HIPX(launch_pass_collapse_stream(db_planes(), d_dig[lane], d_S[lane], nq, part, part_qstride,
                                                 pp.db_rows, C, sp, d_precomp, d_kdig[lane], d_KS[lane], p2,
                                                 part2_stride, pass_done, s));
            } else if (g_split > 0 && g_split < G) {
                HIPX(launch_dbpass_mfma(db_planes(), d_dig[lane], d_S[lane], nq, part + g_split * N, part_qstride, C,
                                        pp.db_rows, C, g_split * N, (G - g_split) * N, sp, s, &pst));
                HIPX(hipEventRecord(pass_ev[lane], s));
                HIPX(launch_dbpass_mfma(db_planes(), d_dig[lane], d_S[lane], nq, part, part_qstride, C, pp.db_rows,
                                        C, 0, g_split * N, sp, s, &pst));
            } else {
                HIPX(launch_dbpass_mfma(db_planes(), d_dig[lane], d_S[lane], nq, part, part_qstride, C, pp.db_rows,
                                        C, 0, C, sp, s, &pst));
            }
            mark_pass(s, false);
            if (ps_s != cs) HIPX(hipStreamWaitEvent(ps_s, coll_ev[lane][0], 0));
            PackStream pk;
            pk.part = part;
            pk.part_qstride = part_qstride;
            pk.slices = sp.slices;
            pk.C = (uint32_t)C;
            pk.pass_done = pass_done;
            pk.pass_target = fused ? pass_collapse_target(sp, nq, G) : (uint32_t)(N / dbplan_cols(sp) * sp.slices);
            pk.packed_done = packed_done;
            pk.spin_limit = stream_spin_limit;
            pk.status = d_status + lane;
            pk.queries = pq;
            HIPX(launch_pack_stream(pk, p2, part2_stride, d_precomp, keys, keys_stride, packed, packed_stride, G, nq,
                                    nc, ps_s));
            HIPX(hipStreamWaitEvent(hs, coll_ev[lane][0], 0));
            if (g_split > 0 && g_split < G) HIPX(hipStreamWaitEvent(hs, pass_ev[lane], 0));
            StreamWait sw;
            sw.done = packed_done;
            sw.queries = pq;
            sw.stride = (uint32_t)G;
            sw.spin_limit = stream_spin_limit;
            sw.status = d_status + lane;
            HIPX(launch_horner(d_precomp + PC_A, PC_GROUP, G, Da, T, packed, packed_stride, d_rgsw + q0 * rgsw_stride,
                               rgsw_stride, out, d_hstate + q0 * T * 4 * N, 0, G, nq, nc, cc, hs, FusedPack{}, sw));
            // the next sub-batch's key digits and flags wait for the collapse and the pack
            HIPX(hipEventRecord(prep_ev[lane], ps_s));
            HIPX(hipStreamWaitEvent(s, prep_ev[lane], 0));
            continue;
        }
        if (shape == 0) {
            const DbPlan p = plan_for(nq, C);
            dbpass_prep(q0, nq, p, s, lane);
            mark_pass(s, true);
            dbpass_range(q0, nq, 0, G, p, s, lane);
            mark_pass(s, false);
            HIPX(launch_collapse_keys(keys, keys_stride, nq, d_kdig[lane], d_KS[lane], s));
            HIPX(launch_collapse(d_precomp, G, d_kdig[lane], d_KS[lane], nq, p2, part2_stride, s, collapse_variant));
            HIPX(launch_pack_finalize(d_interm + q0 * 2 * C, C, p2, part2_stride, d_precomp, 0, G, keys, keys_stride,
                                      packed, packed_stride, nq, nc, s));
            if (!defer_n) defer_lo = q0;
            defer_n += nq;
            continue;
        }
        // shape 2 keeps the collapse on s (two HBM-bound passes side by side
        // share the bandwidth worse than back to back); shape 3 runs it
        // beside; shape 4 runs it for every group on s before the first
        // database-pass chunk, so the chunks then have the device to
        // themselves (and the Horner stream) at their full grid
        // shape 5 is shape 3 with the first database-pass chunk waiting for
        // the first collapse chunk: the collapse stream then stays a chunk
        // ahead, so each chunk's pack and Horner steps wait for its database
        // pass only, and the last ones follow the last pass at once
        const hipStream_t cs = shape == 3 || shape == 5 ? cstream[lane] : s;
        const bool first = shape == 4, ahead = shape == 5;
        joined = true;
        const DbPlan p = chunk_plan(nq, gchunk * N);  // every chunk's plan has this one's kernel and digits
        if (cs != s) {
            HIPX(hipEventRecord(prep_ev[lane], s));  // the keys are in place
            HIPX(hipStreamWaitEvent(cs, prep_ev[lane], 0));
        }
        HIPX(launch_collapse_keys(keys, keys_stride, nq, d_kdig[lane], d_KS[lane], cs));
        if (first) {
            TraceRange trc("collapse of all %zu groups", G);
            HIPX(launch_collapse(d_precomp, G, d_kdig[lane], d_KS[lane], nq, p2, part2_stride, s, collapse_variant));
        }
        dbpass_prep(q0, nq, p, s, lane);
        auto collapse_chunk = [&](size_t k) {
            const size_t g_lo = chunks[k].first, ng = chunks[k].second - g_lo;
            HIPX(launch_collapse(d_precomp + g_lo * PC_GROUP, ng, d_kdig[lane], d_KS[lane], nq, p2 + g_lo * 2 * N,
                                 part2_stride, cs, collapse_variant));
            HIPX(hipEventRecord(coll_ev[lane][k % kChunkEvents], cs));
        };
        if (ahead) {
            collapse_chunk(0);
            HIPX(hipStreamWaitEvent(s, coll_ev[lane][0], 0));
        }
        for (size_t k = 0; k < chunks.size(); k++) {
            const size_t g_lo = chunks[k].first, g_hi = chunks[k].second, ng = g_hi - g_lo;
            TraceRange trk("chunk %zu: groups %zu..%zu (DB pass, collapse, pack, Horner)", k, g_lo, g_hi - 1);
            hipEvent_t dbe = chunk_ev[lane][k % kChunkEvents], ce = coll_ev[lane][k % kChunkEvents];
            mark_pass(s, k == 0);
            dbpass_range(q0, nq, g_lo, g_hi, chunk_plan_at(nq, k, ng, p), s, lane);
            mark_pass(s, false);
            // with the collapse on s its event covers the DB pass too (an
            // event record costs the stream a few microseconds of idle)
            if (cs != s) HIPX(hipEventRecord(dbe, s));
            if (ahead) {
                if (k + 1 < chunks.size()) collapse_chunk(k + 1);  // ce of chunk k was recorded a chunk earlier
            } else {
                if (!first)
                    HIPX(launch_collapse(d_precomp + g_lo * PC_GROUP, ng, d_kdig[lane], d_KS[lane], nq,
                                         p2 + g_lo * 2 * N, part2_stride, cs, collapse_variant));
                HIPX(hipEventRecord(ce, cs));
            }
            if (cs != s) HIPX(hipStreamWaitEvent(hs, dbe, 0));
            HIPX(hipStreamWaitEvent(hs, ce, 0));
            // the pack: fused into the one-workgroup Horner chains, or a kernel of its own
            const bool fuse = pack_wg == 2 && !use_pair(nq * T, true, status_read);
            const FusedPack fp{d_interm + q0 * 2 * C, C, p2, part2_stride, d_precomp, keys, keys_stride};
            if (fuse)
                ;
            else if (nq == 1 ? pack_wg != 0 : pack_wg == 1)
                HIPX(launch_pack_finalize_wg(d_interm + q0 * 2 * C, C, p2, part2_stride, d_precomp, g_lo, ng, keys,
                                             keys_stride, packed, packed_stride, nq, nc, hs));
            else
                HIPX(launch_pack_finalize(d_interm + q0 * 2 * C, C, p2, part2_stride, d_precomp, g_lo, ng, keys,
                                          keys_stride, packed, packed_stride, nq, nc, hs));
            horner_range(nq, out, g_lo, g_hi, hs, q0, lane, status_read, true, fuse ? &fp : nullptr);
        }
        if (cs != s) {  // the next batch's key digits overwrite the lane's scratch
            HIPX(hipEventRecord(prep_ev[lane], cs));
            HIPX(hipStreamWaitEvent(s, prep_ev[lane], 0));
        }
    }
    flush_deferred();
    if (joined) {
        HIPX(hipEventRecord(horner_done[lane], hs));
        HIPX(hipStreamWaitEvent(s, horner_done[lane], 0));
    }
}

void GpuServerCtx::run_step(const void* in, void* out, size_t count, size_t slot0, int lane, int packed,
                            hipStream_t s, const std::function<void()>& fn) {
    if (!graph_mode || !s) {
        fn();
        return;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIPX(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone) {  // inside the caller's own capture
        fn();
        return;
    }
    std::lock_guard<std::mutex> lk(graph_mu);
    for (const StepGraph& g : graphs)
        if (g.in == in && g.out == out && g.count == count && g.slot0 == slot0 && g.lane == lane &&
            g.packed == packed) {
            HIPX(hipGraphLaunch(g.exec, s));
            return;
        }
    // thread-local: other threads' allocations do not break this capture
    HIPX(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipGraph_t graph = nullptr;
    try {
        fn();
    } catch (...) {
        (void)hipStreamEndCapture(s, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    HIPX(hipStreamEndCapture(s, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPX(e);
    if (graphs.size() >= kMaxGraphs) {
        (void)hipDeviceSynchronize();  // the evicted graph may still run on another stream
        (void)hipGraphExecDestroy(graphs.front().exec);
        graphs.erase(graphs.begin());
    }
    graphs.push_back({in, out, count, slot0, lane, packed, exec});
    HIPX(hipGraphLaunch(exec, s));
}

void GpuServerCtx::drop_graphs() {
    std::lock_guard<std::mutex> lk(graph_mu);
    if (graphs.empty()) return;
    (void)hipDeviceSynchronize();  // no replay of a dropped graph may still be queued
    for (const StepGraph& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
}

void GpuServerCtx::enqueue_packed(const uint8_t* d_in, size_t count, uint8_t* d_out, hipStream_t s, size_t slot0,
                                  int lane, bool status_read) {
    uint64_t* dq = d_q + slot0 * qlen;
    uint64_t* dr = d_resp + slot0 * resp_u64s;
    HIPX(launch_query_unpack(d_in, count, pp.db_rows, dq, qlen, s));
    enqueue(dq, count, dr, s, slot0, lane, status_read);
    HIPX(launch_compress(dr, count * T, d_cmp + slot0 * T * 4 * N, d_out, nc, cc, s));
}

void GpuServerCtx::answer_packed_host(const uint8_t* in, size_t count, uint8_t* out) {
    hipStream_t s = stream;
    uint8_t* hq = reinterpret_cast<uint8_t*>(h_q);     // packed queries fit the flat staging
    uint8_t* hr = reinterpret_cast<uint8_t*>(h_resp);  // so do compressed responses
    std::memcpy(hq, in, count * qpack_bytes);
    order_before(s);
    clear_status(s, 0);
    HIPX(hipMemcpyAsync(d_qpack, hq, count * qpack_bytes, hipMemcpyHostToDevice, s));
    run_step(d_qpack, d_rcmp, count, 0, 0, 1, s, [&] { enqueue_packed(d_qpack, count, d_rcmp, s, 0, 0, true); });
    HIPX(hipMemcpyAsync(hr, d_rcmp, count * T * kCompressedCtBytes, hipMemcpyDeviceToHost, s));
    order_after(s);
    HIPX(hipStreamSynchronize(s));
    check_status(s, 0);
    std::memcpy(out, hr, count * T * kCompressedCtBytes);
}

void GpuServerCtx::check_status(hipStream_t s, int lane) {
    uint32_t st = 0;
    HIPX(hipMemcpyAsync(&st, d_status + lane, 4, hipMemcpyDeviceToHost, s));
    HIPX(hipStreamSynchronize(s));
    if (st) {
        HIPX(hipMemsetAsync(d_status + lane, 0, 4, s));
        HIPX(hipStreamSynchronize(s));
        throw std::runtime_error(
            "a Horner chain exceeded its wait bound (pair exchange or streamed database pass); responses of this "
            "call are invalid");
    }
}

void GpuServerCtx::clear_status(hipStream_t s, int lane) {
    // a kernel, not a memset: the call may be captured
    HIPX(launch_fill_u32(d_status + lane, 1, 0, s));
}

void GpuServerCtx::order_before(hipStream_t s) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIPX(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone) return;  // a caller's capture orders its own work
    if (order_stream_set && order_stream != s) HIPX(hipStreamWaitEvent(s, order_ev, 0));
}

void GpuServerCtx::order_after(hipStream_t s) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIPX(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone) return;
    HIPX(hipEventRecord(order_ev, s));
    order_stream = s;
    order_stream_set = true;
}

void GpuServerCtx::answer_host(size_t count, uint64_t* out, hipStream_t s, size_t slot0) {
    if (!s) s = stream;
    uint64_t* hq = h_q + slot0 * qlen;
    uint64_t* hr = h_resp + slot0 * resp_u64s;
    uint64_t* dq = d_q + slot0 * qlen;
    uint64_t* dr = d_resp + slot0 * resp_u64s;
    TraceRange tr("host-staged answer: %zu queries", count);
    order_before(s);
    clear_status(s, 0);
    HIPX(hipMemcpyAsync(dq, hq, count * qlen * 8, hipMemcpyHostToDevice, s));
    run_step(dq, dr, count, slot0, 0, 0, s, [&] { enqueue(dq, count, dr, s, slot0, 0, true); });
    HIPX(hipMemcpyAsync(hr, dr, count * resp_u64s * 8, hipMemcpyDeviceToHost, s));
    order_after(s);
    HIPX(hipStreamSynchronize(s));
    check_status(s, 0);
    if (out) std::memcpy(out, hr, count * resp_u64s * 8);
}

GpuServerCaps gpu_server_caps(const GpuServerCtx* ctx) {
    GpuServerCaps c{};
    c.max_batch = ctx->cfg.max_batch;
    c.resident_bytes = ctx->resident_bytes;
    c.num_cts = ctx->T;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) c.device_free_bytes = fr;
    return c;
}

namespace {
void flatten_query(const QueryMessage& m, size_t R, uint64_t* w) {
    if (m.lwe.b_limb0.size() != R || m.lwe.b_limb1.size() != R)
        throw std::invalid_argument("gpu_answer_batch: query length differs from db_rows");
    std::memcpy(w, m.lwe.b_limb0.data(), R * 8);
    std::memcpy(w + R, m.lwe.b_limb1.data(), R * 8);
    w += 2 * R;
    auto parts = [&](const std::vector<RnsPoly>& v, bool want_ntt) {
        if (v.size() != (size_t)D_EFF) throw std::invalid_argument("gpu_answer_batch: wrong digit count");
        for (const RnsPoly& p0 : v) {
            RnsPoly p = p0;
            if (want_ntt) p.to_ntt(); else p.to_coeff();
            std::memcpy(w, p.limbs[0].data(), N * 8);
            std::memcpy(w + N, p.limbs[1].data(), N * 8);
            w += 2 * N;
        }
    };
    parts(m.ksk_5.b_parts, false);
    parts(m.ksk_neg1.b_parts, false);
    parts(m.rgsw.top.a_parts, true);
    parts(m.rgsw.top.b_parts, true);
    parts(m.rgsw.bottom.a_parts, true);
    parts(m.rgsw.bottom.b_parts, true);
}
}  // namespace

std::vector<std::vector<RlweCt>> gpu_answer_batch(GpuServerCtx* ctx, const QueryMessage* queries, size_t count) {
    if (!ctx) throw std::invalid_argument("gpu_answer_batch: null ctx");
    if (count == 0) return {};
    if (count > ctx->cfg.max_batch) throw std::invalid_argument("gpu_answer_batch: count exceeds max_batch");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->loop_active) throw std::runtime_error("gpu_answer_batch: a serving loop owns this server");
    for (size_t i = 0; i < count; i++) flatten_query(queries[i], ctx->pp.db_rows, ctx->h_q + i * ctx->qlen);
    std::vector<uint64_t> flat(count * ctx->resp_u64s);
    ctx->answer_host(count, flat.data());
    std::vector<std::vector<RlweCt>> out(count, std::vector<RlweCt>(ctx->T));
    const uint64_t* r = flat.data();
    for (auto& resp : out)
        for (auto& ct : resp)
            for (RnsPoly* p : {&ct.a, &ct.b}) {
                std::memcpy(p->limbs[0].data(), r, N * 8);
                std::memcpy(p->limbs[1].data(), r + N, N * 8);
                p->is_ntt = true;
                r += 2 * N;
            }
    return out;
}

std::vector<RlweCt> gpu_answer(GpuServerCtx* ctx, const QueryMessage& qry) {
    return gpu_answer_batch(ctx, &qry, 1)[0];
}

void gpu_free_server(GpuServerCtx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->order_ev) {
        (void)hipEventSynchronize(ctx->order_ev);  // work of a caller stream
        (void)hipEventDestroy(ctx->order_ev);
    }
    ctx->drop_graphs();
    dfree(ctx->d_tw); dfree(ctx->d_db); dfree(ctx->d_precomp);
    dfree(ctx->d_q); dfree(ctx->d_resp); dfree(ctx->d_bq); dfree(ctx->d_keys); dfree(ctx->d_rgsw);
    dfree(ctx->d_part); dfree(ctx->d_interm); dfree(ctx->d_part2); dfree(ctx->d_packed_b);
    dfree(ctx->d_hxbuf); dfree(ctx->d_hflags); dfree(ctx->d_status); dfree(ctx->d_hstate);
    for (auto& f : ctx->d_sflags) dfree(f);
    {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        if ((size_t)ctx->device < kMaxDevices && g_streaming[ctx->device] == ctx) g_streaming[ctx->device] = nullptr;
    }
    dfree(ctx->d_qpack); dfree(ctx->d_cmp); dfree(ctx->d_rcmp);
    for (int l = 0; l < GpuServerCtx::kLanes; l++) {
        if (ctx->hstream[l]) (void)hipStreamSynchronize(ctx->hstream[l]);
        if (ctx->cstream[l]) (void)hipStreamSynchronize(ctx->cstream[l]);
        for (auto& e : ctx->chunk_ev[l])
            if (e) (void)hipEventDestroy(e);
        for (auto& e : ctx->coll_ev[l])
            if (e) (void)hipEventDestroy(e);
        if (ctx->prep_ev[l]) (void)hipEventDestroy(ctx->prep_ev[l]);
        if (ctx->pass_ev[l]) (void)hipEventDestroy(ctx->pass_ev[l]);
        if (ctx->horner_done[l]) (void)hipEventDestroy(ctx->horner_done[l]);
        if (ctx->hstream[l]) (void)hipStreamDestroy(ctx->hstream[l]);
        if (ctx->cstream[l]) (void)hipStreamDestroy(ctx->cstream[l]);
    }
    for (hipEvent_t e : ctx->pass_marks) (void)hipEventDestroy(e);
    for (int l = 0; l < GpuServerCtx::kLanes; l++) {
        dfree(ctx->d_dig[l]); dfree(ctx->d_S[l]); dfree(ctx->d_kdig[l]); dfree(ctx->d_KS[l]);
    }
    if (ctx->h_q) (void)hipHostFree(ctx->h_q);
    if (ctx->h_resp) (void)hipHostFree(ctx->h_resp);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // namespace inspire
