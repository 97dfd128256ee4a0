// Server half of the C ABI (reference src/capi.cu:51-132) plus the
// device-resident batch entry point and the inspection entry points.
// Exceptions never cross the boundary: failures become NULL / nonzero.
#include <cstdio>
#include <cstring>
#include <string>
#include <memory>
#include <vector>

#include "capi.h"
#include "inspire_debug.h"
#include "server_ctx.h"

using namespace inspire;


namespace {
size_t blk() { return (size_t)D_EFF * 2 * N; }

// copy `count` query views into the context's pinned flat staging
void stage_views(GpuServerCtx* ctx, const ipir_query_view* qs, size_t count) {
    const size_t R = ctx->pp.db_rows, b = blk();
    for (size_t i = 0; i < count; i++) {
        uint64_t* w = ctx->h_q + i * ctx->qlen;
        std::memcpy(w, qs[i].lwe_b_limb0, R * 8);
        std::memcpy(w + R, qs[i].lwe_b_limb1, R * 8);
        std::memcpy(w + 2 * R, qs[i].ksk5_b, b * 8);
        std::memcpy(w + 2 * R + b, qs[i].kskneg1_b, b * 8);
        std::memcpy(w + 2 * R + 2 * b, qs[i].rgsw, 4 * b * 8);
    }
}
}  // namespace

extern "C" {

ipir_server* ipir_server_create(size_t n_entries, size_t entry_bytes, size_t db_rows_or_zero, size_t max_batch,
                                const uint16_t* db, const uint8_t* crs_seed_or_null, ipir_params* out) {
    if (!db || !out || max_batch == 0) return nullptr;
    try {
        std::unique_ptr<ipir_server> srv(new ipir_server);
        const size_t rows = db_rows_or_zero ? db_rows_or_zero : 32768;
        srv->pp = crs_seed_or_null ? setup_seeded(n_entries, entry_bytes, rows, crs_seed_or_null)
                                   : setup(n_entries, entry_bytes, rows);
        PreprocessData pd = gpu_preprocess(srv->pp, db);
        GpuServerConfig cfg;
        cfg.max_batch = max_batch;
        try {
            srv->ctx = gpu_setup_server(srv->pp, pd, cfg);
        } catch (...) {
            dfree(pd.d_db_col);
            dfree(pd.d_precomp_arena);
            throw;
        }
        out->n_entries = srv->pp.N_entries;
        out->entry_bytes = srv->pp.w;
        out->db_rows = srv->pp.db_rows;
        out->db_cols = srv->pp.db_cols;
        out->n_packed = srv->pp.n_packed;
        out->interp_d = srv->pp.D;
        out->num_cts = srv->pp.num_cts;
        out->ring_n = N;
        out->d_eff = (size_t)D_EFF;
        std::memcpy(out->seed, srv->pp.seed.data(), SEED_BYTES);
        return srv.release();
    } catch (...) {
        return nullptr;
    }
}

void ipir_server_caps(const ipir_server* srv, ipir_caps* out) {
    if (!srv || !out) return;
    const GpuServerCaps c = gpu_server_caps(srv->ctx);
    out->max_batch = c.max_batch;
    out->resident_bytes = c.resident_bytes;
    out->device_free_bytes = c.device_free_bytes;
    out->num_cts = c.num_cts;
}

int ipir_answer_batch(ipir_server* srv, const ipir_query_view* queries, size_t count, uint64_t* out_resp) {
    if (!srv || !queries || !out_resp) return 1;
    if (count == 0) return 0;
    if (count > srv->ctx->cfg.max_batch) return 2;
    for (size_t i = 0; i < count; i++) {
        const ipir_query_view& v = queries[i];
        if (!v.lwe_b_limb0 || !v.lwe_b_limb1 || !v.ksk5_b || !v.kskneg1_b || !v.rgsw) return 3;
    }
    try {
        std::lock_guard<std::mutex> lk(srv->ctx->mu);
        if (srv->ctx->loop_active) return 5;
        stage_views(srv->ctx, queries, count);
        srv->ctx->answer_host(count, out_resp);
        return 0;
    } catch (...) {
        return 4;
    }
}

int ipir_answer_batch_device(ipir_server* srv, const uint64_t* d_queries, size_t count, uint64_t* d_out_resp,
                             void* hip_stream) {
    if (!srv || !d_queries || !d_out_resp) return 1;
    if (count == 0) return 0;
    if (count > srv->ctx->cfg.max_batch) return 2;
    try {
        GpuServerCtx* c = srv->ctx;
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->loop_active) return 5;
        hipStream_t s = hip_stream ? static_cast<hipStream_t>(hip_stream) : c->stream;
        c->order_before(s);
        // nobody reads a status word after an asynchronous call, so its
        // answers never take the pair Horner kernel (enqueue's default)
        c->run_step(d_queries, d_out_resp, count, 0, 0, 0, s, [&] { c->enqueue(d_queries, count, d_out_resp, s); });
        c->order_after(s);
        return 0;
    } catch (...) {
        return 4;
    }
}

int ipir_device_status(ipir_server* srv) {
    if (!srv) return 1;
    try {
        GpuServerCtx* c = srv->ctx;
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->loop_active) return 5;
        hipStream_t s = c->stream;
        c->order_before(s);  // the last call's work, whatever stream it ran on
        c->check_status(s, 0);
        return 0;
    } catch (...) {
        return 4;
    }
}

int ipir_answer_batch_packed(ipir_server* srv, const uint8_t* packed_queries, size_t count,
                             uint8_t* out_compressed) {
    if (!srv || !packed_queries || !out_compressed) return 1;
    if (count == 0) return 0;
    if (count > srv->ctx->cfg.max_batch) return 2;
    try {
        std::lock_guard<std::mutex> lk(srv->ctx->mu);
        if (srv->ctx->loop_active) return 5;
        srv->ctx->answer_packed_host(packed_queries, count, out_compressed);
        return 0;
    } catch (...) {
        return 4;
    }
}

int ipir_answer_batch_packed_device(ipir_server* srv, const uint8_t* d_packed_queries, size_t count,
                                    uint8_t* d_out_compressed, void* hip_stream) {
    if (!srv || !d_packed_queries || !d_out_compressed) return 1;
    if (count == 0) return 0;
    if (count > srv->ctx->cfg.max_batch) return 2;
    try {
        GpuServerCtx* c = srv->ctx;
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->loop_active) return 5;
        hipStream_t s = hip_stream ? static_cast<hipStream_t>(hip_stream) : c->stream;
        c->order_before(s);
        c->run_step(d_packed_queries, d_out_compressed, count, 0, 0, 1, s,
                    [&] { c->enqueue_packed(d_packed_queries, count, d_out_compressed, s); });
        c->order_after(s);
        return 0;
    } catch (...) {
        return 4;
    }
}

int ipir_server_footprint(size_t n_entries, size_t entry_bytes, size_t db_rows_or_zero, ipir_footprint* out) {
    if (!out) return 1;
    try {
        const PublicParams pp = setup(n_entries, entry_bytes, db_rows_or_zero ? db_rows_or_zero : 32768);
        const ServerFootprint f = server_footprint(pp, default_group_chunk(pp.n_packed));
        out->resident_fixed = f.fixed;
        out->per_slot = f.per_slot;
        out->preprocess_peak = f.preprocess_peak;
        return 0;
    } catch (...) {
        return 1;
    }
}

size_t ipir_fit_max_batch(size_t n_entries, size_t entry_bytes, size_t db_rows_or_zero, size_t reserve_bytes,
                          size_t cap) {
    ipir_footprint f{};
    if (ipir_server_footprint(n_entries, entry_bytes, db_rows_or_zero, &f)) return 0;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
    if (fr < reserve_bytes || fr - reserve_bytes < f.preprocess_peak || fr - reserve_bytes < f.resident_fixed) return 0;
    const size_t b = (fr - reserve_bytes - f.resident_fixed) / f.per_slot;
    return b < cap ? b : cap;
}

void ipir_server_destroy(ipir_server* srv) {
    if (!srv) return;
    gpu_free_server(srv->ctx);
    delete srv;
}

// ---------------------------------------------------------------- inspection
size_t ipir_dbg_precomp_group_words(void) { return ipx::REF_GROUP_WORDS; }

int ipir_dbg_precomp_group(const ipir_server* srv, size_t group, uint32_t* out) {
    if (!srv || !out || group >= srv->ctx->G) return 1;
    using ipx::DD_K; using ipx::DD_SZ; using ipx::HALF; using ipx::LOGN; using ipx::NSTEPS; using ipx::PC_A;
    using ipx::PC_DD; using ipx::PC_DF; using ipx::PC_GROUP; using ipx::RD_SZ;
    constexpr uint32_t N = ipx::N;
    constexpr int D_EFF = ipx::D_EFF;
    const uint32_t* blk_dev = srv->ctx->d_precomp + group * PC_GROUP;
    std::vector<uint8_t> dd(DD_SZ * 4);
    uint32_t* Dp = out + 2 * N;
    uint32_t* Dm = Dp + RD_SZ;
    uint32_t* Df = Dm + RD_SZ;
    if (hipMemcpy(out, blk_dev + PC_A, 2 * N * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(Df, blk_dev + PC_DF, 2 * D_EFF * N * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(dd.data(), blk_dev + PC_DD, dd.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return 4;
    // back from the dot-product form: D+/-[c][s][j][i] = Dd[c][j][p+/-(s, i)][i]
    auto br = [](uint32_t x) {
        uint32_t r = 0;
        for (uint32_t b = 0; b < LOGN; b++) r |= ((x >> b) & 1u) << (LOGN - 1 - b);
        return r;
    };
    auto dd_at = [&](uint32_t c, uint32_t j, uint32_t p, uint32_t i) {
        const size_t k = (size_t)j * N + p;
        uint32_t v = 0;
        for (uint32_t a = 0; a < 3; a++)
            v |= (dd[((((size_t)c * 3 + a) * (DD_K / 16) + k / 16) * N + i) * 16 + k % 16] ^ 0x80u) << (8 * a);
        const uint32_t top = dd[ipx::DD_TOP_BYTES + (((size_t)c * (DD_K / 16) + k / 16) * N + i) * 8 + k % 8];
        return v | ((k % 16 < 8 ? top & 15u : top >> 4) << 24);
    };
    uint32_t t = 1;  // 5^(N/2-2-s) mod 2N, starting at s = 0
    for (uint32_t e = 0; e < HALF - 2; e++) t = (t * 5) & (2 * N - 1);
    for (uint32_t s = 0; s < NSTEPS; s++) {
        for (uint32_t i = 0; i < N; i++) {
            const uint32_t u = (t * (2 * br(i) + 1)) & (2 * N - 1);
            const uint32_t pf = br((u - 1) >> 1), pc = N - 1 - pf;
            for (uint32_t c = 0; c < 2; c++)
                for (uint32_t j = 0; j < (uint32_t)D_EFF; j++) {
                    const size_t o = (((size_t)c * NSTEPS + s) * D_EFF + j) * N + i;
                    Dp[o] = dd_at(c, j, pf, i);
                    Dm[o] = dd_at(c, j, pc, i);
                }
        }
        t = (t * 3277u) & (2 * N - 1);  // times 5^-1
    }
    return 0;
}

int ipir_dbg_db_block(const ipir_server* srv, size_t row0, size_t nrows, size_t col0, size_t ncols,
                      uint16_t* out) {
    if (!srv || !out || row0 + nrows > srv->pp.db_rows || col0 + ncols > srv->pp.db_cols) return 1;
    if (nrows == 0 || ncols == 0) return 0;
    const size_t C = srv->pp.db_cols;
    // resident layout: plane p of (r, c) at byte ((r/16)*2 + p)*C*16 + c*16 + r%16;
    // one 16-row block of both planes at a time, columns [col0, col0+ncols) of each
    const size_t kb0 = row0 / 16, kb1 = (row0 + nrows + 15) / 16;
    std::vector<uint8_t> blk(2 * ncols * 16);
    for (size_t kb = kb0; kb < kb1; kb++) {
        for (size_t p = 0; p < 2; p++)
            if (hipMemcpy(blk.data() + p * ncols * 16, srv->ctx->db_planes() + ((kb * 2 + p) * C + col0) * 16,
                          ncols * 16, hipMemcpyDeviceToHost) != hipSuccess)
                return 4;
        for (size_t r = kb * 16 > row0 ? kb * 16 : row0; r < (kb + 1) * 16 && r < row0 + nrows; r++)
            for (size_t c = 0; c < ncols; c++) {
                const size_t base = c * 16 + r % 16;
                out[(r - row0) * ncols + c] = (uint16_t)(blk[base] | (blk[base + ncols * 16] << 8));
            }
    }
    return 0;
}

int ipir_dbg_db_rows(const ipir_server* srv, size_t row0, size_t nrows, uint16_t* out) {
    return srv ? ipir_dbg_db_block(srv, row0, nrows, 0, srv->pp.db_cols, out) : 1;
}

int ipir_dbg_db_pass(ipir_server* srv, const uint64_t* b, size_t nq, int path, uint32_t* interm) {
    if (!srv || !b || !interm || nq == 0 || nq > srv->ctx->cfg.max_batch || path < -1 || path > 1) return 1;
    GpuServerCtx* c = srv->ctx;
    const size_t R = c->pp.db_rows, C = c->pp.db_cols;
    for (size_t i = 0; i < nq * 2 * R; i++)
        if (b[i] >= ((i / R) % 2 ? ipx::Q1 : ipx::Q0)) return 1;
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->loop_active) return 5;
        c->order_before(c->stream);
        std::vector<double> bd(nq * 2 * R);
        for (size_t i = 0; i < bd.size(); i++) bd[i] = (double)b[i];
        hip_check(hipMemcpy(c->d_bq, bd.data(), bd.size() * 8, hipMemcpyHostToDevice), "dbg db pass upload");
        c->stage_dbpass(nq, c->stream, 0, 0, path);
        c->order_after(c->stream);
        hip_check(hipStreamSynchronize(c->stream), "dbg db pass");
        hip_check(hipMemcpy(interm, c->d_interm, nq * 2 * C * 4, hipMemcpyDeviceToHost), "dbg db pass download");
        return 0;
    } catch (...) {
        return 4;
    }
}

int ipir_dbg_time_dbpass(ipir_server* srv, size_t nq, int iters, double* ms_per_pass) {
    if (!srv || nq == 0 || nq > srv->ctx->cfg.max_batch || nq > ipx::kMfmaMaxQueries || iters <= 0 || !ms_per_pass)
        return 1;
    GpuServerCtx* c = srv->ctx;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->loop_active) {
            rc = 5;
            throw 0;
        }
        c->order_before(c->stream);
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        // the query digits once; then the pass's own launches (the kernel and,
        // for a plan with row slices, the slice reduction)
        const ipx::DbPlan p = c->plan_for(nq, c->pp.db_cols);
        c->dbpass_prep(0, nq, p, c->stream, 0);
        c->dbpass_range(0, nq, 0, c->G, p, c->stream, 0);  // warm-up
        hip_check(hipEventRecord(e0, c->stream), "record");
        for (int i = 0; i < iters; i++) c->dbpass_range(0, nq, 0, c->G, p, c->stream, 0);
        hip_check(hipEventRecord(e1, c->stream), "record");
        hip_check(hipEventSynchronize(e1), "sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        *ms_per_pass = ms / iters;
        c->order_after(c->stream);
    } catch (...) {
        if (!rc) rc = 4;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

int ipir_dbg_dbpass_kernel(const ipir_server* srv, size_t nq, char* out, size_t cap) {
    if (!srv || !out || cap == 0 || nq == 0 || nq > ipx::kMfmaMaxQueries) return 1;
    const GpuServerCtx* c = srv->ctx;
    const ipx::DbPlan p = c->plan_for(nq, c->pp.db_cols);
    const std::string s = ipx::dbplan_kernel(p) + " slices=" + std::to_string(p.slices);
    std::snprintf(out, cap, "%s", s.c_str());
    return 0;
}

int ipir_dbg_tuning(const ipir_server* srv, char* out, size_t cap) {
    if (!srv || !out || cap == 0) return 1;
    const GpuServerCtx* c = srv->ctx;
    const size_t nq = c->cfg.max_batch < ipx::kMfmaMaxQueries ? c->cfg.max_batch : ipx::kMfmaMaxQueries;
    auto plan = [](const ipx::DbPlan& p) {
        return ipx::dbplan_kernel(p) + " slices=" + std::to_string(p.slices);
    };
    std::string s = "batch1: " + plan(c->plan_for(1, c->pp.db_cols)) + "; batch" + std::to_string(nq) + ": " +
                    plan(c->plan_for(nq, c->pp.db_cols)) + " (form " + std::to_string(c->mfma_form) + ", target " +
                    std::to_string(c->mfma_target) + ")";
    if (c->tuned16)
        s += "; batch16: " + plan(c->plan_for(16, c->pp.db_cols)) + " (form " + std::to_string(c->mfma16_form) +
             ", target " + std::to_string(c->mfma16_target) + ")";
    if (c->G / c->gchunk > 1)
        s += "; chunks of " + std::to_string(c->gchunk) + " groups: batch1 " + plan(c->chunk_plan(1, c->gchunk * ipx::N)) +
             ", batch" + std::to_string(nq) + " " + plan(c->chunk_plan(nq, c->gchunk * ipx::N));
    if (c->chunks.size() > c->G / c->gchunk) {
        s += c->tail_split ? " (lowest split in 1/2, 1/4, 1/4" : " (";
        s += c->head_split ? std::string(c->tail_split ? "; " : "") + "highest split in 1/4, 1/4, 1/2)" : ")";
    }
    s += "; collapse variant " + std::to_string(c->collapse_variant) + "; pipeline shapes " +
         std::to_string(c->pipe_single) + "/" + std::to_string(c->pipe_batched);
    if (c->pipe_single >= 6 || c->pipe_batched >= 6) {  // streamed shapes
        auto sp = [&](size_t n, uint32_t sl, uint32_t split) {
            const ipx::DbPlan p = c->stream_plan(n, sl, (n == 1 ? c->pipe_single : c->pipe_batched) == 8);
            return p.kind == ipx::DbPlan::MFMA
                       ? plan(p) + (split ? " split " + std::to_string(split) + "/8" : std::string())
                       : std::string("none");
        };
        s += " (streamed: batch1 " + sp(1, c->stream_slices1, c->stream_split1) + ", batch" + std::to_string(nq) +
             " " + sp(nq, c->stream_slices, c->stream_split) + "; chunked fallback " +
             std::to_string(c->pipe_single_chunked) + "/" + std::to_string(c->pipe_batched_chunked) + ")";
    }
    std::snprintf(out, cap, "%s", s.c_str());
    return 0;
}

int ipir_dbg_answer_taps(ipir_server* srv, const uint64_t* query, uint32_t* interm, uint32_t* packed_b,
                         uint64_t* resp) {
    if (!srv || !query) return 1;
    GpuServerCtx* c = srv->ctx;
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->loop_active) return 5;
        std::memcpy(c->h_q, query, c->qlen * 8);
        std::vector<uint64_t> r(c->resp_u64s);
        c->answer_host(1, r.data());
        if (resp) std::memcpy(resp, r.data(), r.size() * 8);
        if (interm && c->last_part_slices > 1) {
            // a streamed pass left row-slice partials (the pack sums them):
            // the database-pass result is their sum, as the chunked passes write it
            const size_t C = c->pp.db_cols;
            hip_check(ipx::launch_reduce_slices(c->d_part, c->part_stride, c->last_part_slices, C, c->d_interm, 1,
                                                0, C, c->stream),
                      "taps: slice sums");
            hip_check(hipStreamSynchronize(c->stream), "taps");
        }
        if (interm && hipMemcpy(interm, c->d_interm, 2 * c->pp.db_cols * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return 4;
        if (packed_b && hipMemcpy(packed_b, c->d_packed_b, c->packed_stride * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return 4;
        return 0;
    } catch (...) {
        return 4;
    }
}

namespace {
struct Standalone {
    uint32_t* tw = nullptr;
    ipx::NttConst nc{};
    ipx::CrtConst cc{};
};
Standalone* standalone() {
    static Standalone* s = [] {
        auto* x = new Standalone;
        // reuse the context table builder through a throwaway context-free path
        std::vector<uint32_t> h(2 * 4 * (size_t)N);
        for (int l = 0; l < 2; l++) {
            const NttTables& t = ntt_tables(l);
            for (size_t i = 0; i < N; i++) {
                h[(l * 4 + 0) * N + i] = (uint32_t)t.fwd_twiddles[i];
                h[(l * 4 + 1) * N + i] = (uint32_t)t.fwd_primes[i];
                h[(l * 4 + 2) * N + i] = (uint32_t)t.inv_twiddles[i];
                h[(l * 4 + 3) * N + i] = (uint32_t)t.inv_primes[i];
            }
            x->nc.inv_n[l] = (uint32_t)t.inv_n;
            x->nc.inv_n_p[l] = (uint32_t)t.inv_n_prime;
        }
        if (hipMalloc(&x->tw, h.size() * 4) != hipSuccess) return (Standalone*)nullptr;
        if (hipMemcpy(x->tw, h.data(), h.size() * 4, hipMemcpyHostToDevice) != hipSuccess) return (Standalone*)nullptr;
        x->nc.tw = x->tw;
        x->cc.q1inv_q0 = (uint32_t)mod_inv(Q1 % Q0, Q0);
        x->cc.q0inv_q1 = (uint32_t)mod_inv(Q0 % Q1, Q1);
        return x;
    }();
    return s;
}
}  // namespace

int ipir_dbg_ntt(const uint32_t* in, uint32_t* out, size_t npoly, int inverse) {
    if (!in || !out || npoly == 0) return 1;
    Standalone* st = standalone();
    if (!st) return 4;
    uint32_t* d = nullptr;
    const size_t bytes = npoly * 2 * N * 4;
    if (hipMalloc(&d, bytes) != hipSuccess) return 4;
    int rc = 0;
    if (hipMemcpy(d, in, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        ipx::launch_ntt_rns(d, npoly, st->nc, inverse != 0, nullptr) != hipSuccess ||
        hipMemcpy(out, d, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        rc = 4;
    dfree(d);
    return rc;
}

int ipir_dbg_ntt_bench(size_t npoly, int inverse, int iters, double* ms_per_call) {
    if (npoly == 0 || iters <= 0 || !ms_per_call) return 1;
    Standalone* st = standalone();
    if (!st) return 4;
    uint32_t* d = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipMalloc(&d, npoly * 2 * N * 4) || hipMemset(d, 0, npoly * 2 * N * 4) || hipEventCreate(&e0) ||
        hipEventCreate(&e1))
        rc = 4;
    if (!rc && ipx::launch_ntt_rns(d, npoly, st->nc, inverse != 0, nullptr)) rc = 4;  // warm-up
    if (!rc) {
        (void)hipEventRecord(e0, nullptr);
        for (int i = 0; i < iters && !rc; i++)
            if (ipx::launch_ntt_rns(d, npoly, st->nc, inverse != 0, nullptr)) rc = 4;
        (void)hipEventRecord(e1, nullptr);
        float ms = 0;
        if (!rc && (hipEventSynchronize(e1) || hipEventElapsedTime(&ms, e0, e1))) rc = 4;
        *ms_per_call = ms / iters;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(d);
    return rc;
}

int ipir_dbg_copy_bench(size_t bytes, int iters, double* ms_per_copy) {
    if (bytes == 0 || bytes % 16 || iters <= 0 || !ms_per_copy) return 1;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipMalloc(&a, bytes) || hipMalloc(&b, bytes) || hipMemset(a, 1, bytes) || hipEventCreate(&e0) ||
        hipEventCreate(&e1))
        rc = 4;
    if (!rc && ipx::launch_copy16(a, b, bytes, nullptr)) rc = 4;  // warm-up
    if (!rc) {
        (void)hipEventRecord(e0, nullptr);
        for (int i = 0; i < iters && !rc; i++)
            if (ipx::launch_copy16(a, b, bytes, nullptr)) rc = 4;
        (void)hipEventRecord(e1, nullptr);
        float ms = 0;
        if (!rc && (hipEventSynchronize(e1) || hipEventElapsedTime(&ms, e0, e1))) rc = 4;
        *ms_per_copy = ms / iters;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(a);
    dfree(b);
    return rc;
}

int ipir_dbg_horner(const uint64_t* cts, size_t L, const uint64_t* rgsw, uint64_t* out) {
    return ipir_dbg_horner_variant(cts, L, rgsw, out, 0);
}

int ipir_dbg_set_group_chunk(ipir_server* srv, size_t groups) {
    if (!srv || groups == 0 || groups > srv->ctx->G || srv->ctx->G % groups) return 1;
    GpuServerCtx* c = srv->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->loop_active) return 5;
    // the partial-sum buffer was sized for the setup-time chunk width
    const size_t ncols = groups * N;
    for (size_t nq = 1; nq <= ipx::kMfmaMaxQueries; nq++) {
        const ipx::DbPlan x = c->chunk_plan(nq, ncols);
        if (x.slices > 1 && x.slices * 2 * ncols > c->part_stride) return 1;
    }
    c->gchunk = groups;
    c->set_chunks();
    c->drop_graphs();  // their kernels carry the old plans
    return 0;
}

int ipir_dbg_set_tail_split(ipir_server* srv, int on) {
    if (!srv || on < 0 || on > 3) return 1;
    GpuServerCtx* c = srv->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->loop_active) return 5;
    c->tail_split = on & 1;
    c->head_split = (on >> 1) & 1;
    c->set_chunks();
    c->drop_graphs();  // their kernels carry the old plans
    return 0;
}

int ipir_dbg_set_collapse_variant(ipir_server* srv, int variant) {
    if (!srv || variant < 0 || variant >= ipx::kCollapseVariants) return 1;
    std::lock_guard<std::mutex> lk(srv->ctx->mu);
    if (srv->ctx->loop_active) return 5;
    srv->ctx->collapse_variant = variant;
    srv->ctx->drop_graphs();  // their kernels carry the old plans
    return 0;
}

int ipir_dbg_get_collapse_variant(const ipir_server* srv) { return srv ? srv->ctx->collapse_variant : -1; }

int ipir_dbg_set_chunk_slices(ipir_server* srv, size_t nq, uint32_t slices) {
    if (!srv || nq < 1 || nq > ipx::kMfmaMaxQueries || slices > 64 || (slices & (slices - 1))) return 1;
    GpuServerCtx* c = srv->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->loop_active) return 5;
    if (nq == 1) {
        c->chunk_slices1 = slices;
        c->drop_graphs();  // their kernels carry the old plans
        return 0;
    }
    c->chunk_slices = slices;
    c->chunk_param = slices ? c->plan_for(nq, c->pp.db_cols).param : 0;
    c->drop_graphs();  // their kernels carry the old plans
    return 0;
}

int ipir_dbg_set_late_slices(ipir_server* srv, size_t nq, size_t from_chunk, uint32_t slices) {
    if (!srv || nq < 1 || nq > ipx::kMfmaMaxQueries || slices > 64 || (slices & (slices - 1))) return 1;
    GpuServerCtx* c = srv->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->loop_active) return 5;
    if (nq == 1) {
        c->late_from1 = (uint32_t)from_chunk;
        c->late_slices1 = slices;
    } else {
        c->late_from = (uint32_t)from_chunk;
        c->late_slices = slices;
    }
    c->drop_graphs();  // their kernels carry the old plans
    return 0;
}

int ipir_dbg_set_mfma_depth(ipir_server* srv, uint32_t depth, int form) {
    if (!srv || form < 0 || form >= ipx::kMfmaForms) return 1;
    // a depth the 32-query kernel of this form is built with
    if (ipx::dbplan_form(srv->ctx->pp.db_rows, srv->ctx->pp.db_cols, ipx::kMfmaMaxQueries, 1, depth, form).depth !=
        depth)
        return 1;
    for (uint32_t d : ipx::kMfmaDepths)
        if (d == depth) {
            std::lock_guard<std::mutex> lk(srv->ctx->mu);
            if (srv->ctx->loop_active) return 5;
            srv->ctx->mfma_depth = depth;
            srv->ctx->mfma_form = form;
            srv->ctx->tuned16 = false;  // every 9..32-query batch takes the set form
            srv->ctx->drop_graphs();  // their kernels carry the old plans
            return 0;
        }
    return 1;
}

int ipir_dbg_set_stream_slices(ipir_server* srv, size_t nq_class, uint32_t slices) {
    if (!srv || slices == 0 || slices > 16 || (slices & (slices - 1))) return 1;
    std::lock_guard<std::mutex> lk(srv->ctx->mu);
    if (srv->ctx->loop_active) return 5;
    (nq_class == 1 ? srv->ctx->stream_slices1 : srv->ctx->stream_slices) = slices;
    srv->ctx->drop_graphs();
    return 0;
}

int ipir_dbg_set_stream_split(ipir_server* srv, size_t nq_class, uint32_t eighths) {
    if (!srv || eighths > 7) return 1;
    std::lock_guard<std::mutex> lk(srv->ctx->mu);
    if (srv->ctx->loop_active) return 5;
    (nq_class == 1 ? srv->ctx->stream_split1 : srv->ctx->stream_split) = eighths;
    srv->ctx->drop_graphs();
    return 0;
}

int ipir_dbg_stream_ok(const ipir_server* srv) { return srv && srv->ctx->stream_ok ? 1 : 0; }

int ipir_dbg_stream_parts(ipir_server* srv, size_t nq, uint32_t slices, double* out_ms) {
    if (!srv || !out_ms || nq == 0 || nq > ipx::kMfmaMaxQueries || nq > srv->ctx->cfg.max_batch) return 1;
    GpuServerCtx* c = srv->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->loop_active) return 5;
    const ipx::DbPlan p = c->stream_plan(nq, slices);
    if (p.kind != ipx::DbPlan::MFMA) return 1;
    hipStream_t s = c->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    try {
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        const size_t C = c->pp.db_cols, G = c->G;
        uint32_t* pass_done = c->d_sflags[0];
        uint32_t* packed_done = pass_done + G;
        const uint32_t pq = GpuServerCtx::pack_width(nq);
        const size_t nblk = (nq + pq - 1) / pq;
        const uint32_t target = (uint32_t)(ipx::N / ipx::dbplan_cols(p) * p.slices);
        const bool direct = p.slices == 1;
        uint32_t* part = direct ? c->d_interm : c->d_part;
        const size_t part_qstride = direct ? 2 * C : c->part_stride;
        c->order_before(s);
        hip_check(hipMemsetAsync(c->d_bq, 0, nq * 2 * c->pp.db_rows * sizeof(double), s), "zero");
        c->dbpass_prep(0, nq, p, s, 0);
        auto timed = [&](const std::function<void()>& fn) {
            fn();  // warm
            hip_check(hipEventRecord(e0, s), "event");
            for (int r = 0; r < 5; r++) fn();
            hip_check(hipEventRecord(e1, s), "event");
            hip_check(hipEventSynchronize(e1), "sync");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            return (double)ms / 5;
        };
        // 0: the streamed pass alone (nobody waiting)
        out_ms[0] = timed([&] {
            hip_check(ipx::launch_fill_u32(pass_done, G, 0, s), "fill");
            ipx::PassStream pst;
            pst.group_done = pass_done;
            hip_check(ipx::launch_dbpass_mfma(c->db_planes(), c->d_dig[0], c->d_S[0], nq, part, part_qstride, C,
                                              c->pp.db_rows, C, 0, C, p, s, &pst),
                      "streamed pass");
        });
        // 1: the same plan on its 2-D grid, with the slice reduction kernel
        out_ms[1] = timed([&] { c->dbpass_range(0, nq, 0, G, p, s, 0); });
        // 2: the streamed pack alone, every group handed over
        ipx::PackStream pk;
        pk.part = part;
        pk.part_qstride = part_qstride;
        pk.slices = p.slices;
        pk.C = (uint32_t)C;
        pk.pass_done = pass_done;
        pk.pass_target = target;
        pk.packed_done = packed_done;
        pk.spin_limit = c->stream_spin_limit;
        pk.status = c->d_status;
        pk.queries = pq;
        out_ms[2] = timed([&] {
            hip_check(ipx::launch_fill_u32(pass_done, G, target, s), "fill");
            hip_check(ipx::launch_fill_u32(packed_done, nblk * G, 0, s), "fill");
            hip_check(ipx::launch_pack_stream(pk, c->d_part2, c->part2_stride, c->d_precomp, c->d_keys,
                                              c->keys_stride, c->d_packed_b, c->packed_stride, G, nq, c->nc, s),
                      "streamed pack");
        });
        // 3: the streamed chains alone, every group published
        ipx::StreamWait sw;
        sw.done = packed_done;
        sw.queries = pq;
        sw.stride = (uint32_t)G;
        sw.spin_limit = c->stream_spin_limit;
        sw.status = c->d_status;
        hip_check(ipx::launch_fill_u32(packed_done, nblk * G, 1, s), "fill");
        out_ms[3] = timed([&] {
            hip_check(ipx::launch_horner(c->d_precomp + ipx::PC_A, ipx::PC_GROUP, G, c->Da, c->T, c->d_packed_b,
                                         c->packed_stride, c->d_rgsw, c->rgsw_stride, c->d_resp, c->d_hstate, 0, G,
                                         nq, c->nc, c->cc, s, ipx::FusedPack{}, sw),
                      "streamed chains");
        });
        // 4: the plain chains over every group
        out_ms[4] = timed([&] {
            hip_check(ipx::launch_horner(c->d_precomp + ipx::PC_A, ipx::PC_GROUP, G, c->Da, c->T, c->d_packed_b,
                                         c->packed_stride, c->d_rgsw, c->rgsw_stride, c->d_resp, c->d_hstate, 0, G,
                                         nq, c->nc, c->cc, s),
                      "chains");
        });
        c->order_after(s);
        hip_check(hipStreamSynchronize(s), "sync");
        c->check_status(s, 0);
    } catch (...) {
        rc = 4;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

size_t ipir_dbg_streamed_count(const ipir_server* srv) { return srv ? srv->ctx->n_streamed.load() : 0; }

int ipir_dbg_pass_timing(ipir_server* srv, int on, double* total_ms, size_t* passes) {
    if (!srv) return 1;
    GpuServerCtx* c = srv->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->loop_active) return 5;
    if (on) {
        c->time_pass = true;
        c->time_pass_overflow = false;
        c->pass_marks_used = 0;
        c->passes_timed = 0;
        return 0;
    }
    c->time_pass = false;
    try {
        double sum = 0;
        for (size_t k = 0; k + 1 < c->pass_marks_used; k += 2) {
            hip_check(hipEventSynchronize(c->pass_marks[k + 1]), "pass timing");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, c->pass_marks[k], c->pass_marks[k + 1]), "pass timing");
            sum += ms;
        }
        if (total_ms) *total_ms = sum;
        if (passes) *passes = c->passes_timed;
        return c->time_pass_overflow ? 1 : 0;
    } catch (...) {
        return 4;
    }
}

int ipir_dbg_take_streaming(ipir_server* srv) {
    if (!srv) return 1;
    std::lock_guard<std::mutex> lk(srv->ctx->mu);
    if (srv->ctx->loop_active) return 5;
    srv->ctx->take_streaming();
    srv->ctx->drop_graphs();
    return 0;
}

int ipir_dbg_set_pipeline(ipir_server* srv, int single, int batched) {
    auto ok = [](int v) { return v == 0 || (v >= 2 && v <= 8); };
    if (!srv || !ok(single) || !ok(batched)) return 1;
    std::lock_guard<std::mutex> lk(srv->ctx->mu);
    if (srv->ctx->loop_active) return 5;
    srv->ctx->pipe_single = single;
    srv->ctx->pipe_batched = batched;
    srv->ctx->drop_graphs();  // their kernels carry the old plans
    return 0;
}

int ipir_dbg_set_pack(ipir_server* srv, int mode) {
    if (!srv || mode < -1 || mode > 2) return 1;
    std::lock_guard<std::mutex> lk(srv->ctx->mu);
    if (srv->ctx->loop_active) return 5;
    srv->ctx->pack_wg = mode;
    srv->ctx->drop_graphs();  // their kernels carry the old plans
    return 0;
}

int ipir_dbg_set_graph(ipir_server* srv, int on) {
    if (!srv || (on != 0 && on != 1)) return 1;
    std::lock_guard<std::mutex> lk(srv->ctx->mu);
    if (srv->ctx->loop_active) return 5;
    srv->ctx->drop_graphs();
    srv->ctx->graph_mode = on;
    return 0;
}

int ipir_dbg_set_horner(ipir_server* srv, int mode) {
    if (!srv || mode < -1 || mode > 1) return 1;
    std::lock_guard<std::mutex> lk(srv->ctx->mu);
    if (srv->ctx->loop_active) return 5;
    srv->ctx->horner_mode = mode;
    srv->ctx->drop_graphs();  // their kernels carry the old plans
    return 0;
}

int ipir_dbg_set_horner_spin_limit(ipir_server* srv, uint32_t limit) {
    if (!srv) return 1;
    std::lock_guard<std::mutex> lk(srv->ctx->mu);
    if (srv->ctx->loop_active) return 5;
    srv->ctx->horner_spin_limit = limit;
    srv->ctx->drop_graphs();  // their kernels carry the old bound
    return 0;
}

int ipir_dbg_set_stream_spin_limit(ipir_server* srv, uint32_t limit) {
    if (!srv) return 1;
    std::lock_guard<std::mutex> lk(srv->ctx->mu);
    if (srv->ctx->loop_active) return 5;
    srv->ctx->stream_spin_limit = limit;
    srv->ctx->drop_graphs();  // their kernels carry the old bound
    return 0;
}

int ipir_dbg_horner_variant(const uint64_t* cts, size_t L, const uint64_t* rgsw, uint64_t* out, int variant) {
    if (!cts || !rgsw || !out || L == 0 || variant < 0 || variant > 1) return 1;
    Standalone* st = standalone();
    if (!st) return 4;
    const size_t b = blk();
    std::vector<uint32_t> pa(L * 2 * N), pb(L * 2 * N), rg(4 * b);
    for (size_t g = 0; g < L; g++)
        for (size_t x = 0; x < 2 * N; x++) {
            const uint64_t q = x < N ? Q0 : Q1;
            pa[g * 2 * N + x] = (uint32_t)(cts[g * 4 * N + x] % q);
            pb[g * 2 * N + x] = (uint32_t)(cts[g * 4 * N + 2 * N + x] % q);
        }
    for (size_t x = 0; x < 4 * b; x++) rg[x] = (uint32_t)(rgsw[x] % ((x / N) % 2 ? Q1 : Q0));
    uint32_t *d_pa = nullptr, *d_pb = nullptr, *d_rg = nullptr, *d_x = nullptr, *d_f = nullptr, *d_st = nullptr;
    uint32_t* d_hs = nullptr;
    uint64_t* d_o = nullptr;
    int rc = 0;
    if (hipMalloc(&d_pa, pa.size() * 4) || hipMalloc(&d_pb, pb.size() * 4) || hipMalloc(&d_rg, rg.size() * 4) ||
        hipMalloc(&d_o, 4 * N * 8) || hipMalloc(&d_x, ipx::horner_pair_xbuf_words(1) * 4) || hipMalloc(&d_f, 8) ||
        hipMalloc(&d_st, 4) || hipMemset(d_st, 0, 4) || hipMalloc(&d_hs, 4 * N * 4))
        rc = 4;
    if (!rc && (hipMemcpy(d_pa, pa.data(), pa.size() * 4, hipMemcpyHostToDevice) ||
                hipMemcpy(d_pb, pb.data(), pb.size() * 4, hipMemcpyHostToDevice) ||
                hipMemcpy(d_rg, rg.data(), rg.size() * 4, hipMemcpyHostToDevice)))
        rc = 4;
    if (!rc) {
        const hipError_t e =
            variant == 1 ? ipx::launch_horner_pair(d_pa, 2 * N, L, L, 1, d_pb, L * 2 * N, d_rg, 4 * b, d_o, d_hs, 0, L,
                                                   1, d_x, d_f, d_st, GpuServerCtx::kHornerSpinLimit, st->nc, st->cc,
                                                   nullptr)
                         : ipx::launch_horner(d_pa, 2 * N, L, L, 1, d_pb, L * 2 * N, d_rg, 4 * b, d_o, d_hs, 0, L, 1,
                                              st->nc, st->cc, nullptr);
        uint32_t status = 0;
        if (e || hipMemcpy(out, d_o, 4 * N * 8, hipMemcpyDeviceToHost) ||
            hipMemcpy(&status, d_st, 4, hipMemcpyDeviceToHost) || status)
            rc = 4;
    }
    dfree(d_pa); dfree(d_pb); dfree(d_rg); dfree(d_o); dfree(d_x); dfree(d_f); dfree(d_st); dfree(d_hs);
    return rc;
}

int ipir_dbg_profile_batch(ipir_server* srv, const uint64_t* d_queries, size_t count, uint64_t* d_out,
                           double* stage_ms) {
    if (!srv || !d_queries || !d_out || !stage_ms || count == 0 || count > srv->ctx->cfg.max_batch) return 1;
    GpuServerCtx* c = srv->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->loop_active) return 5;
    hipEvent_t ev[6];
    for (auto& e : ev) (void)hipEventCreate(&e);
    hipStream_t s = c->stream;
    int rc = 0;
    try {
        c->order_before(s);
        c->clear_status(s, 0);
        (void)hipEventRecord(ev[0], s);
        c->stage_prep(d_queries, count, s, 0);
        (void)hipEventRecord(ev[1], s);
        c->stage_dbpass(count, s, 0);
        (void)hipEventRecord(ev[2], s);
        c->stage_collapse(count, s, 0);
        (void)hipEventRecord(ev[3], s);
        c->stage_finalize(count, s, 0);
        (void)hipEventRecord(ev[4], s);
        c->stage_horner(count, d_out, s, 0, true);
        (void)hipEventRecord(ev[5], s);
        c->order_after(s);
        hip_check(hipEventSynchronize(ev[5]), "sync");
        c->check_status(s, 0);
        for (int i = 0; i < 5; i++) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
            stage_ms[i] = ms;
        }
    } catch (...) {
        rc = 4;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return rc;
}

void* ipir_dev_alloc(size_t bytes) {
    void* p = nullptr;
    return hipMalloc(&p, bytes ? bytes : 1) == hipSuccess ? p : nullptr;
}

int ipir_dev_free(void* p) { return p && hipFree(p) == hipSuccess ? 0 : 1; }

int ipir_dev_copy(void* dst, const void* src, size_t bytes, int kind) {
    if (!dst || !src) return 1;
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost
                          : kind == 3 ? hipMemcpyDeviceToDevice : hipMemcpyDefault;
    return hipMemcpy(dst, src, bytes, k) == hipSuccess ? 0 : 4;
}

int ipir_dev_sync(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : 4; }

void* ipir_dev_stream_create(void) {
    hipStream_t s = nullptr;
    return hipStreamCreate(&s) == hipSuccess ? static_cast<void*>(s) : nullptr;
}

int ipir_dev_stream_sync(void* stream) {
    return hipStreamSynchronize(static_cast<hipStream_t>(stream)) == hipSuccess ? 0 : 4;
}

int ipir_dev_stream_destroy(void* stream) {
    return stream && hipStreamDestroy(static_cast<hipStream_t>(stream)) == hipSuccess ? 0 : 1;
}

int ipir_dev_mem_info(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return 1;
    return hipMemGetInfo(free_bytes, total_bytes) == hipSuccess ? 0 : 4;
}

}  // extern "C"
