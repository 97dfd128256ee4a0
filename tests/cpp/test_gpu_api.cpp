// The C++ server API as the reference's own callers drive it
// (tests/test_gpu_batch.cu:29-115, tests/test_e2e_gpu.cu:95-104):
//   setup -> gpu_preprocess -> gpu_setup_server(max_batch 10) -> gpu_server_caps
//   -> gpu_answer / gpu_answer_batch -> extract -> gpu_free_server
// plus the adoption of per-group precomp buffers in the reference's layout
// (PreprocessData with no arena), its failure path (the caller keeps its
// buffers) and its answers against the arena server's.
//
// Built in-tree by inspire-gpu_amd/Makefile and run by a GPU test. Prints
// one line per check; exit status 0 when every check passed.
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

#include "capi.h"
#include "gpu_protocol.h"
#include "inspire_debug.h"

using namespace inspire;

namespace {
int fails = 0;

void check(bool ok, const char* what) {
    std::printf("%s: %s\n", ok ? "ok  " : "FAIL", what);
    std::fflush(stdout);
    if (!ok) fails++;
}

bool ct_equal(const RlweCt& x, const RlweCt& y) {
    for (int l = 0; l < 2; l++)
        for (size_t i = 0; i < N; i++)
            if (x.a.limbs[l][i] != y.a.limbs[l][i] || x.b.limbs[l][i] != y.b.limbs[l][i]) return false;
    return true;
}

bool resp_equal(const std::vector<RlweCt>& x, const std::vector<RlweCt>& y) {
    if (x.size() != y.size()) return false;
    for (size_t g = 0; g < x.size(); g++)
        if (!ct_equal(x[g], y[g])) return false;
    return true;
}

// Group g's precomp in the reference layout, uploaded into per-group device
// buffers: a [2][N], D+ / D- [2][N/2-1][2][N], D_final [2][2][N] (u32).
void upload_group(ipir_server* srv, size_t g, PreprocessData& pd) {
    std::vector<uint32_t> w(ipir_dbg_precomp_group_words());
    if (ipir_dbg_precomp_group(srv, g, w.data())) throw std::runtime_error("precomp group");
    const size_t a = 2 * N, d = 2 * (N / 2 - 1) * D_EFF * N, f = 2 * D_EFF * N;
    auto up = [&](size_t off, size_t words) {
        void* p = ipir_dev_alloc(words * 4);
        if (!p || ipir_dev_copy(p, w.data() + off, words * 4, 1)) throw std::runtime_error("upload");
        return static_cast<uint32_t*>(p);
    };
    pd.d_precomp_a[g] = up(0, a);
    pd.d_precomp_D_plus[g] = up(a, d);
    pd.d_precomp_D_minus[g] = up(a + d, d);
    pd.d_precomp_D_final[g] = up(a + 2 * d, f);
}
}  // namespace

int main() {
    const size_t n_entries = 524288, entry_bytes = 120, db_rows = 2048;
    uint8_t seed[SEED_BYTES] = {0x42};
    const PublicParams pp = setup_seeded(n_entries, entry_bytes, db_rows, seed);
    const size_t cpe = (pp.w * 8 + P_BITS - 1) / P_BITS;
    std::printf("geometry: %zu x %zu, n_packed %zu, num_cts %zu\n", pp.db_rows, pp.db_cols, pp.n_packed,
                pp.num_cts);
    std::mt19937_64 rng(42);
    std::vector<uint16_t> db(pp.db_rows * pp.db_cols);
    for (auto& v : db) v = (uint16_t)(rng() % P);

    PreprocessData data = gpu_preprocess(pp, db.data());
    GpuServerConfig cfg;
    cfg.max_batch = 10;
    GpuServerCtx* ctx = gpu_setup_server(pp, data, cfg);
    check(data.d_db_col == nullptr && data.d_precomp_arena == nullptr && data.d_precomp_a.empty(),
          "gpu_setup_server adopts the preprocess output (source fields cleared)");
    const GpuServerCaps caps = gpu_server_caps(ctx);
    check(caps.max_batch == 10 && caps.num_cts == pp.num_cts && caps.resident_bytes > 0 && caps.device_free_bytes > 0,
          "caps: max_batch, num_cts, resident and free bytes");

    // 5 indices, each queried twice with independent keys (fresh secrets)
    const std::vector<size_t> idx = {0, 42, 1000, 100000, 524287, 0, 42, 1000, 100000, 524287};
    std::vector<QueryState> states;
    std::vector<QueryMessage> qs;
    for (size_t i : idx) {
        auto [st, q] = query(pp, i);
        states.push_back(st);
        qs.push_back(q);
    }
    std::vector<std::vector<RlweCt>> single;
    for (const auto& q : qs) single.push_back(gpu_answer(ctx, q));
    auto batch = gpu_answer_batch(ctx, qs);
    bool ok = batch.size() == qs.size();
    for (size_t i = 0; ok && i < qs.size(); i++) ok = resp_equal(batch[i], single[i]);
    check(ok, "gpu_answer_batch of 10 equals 10 gpu_answer calls");
    ok = true;
    for (size_t i = 0; i < qs.size(); i++) {
        const auto slots = extract(pp, states[i], batch[i]);
        const size_t r = idx[i] % pp.db_rows, col = (idx[i] / pp.db_rows) * cpe;
        ok = ok && slots.size() == cpe;
        for (size_t c = 0; ok && c < cpe; c++) ok = slots[c] == db[r * pp.db_cols + col + c];
    }
    check(ok, "extract() returns every queried entry");
    auto again = gpu_answer_batch(ctx, qs);
    ok = true;
    for (size_t i = 0; i < qs.size(); i++) ok = ok && resp_equal(again[i], batch[i]);
    check(ok, "rerun of the same batch is bit-identical");
    for (size_t k : {3, 2}) {
        auto part = gpu_answer_batch(ctx, qs.data(), k);
        ok = part.size() == k;
        for (size_t i = 0; ok && i < k; i++) ok = resp_equal(part[i], single[i]);
        check(ok, k == 3 ? "partial batch of 3" : "partial batch of 2");
    }
    std::vector<QueryMessage> eleven(qs);
    eleven.push_back(qs[0]);
    bool threw = false;
    try {
        gpu_answer_batch(ctx, eleven);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    check(threw, "count > max_batch throws std::invalid_argument");
    QueryMessage bad = qs[1];
    bad.lwe.b_limb0.pop_back();
    threw = false;
    try {
        gpu_answer(ctx, bad);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    check(threw, "a query of the wrong length throws std::invalid_argument");
    check(resp_equal(gpu_answer(ctx, qs[1]), single[1]), "the server answers after rejected calls");

    // per-group precomp in the reference layout, from the same preprocess
    ipir_params prm{};
    ipir_server* srv = ipir_server_create(n_entries, entry_bytes, db_rows, 2, db.data(), seed, &prm);
    check(srv != nullptr, "C ABI server with the same seed");
    if (!srv) return 1;
    PreprocessData pd = gpu_preprocess(pp, db.data());
    ipir_dev_free(pd.d_precomp_arena);  // keep the encoded DB, replace the arena
    pd.d_precomp_arena = nullptr;
    for (size_t g = 0; g < pp.n_packed; g++) upload_group(srv, g, pd);
    ipir_server_destroy(srv);
    uint16_t* const db_dev = pd.d_db_col;
    uint32_t* const df1 = pd.d_precomp_D_final[1];
    pd.d_precomp_D_final[1] = nullptr;
    threw = false;
    try {
        gpu_setup_server(pp, pd, cfg);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    check(threw && pd.d_db_col == db_dev && pd.d_precomp_a.size() == pp.n_packed,
          "a missing per-group buffer throws and the caller keeps every buffer");
    pd.d_precomp_D_final[1] = df1;
    GpuServerCtx* ctx2 = gpu_setup_server(pp, pd, cfg);
    check(pd.d_db_col == nullptr && pd.d_precomp_D_plus.empty(), "per-group buffers adopted (source cleared)");
    auto batch2 = gpu_answer_batch(ctx2, qs);
    ok = true;
    for (size_t i = 0; i < qs.size(); i++) ok = ok && resp_equal(batch2[i], batch[i]);
    check(ok, "answers from reference-layout per-group precomp equal the arena server's");
    check(resp_equal(gpu_answer(ctx2, qs[4]), single[4]), "single answer on the per-group server");
    gpu_free_server(ctx2);
    gpu_free_server(ctx);
    std::printf("%s (%d failed)\n", fails ? "FAILED" : "all C++ API checks passed", fails);
    return fails ? 1 : 0;
}
