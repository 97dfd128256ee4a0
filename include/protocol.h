// Protocol types and the client query/extract path.
//
// Drop-in for the reference's src/protocol.h (PreprocessData :12-24,
// QueryState :30-35, QueryMessage :37-42, query :44, extract :53-55,
// compress_response / extract_compressed :63-67).
#pragma once
#include <utility>

#include "crypto.h"

namespace inspire {

// Device-resident preprocess output, produced by gpu_preprocess and adopted
// (ownership transfer, source fields cleared) by gpu_setup_server.
//
// gpu_preprocess allocates the precomp as one arena (d_precomp_arena) in the
// server's resident layout; the per-group pointers are then views into it:
// a and D_final as in the reference, D_plus the group's collapse tensors in
// dot-product form (D+ and D- together), D_minus null. When d_precomp_arena
// is null the per-group pointers own their memory and hold the reference's
// layout (D+ / D- each [limb][step][digit][coeff]); gpu_setup_server
// converts them and frees them. gpu_setup_server takes ownership only when
// it succeeds; if it throws, every buffer is still the caller's.
struct PreprocessData {
    std::vector<uint32_t*> d_precomp_a;        // [n_packed] -> 2*N u32, NTT form
    std::vector<uint32_t*> d_precomp_D_plus;   // [n_packed] (see above)
    std::vector<uint32_t*> d_precomp_D_minus;  // [n_packed] (see above)
    std::vector<uint32_t*> d_precomp_D_final;  // [n_packed] -> 2*D_EFF*N u32
    // encoded DB in the server's resident layout, as gpu_preprocess leaves
    // it: two byte planes (low, high byte) in 16-row blocks, plane p of
    // (r, c) at byte ((r/16)*2 + p)*cols*16 + c*16 + r%16 (db_rows*db_cols*2
    // bytes). gpu_setup_server adopts it as is, also with per-group
    // precomp buffers (it does not accept the reference's row-major DB).
    uint16_t* d_db_col = nullptr;
    void* d_precomp_arena = nullptr;
};

struct QueryState {
    RnsPoly s;        // secret, coefficient form
    RnsPoly s_ntt;    // secret, NTT form
    size_t k_star;    // response ciphertext index
    size_t ell_star;  // slot offset within that ciphertext
};

struct QueryMessage {
    Ksk ksk_5;
    Ksk ksk_neg1;
    LweQuery lwe;
    RgswCt rgsw;
};

// Fresh secret per call (std::random_device).
std::pair<QueryState, QueryMessage> query(const PublicParams& pp, size_t idx);
// Same with a caller-seeded generator (no counterpart in the reference;
// for reproducible tests and benchmarks).
std::pair<QueryState, QueryMessage> query_seeded(const PublicParams& pp, size_t idx, uint64_t rng_seed);

std::vector<uint16_t> extract(const PublicParams& pp, const QueryState& qst,
                              const std::vector<RlweCt>& resp);
std::vector<CompressedCt> compress_response(const std::vector<RlweCt>& resp);
std::vector<uint16_t> extract_compressed(const PublicParams& pp, const QueryState& qst,
                                         const std::vector<CompressedCt>& resp);

}  // namespace inspire
