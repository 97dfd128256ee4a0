// Client query construction and response extraction.
//
// Index decomposition, key generation order and the RGSW message
// mu = X^(2N j*/D_actual) follow the reference query() (src/protocol.cpp:
// 12-52); extract() decrypts response k* and returns slots ell*..ell*+cpe-1
// (src/protocol.cpp:58-71).
#include <cassert>

#include "protocol.h"

namespace inspire {

namespace {
std::pair<QueryState, QueryMessage> build_query(const PublicParams& pp, size_t idx, std::mt19937_64& rng) {
    const size_t cpe = (pp.w * 8 + P_BITS - 1) / P_BITS;
    const size_t i_star = idx % pp.db_rows;
    const size_t col = (idx / pp.db_rows) * cpe;
    const size_t da = pp.D < pp.n_packed ? pp.D : pp.n_packed;
    const size_t j_star = (col / N) % da;
    const size_t k_star = col / (N * da);
    const size_t ell_star = col % N;

    RnsPoly s = sk_gen(rng);
    RnsPoly s_ntt = s;
    s_ntt.to_ntt();
    QueryMessage m;
    m.ksk_5 = ksk_gen_b(s, AUTO_GEN, pp.seed.data(), rng);
    m.ksk_neg1 = ksk_gen_b(s, 2 * N - 1, pp.seed.data(), rng);
    m.lwe = lwe_enc_b(s, i_star, pp.db_rows, pp.seed.data(), rng);
    RnsPoly mu;
    const size_t e = (j_star * 2 * N) / da;
    if (e < N) mu.set_coeff(e, 1, 1);
    else mu.set_coeff(e - N, Q0 - 1, Q1 - 1);
    m.rgsw = rgsw_enc_b(s, mu, pp.seed.data(), rng);
    QueryState st{s, s_ntt, k_star, ell_star};
    return {std::move(st), std::move(m)};
}

std::vector<uint16_t> slots_of(const PublicParams& pp, size_t ell, const std::vector<uint64_t>& m) {
    const size_t cpe = (pp.w * 8 + P_BITS - 1) / P_BITS;
    std::vector<uint16_t> out;
    out.reserve(cpe);
    for (size_t c = 0; c < cpe && ell + c < N; c++) out.push_back((uint16_t)(m[ell + c] % P));
    return out;
}
}  // namespace

std::pair<QueryState, QueryMessage> query(const PublicParams& pp, size_t idx) {
    std::mt19937_64 rng(std::random_device{}());
    return build_query(pp, idx, rng);
}

std::pair<QueryState, QueryMessage> query_seeded(const PublicParams& pp, size_t idx, uint64_t rng_seed) {
    std::mt19937_64 rng(rng_seed);
    return build_query(pp, idx, rng);
}

std::vector<uint16_t> extract(const PublicParams& pp, const QueryState& qst, const std::vector<RlweCt>& resp) {
    assert(qst.k_star < resp.size());
    return slots_of(pp, qst.ell_star, rlwe_dec(qst.s_ntt, resp[qst.k_star]));
}

std::vector<CompressedCt> compress_response(const std::vector<RlweCt>& resp) {
    std::vector<CompressedCt> out;
    out.reserve(resp.size());
    for (const RlweCt& ct : resp) out.push_back(compress_ct(ct));
    return out;
}

std::vector<uint16_t> extract_compressed(const PublicParams& pp, const QueryState& qst,
                                         const std::vector<CompressedCt>& resp) {
    assert(qst.k_star < resp.size());
    return slots_of(pp, qst.ell_star, dec_compressed(qst.s, resp[qst.k_star]));
}

}  // namespace inspire
