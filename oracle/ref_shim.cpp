// ref_shim.cpp — TEST INFRASTRUCTURE ONLY.
//
// extern "C" entry points over the reference's own CPU library
// (src/params.cpp, ntt.cpp, ring.cpp, crypto.cpp, protocol.cpp,
// capi_client.cpp), compiled in place from the reference tree by
// oracle/Makefile into oracle/_ref/. Used by tests to pin the oracle
// against the reference's actual code; never linked into the product.
#include "params.h"
#include "ntt.h"
#include "ring.h"
#include "crypto.h"
#include "protocol.h"
#include <cstring>
#include <stdexcept>

using namespace inspire;

static RnsPoly from_flat(const uint64_t* p, bool ntt) {
    RnsPoly r;
    for (size_t i = 0; i < N; i++) { r.limbs[0][i] = p[i]; r.limbs[1][i] = p[N + i]; }
    r.is_ntt = ntt;
    return r;
}
static void to_flat(const RnsPoly& r, uint64_t* p) {
    for (size_t i = 0; i < N; i++) { p[i] = r.limbs[0][i]; p[N + i] = r.limbs[1][i]; }
}
static RgswCt rgsw_from_flat(const uint64_t* w) {
    const size_t blk = (size_t)D_EFF * 2 * N;
    RgswCt g;
    std::vector<RnsPoly>* parts[4] = {&g.top.a_parts, &g.top.b_parts, &g.bottom.a_parts, &g.bottom.b_parts};
    for (int k = 0; k < 4; k++) {
        parts[k]->resize(D_EFF);
        for (int j = 0; j < D_EFF; j++) (*parts[k])[j] = from_flat(w + k * blk + (size_t)j * 2 * N, true);
    }
    return g;
}
static void rgsw_to_flat(const RgswCt& g, uint64_t* w) {
    const size_t blk = (size_t)D_EFF * 2 * N;
    const std::vector<RnsPoly>* parts[4] = {&g.top.a_parts, &g.top.b_parts, &g.bottom.a_parts, &g.bottom.b_parts};
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < D_EFF; j++) to_flat((*parts[k])[j], w + k * blk + (size_t)j * 2 * N);
}

extern "C" {

void ref_ntt_tables(int limb, uint64_t* fwd, uint64_t* inv, uint64_t* fwdp, uint64_t* invp, uint64_t* scal4) {
    const NttTables& t = ntt_tables(limb);
    for (size_t i = 0; i < N; i++) { fwd[i] = t.fwd_twiddles[i]; inv[i] = t.inv_twiddles[i]; fwdp[i] = t.fwd_primes[i]; invp[i] = t.inv_primes[i]; }
    scal4[0] = t.psi; scal4[1] = t.inv_n; scal4[2] = t.inv_n_prime; scal4[3] = t.q;
}
void ref_ntt_fwd(uint64_t* a, int limb) { ntt_forward(a, limb); }
void ref_ntt_inv(uint64_t* a, int limb) { ntt_inverse(a, limb); }
void ref_automorphism(const uint64_t* f, uint64_t k, uint64_t* out) { to_flat(automorphism(from_flat(f, false), k), out); }
void ref_automorphism_ntt(const uint64_t* f, uint64_t k, uint64_t* out) { to_flat(automorphism_ntt(from_flat(f, true), k), out); }
void ref_decomp(const uint64_t* f, uint64_t* out) {
    auto d = decomp(from_flat(f, false));
    for (int j = 0; j < D_EFF; j++) to_flat(d[j], out + (size_t)j * 2 * N);
}
void ref_expand_seed(const uint8_t* seed, const char* domain, uint64_t index, uint64_t* out) {
    to_flat(expand_seed(seed, domain, index), out);
}
void ref_ext_prod(const uint64_t* rgsw, const uint64_t* ct, uint64_t* res) {
    RlweCt c{from_flat(ct, true), from_flat(ct + 2 * N, true)};
    RlweCt r = ext_prod(rgsw_from_flat(rgsw), c);
    to_flat(r.a, res); to_flat(r.b, res + 2 * N);
}
void ref_horner(const uint64_t* cts, uint64_t L, const uint64_t* rgsw, uint64_t* out) {
    std::vector<RlweCt> v(L);
    for (uint64_t i = 0; i < L; i++) v[i] = RlweCt{from_flat(cts + i * 4 * N, true), from_flat(cts + i * 4 * N + 2 * N, true)};
    RlweCt r = horner_eval(v, rgsw_from_flat(rgsw), L);
    to_flat(r.a, out); to_flat(r.b, out + 2 * N);
}
void ref_rlwe_dec(const uint64_t* s_ntt, const uint64_t* ct, uint64_t* m) {
    RlweCt c{from_flat(ct, true), from_flat(ct + 2 * N, true)};
    auto v = rlwe_dec(from_flat(s_ntt, true), c);
    for (size_t i = 0; i < N; i++) m[i] = v[i];
}
// setup(): 0 ok (out filled), 1 on std::invalid_argument.
int ref_setup(uint64_t n, uint64_t w, uint64_t rows, uint64_t* out) {
    try {
        PublicParams pp = setup(n, w, rows);
        out[0] = pp.db_rows; out[1] = pp.db_cols; out[2] = pp.n_packed; out[3] = pp.D; out[4] = pp.num_cts;
        return 0;
    } catch (const std::invalid_argument&) { return 1; }
}
// query() with mt19937_64(rng_seed) in place of random_device; same call
// sequence and index math as src/protocol.cpp:12-52. Output layout matches
// orc_query in oracle/inspire_oracle.cpp.
void ref_query_seeded(uint64_t R, uint64_t C, uint64_t D, uint64_t w, const uint8_t* seed,
                      uint64_t idx, uint64_t rng_seed, uint64_t* out_query, uint64_t* out_state) {
    PublicParams pp{};
    pp.db_rows = R; pp.db_cols = C; pp.n_packed = C / N; pp.D = D; pp.w = w;
    pp.num_cts = std::max<size_t>(1, pp.n_packed / D);
    std::memcpy(pp.seed.data(), seed, SEED_BYTES);
    std::mt19937_64 rng(rng_seed);
    size_t i_star = idx % pp.db_rows;
    size_t cpe = (pp.w * 8 + P_BITS - 1) / P_BITS;
    size_t col = (idx / pp.db_rows) * cpe;
    size_t Da = std::min(pp.D, pp.n_packed);
    size_t j_star = (col / N) % Da, k_star = col / (N * Da), ell_star = col % N;
    RnsPoly s = sk_gen(rng);
    Ksk k5 = ksk_gen_b(s, AUTO_GEN, pp.seed.data(), rng);
    Ksk kn = ksk_gen_b(s, 2 * N - 1, pp.seed.data(), rng);
    LweQuery lwe = lwe_enc_b(s, i_star, pp.db_rows, pp.seed.data(), rng);
    RnsPoly mu = RnsPoly::zero();
    size_t e = (j_star * 2 * N) / Da;
    if (e < N) mu.set_coeff(e, 1, 1); else mu.set_coeff(e - N, Q0 - 1, Q1 - 1);
    RgswCt g = rgsw_enc_b(s, mu, pp.seed.data(), rng);
    uint64_t* w0 = out_query;
    std::memcpy(w0, lwe.b_limb0.data(), R * 8); std::memcpy(w0 + R, lwe.b_limb1.data(), R * 8);
    const size_t blk = (size_t)D_EFF * 2 * N;
    for (int j = 0; j < D_EFF; j++) {
        to_flat(k5.b_parts[j], w0 + 2 * R + (size_t)j * 2 * N);
        to_flat(kn.b_parts[j], w0 + 2 * R + blk + (size_t)j * 2 * N);
    }
    rgsw_to_flat(g, w0 + 2 * R + 2 * blk);
    to_flat(s, out_state);
    RnsPoly sn = s; sn.to_ntt(); to_flat(sn, out_state + 2 * N);
    out_state[4 * N] = k_star; out_state[4 * N + 1] = ell_star;
}
void ref_compress_ct(const uint64_t* ct, uint32_t* a_out, uint32_t* b_out) {
    RlweCt c{from_flat(ct, true), from_flat(ct + 2 * N, true)};
    CompressedCt cc = compress_ct(c);
    for (size_t i = 0; i < N; i++) { a_out[i] = cc.a[i]; b_out[i] = cc.b[i]; }
}

} // extern "C"
