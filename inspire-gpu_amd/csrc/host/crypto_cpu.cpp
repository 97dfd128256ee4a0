// Client-side RLWE primitives on the CPU: key generation, LWE query rows,
// key-switching keys, RGSW encryption, decryption, external product,
// Horner evaluation and mod-switch compression.
//
// Each follows the corresponding reference algorithm so queries and
// responses interoperate: rlwe_dec (src/crypto.cpp:19-58), lwe_enc_b
// (:65-101), ksk_gen_b (:108-145, domain "ksk", index k*100+digit),
// rgsw_enc_b (:151-213, domains "rgsw0"/"rgsw1"), ext_prod (:219-257),
// horner_eval (:264-279), compress_ct / dec_compressed (:285-352).
#include <cassert>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "crypto.h"

namespace inspire {

namespace {
const uint64_t kQ1invQ0 = mod_inv(Q1 % Q0, Q0);
const uint64_t kQ0invQ1 = mod_inv(Q0 % Q1, Q1);
inline uint64_t crt_lift(uint64_t r0, uint64_t r1) {
    uint64_t v = (uint64_t)((unsigned __int128)r0 * kQ1invQ0 % Q0) * Q1 +
                 (uint64_t)((unsigned __int128)r1 * kQ0invQ1 % Q1) * Q0;
    return v >= Q_FULL ? v - Q_FULL : v;
}
uint64_t pow_base(int digit) {
    uint64_t b = 1;
    for (int t = 0; t < digit; t++) b *= BASE;
    return b;
}
// a * s in coefficient form, a given in NTT form
RnsPoly times_s(const RnsPoly& a_ntt, const RnsPoly& s_ntt) {
    RnsPoly t = a_ntt;
    t.mul_inplace(s_ntt);
    t.to_coeff();
    return t;
}

// The CRS polynomials a client multiplies by its secret are fixed by the
// public seed: expand_seed(seed, domain, index) in NTT form, computed once
// per (seed, domain, index) and shared by every later query of the process
// (the reference expands them again for each query; its docs note they are
// cacheable, docs/ARCHITECTURE.md:118-120). At most kCrsCacheEntries are
// kept; a full cache is cleared, so memory stays bounded for clients that
// talk to many servers.
constexpr size_t kCrsCacheEntries = 1 << 12;

std::shared_ptr<const RnsPoly> crs_ntt(const uint8_t seed[SEED_BYTES], const char* domain, uint64_t index) {
    static std::mutex mu;
    static std::map<std::string, std::shared_ptr<const RnsPoly>> cache;
    std::string key(reinterpret_cast<const char*>(seed), SEED_BYTES);
    key += domain;
    key.push_back('\0');
    key.append(reinterpret_cast<const char*>(&index), sizeof index);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    auto p = std::make_shared<RnsPoly>(expand_seed(seed, domain, index));
    p->to_ntt();
    std::lock_guard<std::mutex> lk(mu);
    if (cache.size() >= kCrsCacheEntries) cache.clear();
    return cache.emplace(key, p).first->second;
}
}  // namespace

RnsPoly sk_gen(std::mt19937_64& rng) { return sample_ternary(rng); }

std::vector<uint64_t> rlwe_dec(const RnsPoly& s_ntt, const RlweCt& ct) {
    RnsPoly as = ct.a;
    as.to_ntt();
    as.mul_inplace(s_ntt);
    RnsPoly v = ct.b;
    v.to_ntt();
    v.sub_inplace(as);
    v.to_coeff();
    std::vector<uint64_t> m(N);
    for (size_t i = 0; i < N; i++) {
        const uint64_t val = crt_lift(v.limbs[0][i], v.limbs[1][i]);
        const int64_t c = val > Q_FULL / 2 ? (int64_t)(val - Q_FULL) : (int64_t)val;
        const int64_t r = (int64_t)std::round((double)c * (double)P / (double)Q_FULL);
        m[i] = (uint64_t)(((r % (int64_t)P) + (int64_t)P) % (int64_t)P);
    }
    return m;
}

LweQuery lwe_enc_b(const RnsPoly& s, size_t i_star, size_t ell, const uint8_t seed[SEED_BYTES],
                   std::mt19937_64& rng) {
    LweQuery q;
    q.b_limb0.resize(ell);
    q.b_limb1.resize(ell);
    RnsPoly s_ntt = s;
    s_ntt.to_ntt();
    for (size_t r = 0; r < ell / N; r++) {
        RnsPoly b = times_s(*crs_ntt(seed, "rlwe", r), s_ntt);
        b.add_inplace(sample_error(rng));
        if (r == i_star / N) {
            RnsPoly m;
            m.set_coeff(i_star % N, DELTA_MOD_Q0, DELTA_MOD_Q1);
            b.add_inplace(m);
        }
        std::copy(b.limbs[0].begin(), b.limbs[0].end(), q.b_limb0.begin() + r * N);
        std::copy(b.limbs[1].begin(), b.limbs[1].end(), q.b_limb1.begin() + r * N);
    }
    return q;
}

Ksk ksk_gen_b(const RnsPoly& s, uint64_t k, const uint8_t seed[SEED_BYTES], std::mt19937_64& rng) {
    Ksk ksk;
    ksk.b_parts.resize(D_EFF);
    RnsPoly s_ntt = s;
    s_ntt.to_ntt();
    const RnsPoly s_auto = automorphism(s, k);
    for (int j = 0; j < D_EFF; j++) {
        const int digit = j + 1;
        RnsPoly b = times_s(*crs_ntt(seed, "ksk", k * 100 + digit), s_ntt);
        b.add_inplace(sample_error(rng));
        RnsPoly sk = s_auto;
        const uint64_t bj = pow_base(digit);
        sk.scalar_mul_inplace(bj % Q0, bj % Q1);
        b.add_inplace(sk);
        ksk.b_parts[j] = b;
    }
    return ksk;
}

RgswCt rgsw_enc_b(const RnsPoly& s, const RnsPoly& mu, const uint8_t seed[SEED_BYTES], std::mt19937_64& rng) {
    RgswCt g;
    for (auto* v : {&g.top.a_parts, &g.top.b_parts, &g.bottom.a_parts, &g.bottom.b_parts}) v->resize(D_EFF);
    RnsPoly s_ntt = s;
    s_ntt.to_ntt();
    RnsPoly mu_ntt = mu;
    mu_ntt.to_ntt();
    for (int j = 0; j < D_EFF; j++) {
        const int digit = j + 1;
        const uint64_t bj = pow_base(digit);
        RnsPoly s_mu = s;  // s * B^j * mu, coefficient form
        s_mu.scalar_mul_inplace(bj % Q0, bj % Q1);
        s_mu.to_ntt();
        s_mu.mul_inplace(mu_ntt);
        s_mu.to_coeff();
        RnsPoly b_mu = mu;  // B^j * mu
        b_mu.scalar_mul_inplace(bj % Q0, bj % Q1);

        const auto a0 = crs_ntt(seed, "rgsw0", digit);
        RnsPoly b = times_s(*a0, s_ntt);
        b.add_inplace(sample_error(rng));
        b.sub_inplace(s_mu);
        b.to_ntt();
        g.top.a_parts[j] = *a0; g.top.b_parts[j] = b;

        const auto a1 = crs_ntt(seed, "rgsw1", digit);
        b = times_s(*a1, s_ntt);
        b.add_inplace(sample_error(rng));
        b.add_inplace(b_mu);
        b.to_ntt();
        g.bottom.a_parts[j] = *a1; g.bottom.b_parts[j] = b;
    }
    return g;
}

RlweCt ext_prod(const RgswCt& rgsw, const RlweCt& ct) {
    RnsPoly ac = ct.a, bc = ct.b;
    ac.to_coeff();
    bc.to_coeff();
    std::vector<RnsPoly> da = decomp(ac), db = decomp(bc);
    RlweCt r;
    r.a.is_ntt = r.b.is_ntt = true;
    for (int j = 0; j < D_EFF; j++) {
        da[j].to_ntt();
        db[j].to_ntt();
        for (int l = 0; l < 2; l++) {
            const uint64_t q = l ? Q1 : Q0;
            for (size_t i = 0; i < N; i++) {
                const unsigned __int128 xa = (unsigned __int128)da[j].limbs[l][i] * rgsw.top.a_parts[j].limbs[l][i] +
                                             (unsigned __int128)db[j].limbs[l][i] * rgsw.bottom.a_parts[j].limbs[l][i];
                const unsigned __int128 xb = (unsigned __int128)da[j].limbs[l][i] * rgsw.top.b_parts[j].limbs[l][i] +
                                             (unsigned __int128)db[j].limbs[l][i] * rgsw.bottom.b_parts[j].limbs[l][i];
                r.a.limbs[l][i] = (uint64_t)((r.a.limbs[l][i] + xa) % q);
                r.b.limbs[l][i] = (uint64_t)((r.b.limbs[l][i] + xb) % q);
            }
        }
    }
    return r;
}

RlweCt horner_eval(const std::vector<RlweCt>& cts, const RgswCt& rgsw, size_t D) {
    assert(cts.size() >= D);
    RlweCt acc = cts[D - 1];
    acc.a.to_ntt();
    acc.b.to_ntt();
    for (int i = (int)D - 2; i >= 0; i--) {
        acc = ext_prod(rgsw, acc);
        RlweCt c = cts[i];
        c.a.to_ntt();
        c.b.to_ntt();
        acc.a.add_inplace(c.a);
        acc.b.add_inplace(c.b);
    }
    return acc;
}

CompressedCt compress_ct(const RlweCt& ct) {
    RnsPoly a = ct.a, b = ct.b;
    a.to_coeff();
    b.to_coeff();
    // round(v * q' / Q) mod q', exactly
    auto down = [](uint64_t v, uint64_t qp) {
        const unsigned __int128 num = (unsigned __int128)v * qp + Q_FULL / 2;
        return (uint32_t)((uint64_t)(num / Q_FULL) % qp);
    };
    CompressedCt c;
    c.a.resize(N);
    c.b.resize(N);
    for (size_t i = 0; i < N; i++) {
        c.a[i] = down(crt_lift(a.limbs[0][i], a.limbs[1][i]), Q_PRIME_1);
        c.b[i] = down(crt_lift(b.limbs[0][i], b.limbs[1][i]), Q_PRIME_0);
    }
    return c;
}

std::vector<uint64_t> dec_compressed(const RnsPoly& s, const CompressedCt& cct) {
    // a' * s over the integers (negacyclic), walking the nonzero key entries
    std::vector<int64_t> c(N, 0);
    for (size_t j = 0; j < N; j++) {
        const uint64_t v = s.limbs[0][j];
        if (v == 0) continue;
        const int64_t sj = v == 1 ? 1 : -1;
        for (size_t i = 0; i < N; i++) {
            const int64_t t = sj * (int64_t)cct.a[i];
            if (i + j < N) c[i + j] += t; else c[i + j - N] -= t;
        }
    }
    const __int128 QP = (__int128)Q_PRIME_0 * Q_PRIME_1;
    std::vector<uint64_t> m(N);
    for (size_t i = 0; i < N; i++) {
        __int128 num = ((__int128)cct.b[i] * Q_PRIME_1 - (__int128)c[i] * Q_PRIME_0) % QP;
        if (num < 0) num += QP;
        if (num >= QP / 2) num -= QP;
        const __int128 sc = num * (__int128)P;
        const __int128 rd = sc >= 0 ? (sc + QP / 2) / QP : -((-sc + QP / 2) / QP);
        const int64_t r = (int64_t)(rd % (__int128)P);
        m[i] = (uint64_t)((r % (int64_t)P + (int64_t)P) % (int64_t)P);
    }
    return m;
}

}  // namespace inspire
