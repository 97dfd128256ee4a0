// Client half of the C ABI: query build / extract over the flat layouts,
// the lossless 53-bit CRT wire packing and the mod-switch response
// compression. CPU only.
//
// Byte formats are those of the reference (src/capi_client.cpp):
// parameter validation (:21-38), flat query layout (:84-94), Garner CRT
// combine (:174-179), MSB-first bit streams with a zero-padded final byte
// (:184-222), compressed responses of N x 28-bit then N x 20-bit values
// flushed per ciphertext (:310-385).
#include <array>
#include <cstring>
#include <new>

#include "capi.h"
#include "protocol.h"

using namespace inspire;

struct ipir_client_query {
    QueryState st;
};

namespace {

bool params_ok(const ipir_params* p) {
    if (!p || p->ring_n != N || p->d_eff != (size_t)D_EFF) return false;
    if (p->db_rows == 0 || p->db_rows > (1ULL << 26)) return false;
    if (p->n_entries == 0 || p->entry_bytes == 0 || p->entry_bytes > (1ULL << 20)) return false;
    if (p->n_packed == 0 || p->n_packed > (1ULL << 20)) return false;
    if (p->interp_d == 0 || p->interp_d > (1ULL << 20)) return false;
    if (p->num_cts == 0 || p->num_cts > (1ULL << 16)) return false;
    const size_t cpe = (p->entry_bytes * 8 + P_BITS - 1) / P_BITS;
    if (N % cpe != 0) return false;
    if (p->db_cols != p->n_packed * N) return false;
    const size_t da = p->interp_d < p->n_packed ? p->interp_d : p->n_packed;
    if ((2 * N) % da != 0) return false;
    if (p->n_packed > p->interp_d && p->n_packed % p->interp_d != 0) return false;
    const size_t ct = p->n_packed / p->interp_d > 1 ? p->n_packed / p->interp_d : 1;
    return p->num_cts == ct;
}

PublicParams pp_of(const ipir_params* p) {
    PublicParams pp{};
    pp.N_entries = p->n_entries; pp.w = p->entry_bytes; pp.db_rows = p->db_rows; pp.db_cols = p->db_cols;
    pp.n_packed = p->n_packed; pp.D = p->interp_d; pp.num_cts = p->num_cts;
    std::memcpy(pp.seed.data(), p->seed, SEED_BYTES);
    return pp;
}

uint64_t* put_parts(uint64_t* w, const std::vector<RnsPoly>& parts, bool want_ntt) {
    for (const RnsPoly& p0 : parts) {
        RnsPoly p = p0;
        if (want_ntt) p.to_ntt(); else p.to_coeff();
        std::memcpy(w, p.limbs[0].data(), N * 8); w += N;
        std::memcpy(w, p.limbs[1].data(), N * 8); w += N;
    }
    return w;
}

std::vector<RlweCt> cts_from_flat(size_t num_cts, const uint64_t* r) {
    std::vector<RlweCt> cts(num_cts);
    for (auto& ct : cts) {
        for (RnsPoly* p : {&ct.a, &ct.b}) {
            std::memcpy(p->limbs[0].data(), r, N * 8); r += N;
            std::memcpy(p->limbs[1].data(), r, N * 8); r += N;
            p->is_ntt = true;
        }
    }
    return cts;
}

// one value mod Q0*Q1 from its residues (Garner)
uint64_t garner(uint64_t v0, uint64_t v1) {
    static const uint64_t inv = mod_inv(Q1 % Q0, Q0);
    const uint64_t d = (v0 + Q0 - v1 % Q0) % Q0;
    return v1 + Q1 * (uint64_t)((unsigned __int128)d * inv % Q0);
}

// MSB-first bit stream writer/reader over fields of up to 53 bits
class BitOut {
    uint8_t* p_; uint64_t acc_ = 0; int n_ = 0;
public:
    explicit BitOut(uint8_t* p) : p_(p) {}
    void put(uint64_t v, int bits) {
        acc_ = (acc_ << bits) | v;
        n_ += bits;
        while (n_ >= 8) { n_ -= 8; *p_++ = (uint8_t)(acc_ >> n_); }
        acc_ &= (1ULL << n_) - 1;
    }
    void flush() { if (n_) { *p_++ = (uint8_t)(acc_ << (8 - n_)); n_ = 0; acc_ = 0; } }
};
class BitIn {
    const uint8_t* p_; uint64_t acc_ = 0; int n_ = 0;
public:
    explicit BitIn(const uint8_t* p) : p_(p) {}
    uint64_t get(int bits) {
        while (n_ < bits) { acc_ = (acc_ << 8) | *p_++; n_ += 8; }
        n_ -= bits;
        const uint64_t v = acc_ >> n_;
        acc_ &= (1ULL << n_) - 1;
        return v;
    }
};

size_t query_values(const ipir_params* p) { return p->db_rows + 6 * p->d_eff * p->ring_n; }
size_t response_values(const ipir_params* p) { return p->num_cts * 2 * p->ring_n; }

// walk the (limb0, limb1) pairs of a flat query / response
template <class F> void walk_query(const ipir_params* p, F f) {
    const size_t R = p->db_rows, n = p->ring_n;
    for (size_t i = 0; i < R; i++) f(i, R + i);
    for (size_t j = 0, base = 2 * R; j < 6 * p->d_eff; j++, base += 2 * n)
        for (size_t i = 0; i < n; i++) f(base + i, base + n + i);
}
template <class F> void walk_response(const ipir_params* p, F f) {
    const size_t n = p->ring_n;
    for (size_t j = 0, base = 0; j < p->num_cts * 2; j++, base += 2 * n)
        for (size_t i = 0; i < n; i++) f(base + i, base + n + i);
}

ipir_client_query* build(const ipir_params* params, uint64_t idx, uint64_t* out, const uint64_t* rng_seed) {
    if (!params_ok(params) || !out || idx >= params->n_entries) return nullptr;
    try {
        const PublicParams pp = pp_of(params);
        auto qm = rng_seed ? query_seeded(pp, (size_t)idx, *rng_seed) : query(pp, (size_t)idx);
        auto* cq = new (std::nothrow) ipir_client_query;
        if (!cq) return nullptr;
        cq->st = std::move(qm.first);
        const QueryMessage& m = qm.second;
        uint64_t* w = out;
        std::memcpy(w, m.lwe.b_limb0.data(), params->db_rows * 8); w += params->db_rows;
        std::memcpy(w, m.lwe.b_limb1.data(), params->db_rows * 8); w += params->db_rows;
        w = put_parts(w, m.ksk_5.b_parts, false);
        w = put_parts(w, m.ksk_neg1.b_parts, false);
        w = put_parts(w, m.rgsw.top.a_parts, true);
        w = put_parts(w, m.rgsw.top.b_parts, true);
        w = put_parts(w, m.rgsw.bottom.a_parts, true);
        put_parts(w, m.rgsw.bottom.b_parts, true);
        return cq;
    } catch (...) {
        return nullptr;
    }
}

}  // namespace

extern "C" {

int ipir_geometry(size_t n_entries, size_t entry_bytes, size_t db_rows_or_zero, ipir_params* out) {
    if (!out) return 1;
    try {
        const PublicParams pp = setup_seeded(n_entries, entry_bytes, db_rows_or_zero ? db_rows_or_zero : 32768,
                                             std::array<uint8_t, SEED_BYTES>{}.data());
        out->n_entries = pp.N_entries;
        out->entry_bytes = pp.w;
        out->db_rows = pp.db_rows;
        out->db_cols = pp.db_cols;
        out->n_packed = pp.n_packed;
        out->interp_d = pp.D;
        out->num_cts = pp.num_cts;
        out->ring_n = N;
        out->d_eff = (size_t)D_EFF;
        std::memset(out->seed, 0, SEED_BYTES);
        return 0;
    } catch (...) {
        return 1;
    }
}

size_t ipir_query_u64s(const ipir_params* p) { return params_ok(p) ? 2 * p->db_rows + 6 * p->d_eff * 2 * p->ring_n : 0; }
size_t ipir_response_u64s(const ipir_params* p) { return params_ok(p) ? p->num_cts * 4 * p->ring_n : 0; }
size_t ipir_entry_slots(const ipir_params* p) { return params_ok(p) ? (p->entry_bytes * 8 + P_BITS - 1) / P_BITS : 0; }

int ipir_query_view_from_flat(const ipir_params* p, const uint64_t* flat, ipir_query_view* v) {
    if (!params_ok(p) || !flat || !v) return 1;
    const size_t blk = p->d_eff * 2 * p->ring_n;
    v->lwe_b_limb0 = flat;
    v->lwe_b_limb1 = flat + p->db_rows;
    v->ksk5_b = flat + 2 * p->db_rows;
    v->kskneg1_b = flat + 2 * p->db_rows + blk;
    v->rgsw = flat + 2 * p->db_rows + 2 * blk;
    return 0;
}

ipir_client_query* ipir_query_build(const ipir_params* p, uint64_t idx, uint64_t* out) {
    return build(p, idx, out, nullptr);
}

ipir_client_query* ipir_query_build_seeded(const ipir_params* p, uint64_t idx, uint64_t rng_seed, uint64_t* out) {
    return build(p, idx, out, &rng_seed);
}

int ipir_extract(const ipir_params* p, const ipir_client_query* q, const uint64_t* resp, uint16_t* out) {
    if (!params_ok(p) || !q || !resp || !out) return 1;
    try {
        const PublicParams pp = pp_of(p);
        const std::vector<uint16_t> s = extract(pp, q->st, cts_from_flat(pp.num_cts, resp));
        if (s.size() != ipir_entry_slots(p)) return 2;
        std::memcpy(out, s.data(), s.size() * 2);
        return 0;
    } catch (...) {
        return 3;
    }
}

void ipir_client_query_destroy(ipir_client_query* q) { delete q; }

size_t ipir_query_packed_bytes(const ipir_params* p) { return params_ok(p) ? (query_values(p) * BETA + 7) / 8 : 0; }
size_t ipir_response_packed_bytes(const ipir_params* p) { return params_ok(p) ? (response_values(p) * BETA + 7) / 8 : 0; }

int ipir_query_pack(const ipir_params* p, const uint64_t* flat, uint8_t* out) {
    if (!params_ok(p) || !flat || !out) return 1;
    BitOut w(out);
    walk_query(p, [&](size_t i0, size_t i1) { w.put(garner(flat[i0], flat[i1]), BETA); });
    w.flush();
    return 0;
}

int ipir_query_unpack(const ipir_params* p, const uint8_t* in, uint64_t* flat) {
    if (!params_ok(p) || !in || !flat) return 1;
    BitIn r(in);
    walk_query(p, [&](size_t i0, size_t i1) { const uint64_t v = r.get(BETA); flat[i0] = v % Q0; flat[i1] = v % Q1; });
    return 0;
}

int ipir_response_pack(const ipir_params* p, const uint64_t* flat, uint8_t* out) {
    if (!params_ok(p) || !flat || !out) return 1;
    BitOut w(out);
    walk_response(p, [&](size_t i0, size_t i1) { w.put(garner(flat[i0], flat[i1]), BETA); });
    w.flush();
    return 0;
}

int ipir_response_unpack(const ipir_params* p, const uint8_t* in, uint64_t* flat) {
    if (!params_ok(p) || !in || !flat) return 1;
    BitIn r(in);
    walk_response(p, [&](size_t i0, size_t i1) { const uint64_t v = r.get(BETA); flat[i0] = v % Q0; flat[i1] = v % Q1; });
    return 0;
}

size_t ipir_response_compressed_bytes(const ipir_params* p) {
    return params_ok(p) ? p->num_cts * (p->ring_n * (28 + 20)) / 8 : 0;
}

int ipir_response_compress(const ipir_params* p, const uint64_t* flat, uint8_t* out) {
    if (!params_ok(p) || !flat || !out) return 1;
    try {
        BitOut w(out);
        for (const RlweCt& ct : cts_from_flat(p->num_cts, flat)) {
            const CompressedCt c = compress_ct(ct);
            for (size_t i = 0; i < N; i++) w.put(c.a[i], 28);
            for (size_t i = 0; i < N; i++) w.put(c.b[i], 20);
            w.flush();
        }
        return 0;
    } catch (...) {
        return 2;
    }
}

int ipir_extract_compressed(const ipir_params* p, const ipir_client_query* q, const uint8_t* in, uint16_t* out) {
    if (!params_ok(p) || !q || !in || !out) return 1;
    try {
        const PublicParams pp = pp_of(p);
        std::vector<CompressedCt> cts(pp.num_cts);
        BitIn r(in);
        for (auto& c : cts) {
            c.a.resize(N);
            c.b.resize(N);
            for (size_t i = 0; i < N; i++) c.a[i] = (uint32_t)r.get(28);
            for (size_t i = 0; i < N; i++) c.b[i] = (uint32_t)r.get(20);
        }
        const std::vector<uint16_t> s = extract_compressed(pp, q->st, cts);
        if (s.size() != ipir_entry_slots(p)) return 2;
        std::memcpy(out, s.data(), s.size() * 2);
        return 0;
    } catch (...) {
        return 3;
    }
}

}  // extern "C"
