// CPU ring arithmetic for the client and for the server's host-side table
// builders: geometry (setup), NTT tables, negacyclic NTT, RnsPoly,
// automorphisms, gadget decomposition, SHAKE-256 seed expansion, samplers.
//
// Conventions follow the reference so that ciphertexts interoperate:
// psi selection and twiddle order (src/ntt.cpp:53-96), CT/GS butterflies
// (src/ntt.cpp:107-149), decomposition (src/ring.cpp:205-246), expand_seed
// (src/ring.cpp:252-279), samplers (src/ring.cpp:285-334), setup
// (src/params.cpp:23-68).
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "params.h"
#include "ring.h"

namespace inspire {

namespace {

inline uint64_t addq(uint64_t a, uint64_t b, uint64_t q) { uint64_t s = a + b; return s >= q ? s - q : s; }
inline uint64_t subq(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + q - b; }
inline uint64_t mulq(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((unsigned __int128)a * b % q); }
// x * w mod q with w' = floor(w * 2^32 / q), x < 2^32: result in [0, q).
inline uint64_t shoup(uint64_t x, uint64_t w, uint64_t wp, uint64_t q) {
    uint64_t r = x * w - ((x * wp) >> 32) * q;
    return r >= q ? r - q : r;
}
uint64_t powq(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1;
    b %= q;
    for (; e; e >>= 1, b = mulq(b, b, q))
        if (e & 1) r = mulq(r, b, q);
    return r;
}
uint32_t rev_bits(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; i++) r = (r << 1) | ((x >> i) & 1);
    return r;
}

NttTables build(uint64_t q) {
    NttTables t{};
    t.q = q;
    const uint64_t e = (q - 1) / (2 * N);
    for (uint64_t g = 2; g < q; g++) {
        uint64_t psi = powq(g, e, q);
        if (powq(psi, N, q) == q - 1) { t.psi = psi; break; }
    }
    for (size_t i = 0; i < N; i++) t.fwd_twiddles[i] = powq(t.psi, rev_bits((uint32_t)i, 11), q);
    t.inv_twiddles[0] = 1;
    for (size_t i = 1; i < N; i++) t.inv_twiddles[i] = powq(t.fwd_twiddles[i], q - 2, q);
    t.inv_n = powq(N, q - 2, q);
    for (size_t i = 0; i < N; i++) {
        t.fwd_primes[i] = (t.fwd_twiddles[i] << 32) / q;
        t.inv_primes[i] = (t.inv_twiddles[i] << 32) / q;
    }
    t.inv_n_prime = (t.inv_n << 32) / q;
    return t;
}

inline uint64_t qof(int l) { return l ? Q1 : Q0; }

}  // namespace

// ---------------------------------------------------------------- geometry
static PublicParams geometry(size_t n_entries, size_t w, size_t db_rows) {
    PublicParams pp{};
    pp.N_entries = n_entries;
    pp.w = w;
    pp.D = 256;
    pp.db_rows = db_rows;
    if (db_rows == 0 || db_rows % N != 0)
        throw std::invalid_argument("setup: db_rows must be a positive multiple of N=2048");
    const size_t cpe = (w * 8 + P_BITS - 1) / P_BITS;
    const size_t cols = (n_entries * cpe + db_rows - 1) / db_rows;
    pp.db_cols = ((cols + N - 1) / N) * N;
    if (pp.db_cols == 0) pp.db_cols = N;
    pp.n_packed = pp.db_cols / N;
    pp.num_cts = pp.n_packed / pp.D > 1 ? pp.n_packed / pp.D : 1;
    if (cpe == 0 || N % cpe != 0)
        throw std::invalid_argument("setup: ceil(8*w/15) must divide N=2048");
    const size_t da = pp.D < pp.n_packed ? pp.D : pp.n_packed;
    if ((2 * N) % da != 0)
        throw std::invalid_argument("setup: n_packed must be a power of two");
    if (pp.n_packed > pp.D && pp.n_packed % pp.D != 0)
        throw std::invalid_argument("setup: n_packed must be a multiple of D");
    return pp;
}

PublicParams setup(size_t n_entries, size_t w) { return setup(n_entries, w, 32768); }

PublicParams setup(size_t n_entries, size_t w, size_t db_rows) {
    PublicParams pp = geometry(n_entries, w, db_rows);
    std::random_device rd;
    std::mt19937_64 rng(rd());
    for (auto& b : pp.seed) b = rng() & 0xFF;
    return pp;
}

PublicParams setup_seeded(size_t n_entries, size_t w, size_t db_rows, const uint8_t seed[SEED_BYTES]) {
    PublicParams pp = geometry(n_entries, w, db_rows);
    std::memcpy(pp.seed.data(), seed, SEED_BYTES);
    return pp;
}

// ---------------------------------------------------------------- NTT
const NttTables& ntt_tables(int limb) {
    static const NttTables t0 = build(Q0), t1 = build(Q1);
    return limb ? t1 : t0;
}

void ntt_forward(uint64_t* a, int limb) {
    const NttTables& t = ntt_tables(limb);
    const uint64_t q = t.q;
    for (size_t m = 1, st = N / 2; m < N; m <<= 1, st >>= 1)
        for (size_t g = 0; g < m; g++) {
            const uint64_t w = t.fwd_twiddles[m + g], wp = t.fwd_primes[m + g];
            uint64_t* x = a + 2 * g * st;
            for (size_t p = 0; p < st; p++) {
                uint64_t u = x[p], v = shoup(x[p + st], w, wp, q);
                x[p] = addq(u, v, q);
                x[p + st] = subq(u, v, q);
            }
        }
}

void ntt_inverse(uint64_t* a, int limb) {
    const NttTables& t = ntt_tables(limb);
    const uint64_t q = t.q;
    for (size_t m = N / 2, st = 1; m >= 1; m >>= 1, st <<= 1) {
        for (size_t g = 0; g < m; g++) {
            const uint64_t w = t.inv_twiddles[m + g], wp = t.inv_primes[m + g];
            uint64_t* x = a + 2 * g * st;
            for (size_t p = 0; p < st; p++) {
                uint64_t u = x[p], v = x[p + st];
                x[p] = addq(u, v, q);
                x[p + st] = shoup(subq(u, v, q), w, wp, q);
            }
        }
        if (m == 1) break;
    }
    for (size_t i = 0; i < N; i++) a[i] = shoup(a[i], t.inv_n, t.inv_n_prime, q);
}

// ---------------------------------------------------------------- RnsPoly
RnsPoly::RnsPoly() { for (auto& l : limbs) l.assign(N, 0); }
RnsPoly RnsPoly::zero() { return RnsPoly(); }
void RnsPoly::to_ntt() {
    if (is_ntt) return;
    ntt_forward(limbs[0].data(), 0); ntt_forward(limbs[1].data(), 1);
    is_ntt = true;
}
void RnsPoly::to_coeff() {
    if (!is_ntt) return;
    ntt_inverse(limbs[0].data(), 0); ntt_inverse(limbs[1].data(), 1);
    is_ntt = false;
}
void RnsPoly::add_inplace(const RnsPoly& o) {
    for (int l = 0; l < 2; l++) for (size_t i = 0; i < N; i++) limbs[l][i] = addq(limbs[l][i], o.limbs[l][i], qof(l));
}
void RnsPoly::sub_inplace(const RnsPoly& o) {
    for (int l = 0; l < 2; l++) for (size_t i = 0; i < N; i++) limbs[l][i] = subq(limbs[l][i], o.limbs[l][i], qof(l));
}
void RnsPoly::mul_inplace(const RnsPoly& o) {
    for (int l = 0; l < 2; l++) for (size_t i = 0; i < N; i++) limbs[l][i] = mulq(limbs[l][i], o.limbs[l][i], qof(l));
}
void RnsPoly::scalar_mul_inplace(uint64_t s0, uint64_t s1) {
    for (size_t i = 0; i < N; i++) { limbs[0][i] = mulq(limbs[0][i], s0, Q0); limbs[1][i] = mulq(limbs[1][i], s1, Q1); }
}
void RnsPoly::negate_inplace() {
    for (int l = 0; l < 2; l++) for (size_t i = 0; i < N; i++) if (limbs[l][i]) limbs[l][i] = qof(l) - limbs[l][i];
}
void RnsPoly::mul_sub_inplace(const RnsPoly& a, const RnsPoly& b) {
    for (int l = 0; l < 2; l++)
        for (size_t i = 0; i < N; i++) limbs[l][i] = subq(limbs[l][i], mulq(a.limbs[l][i], b.limbs[l][i], qof(l)), qof(l));
}
void RnsPoly::inner_product_sub(const RnsPoly* a, const RnsPoly* b, int d) {
    for (int l = 0; l < 2; l++)
        for (size_t i = 0; i < N; i++) {
            unsigned __int128 s = 0;
            for (int j = 0; j < d; j++) s += (unsigned __int128)a[j].limbs[l][i] * b[j].limbs[l][i];
            limbs[l][i] = subq(limbs[l][i], (uint64_t)(s % qof(l)), qof(l));
        }
}
void RnsPoly::set_coeff(size_t idx, uint64_t v0, uint64_t v1) { limbs[0][idx] = v0 % Q0; limbs[1][idx] = v1 % Q1; }

// ---------------------------------------------------------------- automorphisms
RnsPoly automorphism(const RnsPoly& f, uint64_t k) {
    RnsPoly r;
    for (size_t j = 0; j < N; j++) {
        const size_t d = (size_t)((j * k) % (2 * N));
        for (int l = 0; l < 2; l++) {
            if (d < N) r.limbs[l][d] = addq(r.limbs[l][d], f.limbs[l][j], qof(l));
            else r.limbs[l][d - N] = subq(r.limbs[l][d - N], f.limbs[l][j], qof(l));
        }
    }
    r.is_ntt = false;
    return r;
}
RnsPoly automorphism_ntt(const RnsPoly& f_ntt, uint64_t k) {
    RnsPoly f = f_ntt;
    f.is_ntt = true;
    f.to_coeff();
    RnsPoly g = automorphism(f, k);
    g.to_ntt();
    return g;
}

// ---------------------------------------------------------------- decomposition
std::vector<RnsPoly> decomp(const RnsPoly& f) {
    static const uint64_t i10 = powq(Q1 % Q0, Q0 - 2, Q0), i01 = powq(Q0 % Q1, Q1 - 2, Q1);
    std::vector<RnsPoly> out(D_EFF);
    for (size_t i = 0; i < N; i++) {
        uint64_t v = mulq(f.limbs[0][i], i10, Q0) * Q1 + mulq(f.limbs[1][i], i01, Q1) * Q0;
        if (v >= Q_FULL) v -= Q_FULL;
        int64_t r = v <= Q_FULL / 2 ? (int64_t)v : (int64_t)v - (int64_t)Q_FULL;
        int64_t dg[D_GSW];
        for (int j = 0; j < D_GSW; j++) {
            int64_t d = (int64_t)((uint64_t)r << (64 - BASE_LOG)) >> (64 - BASE_LOG);
            dg[j] = d;
            r = (r - d) >> BASE_LOG;
        }
        for (int j = 0; j < D_EFF; j++) {
            const int64_t d = dg[j + 1];
            out[j].limbs[0][i] = d >= 0 ? (uint64_t)d : (uint64_t)((int64_t)Q0 + d);
            out[j].limbs[1][i] = d >= 0 ? (uint64_t)d : (uint64_t)((int64_t)Q1 + d);
        }
    }
    return out;
}

// ---------------------------------------------------------------- SHAKE-256
namespace {
constexpr uint64_t KRC[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
    0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
    0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};
// rotation offsets and lane order of the combined rho+pi walk (FIPS 202 3.2.2-3.2.3)
constexpr int KROT[24] = {1, 3, 6, 10, 15, 21, 28, 36, 45, 55, 2, 14, 27, 41, 56, 8, 25, 43, 62, 18, 39, 61, 20, 44};
constexpr int KPI[24] = {10, 7, 11, 17, 18, 3, 5, 16, 8, 21, 24, 4, 15, 23, 19, 13, 12, 2, 20, 14, 22, 9, 6, 1};
inline uint64_t rotl(uint64_t x, int s) { return (x << s) | (x >> (64 - s)); }
void permute(uint64_t st[25]) {
    for (int round = 0; round < 24; round++) {
        uint64_t bc[5];
        for (int i = 0; i < 5; i++) bc[i] = st[i] ^ st[i + 5] ^ st[i + 10] ^ st[i + 15] ^ st[i + 20];
        for (int i = 0; i < 5; i++) {
            const uint64_t t = bc[(i + 4) % 5] ^ rotl(bc[(i + 1) % 5], 1);
            for (int j = 0; j < 25; j += 5) st[j + i] ^= t;
        }
        uint64_t cur = st[1];
        for (int i = 0; i < 24; i++) {
            const int j = KPI[i];
            const uint64_t tmp = st[j];
            st[j] = rotl(cur, KROT[i]);
            cur = tmp;
        }
        for (int j = 0; j < 25; j += 5) {
            uint64_t row[5];
            for (int i = 0; i < 5; i++) row[i] = st[j + i];
            for (int i = 0; i < 5; i++) st[j + i] = row[i] ^ (~row[(i + 1) % 5] & row[(i + 2) % 5]);
        }
        st[0] ^= KRC[round];
    }
}
}  // namespace

void shake256(const uint8_t* in, size_t inlen, uint8_t* out, size_t outlen) {
    constexpr size_t RATE = 136;
    uint64_t st[25] = {0};
    uint8_t* sb = reinterpret_cast<uint8_t*>(st);  // little-endian lane bytes
    size_t pos = 0;
    for (size_t i = 0; i < inlen; i++) {
        sb[pos++] ^= in[i];
        if (pos == RATE) { permute(st); pos = 0; }
    }
    sb[pos] ^= 0x1F;
    sb[RATE - 1] ^= 0x80;
    permute(st);
    pos = 0;
    for (size_t i = 0; i < outlen; i++) {
        if (pos == RATE) { permute(st); pos = 0; }
        out[i] = sb[pos++];
    }
}

RnsPoly expand_seed(const uint8_t seed[SEED_BYTES], const char* domain, uint64_t index) {
    std::vector<uint8_t> msg(seed, seed + SEED_BYTES);
    msg.insert(msg.end(), domain, domain + std::strlen(domain));
    for (int i = 0; i < 8; i++) msg.push_back((uint8_t)(index >> (8 * i)));
    std::vector<uint64_t> words(2 * N);
    shake256(msg.data(), msg.size(), reinterpret_cast<uint8_t*>(words.data()), 2 * N * 8);
    RnsPoly r;
    for (size_t i = 0; i < N; i++) { r.limbs[0][i] = words[i] % Q0; r.limbs[1][i] = words[N + i] % Q1; }
    return r;
}

// ---------------------------------------------------------------- samplers
RnsPoly sample_ternary(std::mt19937_64& rng) {
    RnsPoly r;
    std::uniform_int_distribution<int> dist(-1, 1);
    for (size_t i = 0; i < N; i++) {
        const int v = dist(rng);
        r.limbs[0][i] = v >= 0 ? (uint64_t)v : Q0 - 1;
        r.limbs[1][i] = v >= 0 ? (uint64_t)v : Q1 - 1;
    }
    return r;
}

RnsPoly sample_error(std::mt19937_64& rng) {
    // cumulative table of the half-weighted discrete Gaussian on [0, 19]
    static const std::vector<double> cdt = [] {
        std::vector<double> t;
        double c = 0;
        for (int x = 0; x <= 19; x++) {
            double p = std::exp(-(double)x * x / (2.0 * SIGMA * SIGMA));
            if (x == 0) p *= 0.5;
            c += p;
            t.push_back(c);
        }
        for (auto& v : t) v /= 2.0 * c;
        return t;
    }();
    RnsPoly r;
    std::uniform_real_distribution<double> ud(0.0, 0.5);
    std::uniform_int_distribution<int> sd(0, 1);
    for (size_t i = 0; i < N; i++) {
        const double u = ud(rng);
        int x = 0;
        while (x < (int)cdt.size() && u >= cdt[x]) x++;
        if (sd(rng) && x > 0) x = -x;
        r.limbs[0][i] = x >= 0 ? (uint64_t)x : Q0 - (uint64_t)(-x);
        r.limbs[1][i] = x >= 0 ? (uint64_t)x : Q1 - (uint64_t)(-x);
    }
    return r;
}

// ---------------------------------------------------------------- Galois helpers
uint64_t galois_power(uint64_t g, size_t k) {
    uint64_t r = 1, b = g % (2 * N);
    for (; k; k >>= 1, b = b * b % (2 * N))
        if (k & 1) r = r * b % (2 * N);
    return r;
}
const std::vector<uint64_t>& gen_powers() {
    static const std::vector<uint64_t> v = [] {
        std::vector<uint64_t> p(N);
        for (size_t i = 0; i < N; i++) p[i] = galois_power(AUTO_GEN, i);
        return p;
    }();
    return v;
}
uint64_t mod_inv(uint64_t a, uint64_t m) {
    int64_t r0 = (int64_t)a, r1 = (int64_t)m, s0 = 1, s1 = 0;
    while (r1) {
        int64_t qq = r0 / r1, t = r1;
        r1 = r0 - qq * r1; r0 = t;
        t = s1; s1 = s0 - qq * s1; s0 = t;
    }
    return (uint64_t)((s0 % (int64_t)m + (int64_t)m) % (int64_t)m);
}

void get_ntt_twiddles(int limb, std::vector<uint32_t>& fwd, std::vector<uint32_t>& inv, uint32_t& inv_n) {
    const NttTables& t = ntt_tables(limb);
    fwd.resize(N); inv.resize(N);
    for (size_t i = 0; i < N; i++) { fwd[i] = (uint32_t)t.fwd_twiddles[i]; inv[i] = (uint32_t)t.inv_twiddles[i]; }
    inv_n = (uint32_t)t.inv_n;
}
void get_ntt_twiddles_and_primes(int limb, std::vector<uint32_t>& fwd, std::vector<uint32_t>& fwdp,
                                 std::vector<uint32_t>& inv, std::vector<uint32_t>& invp,
                                 uint32_t& inv_n, uint32_t& inv_n_prime) {
    const NttTables& t = ntt_tables(limb);
    fwd.resize(N); fwdp.resize(N); inv.resize(N); invp.resize(N);
    for (size_t i = 0; i < N; i++) {
        fwd[i] = (uint32_t)t.fwd_twiddles[i]; fwdp[i] = (uint32_t)t.fwd_primes[i];
        inv[i] = (uint32_t)t.inv_twiddles[i]; invp[i] = (uint32_t)t.inv_primes[i];
    }
    inv_n = (uint32_t)t.inv_n; inv_n_prime = (uint32_t)t.inv_n_prime;
}

}  // namespace inspire
