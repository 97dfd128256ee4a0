// inspire_oracle.cpp — TEST INFRASTRUCTURE ONLY.
//
// A scalar CPU restatement of the InsPIRe server path and the client
// primitives it depends on. It exists to check the HIP implementation bit for
// bit; nothing in the product links, loads or calls it. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
//
// Every function cites the reference file:line whose behaviour it restates.
// All arithmetic is exact: 27-bit residues, products in u64, CRT in u128.
// Values are canonical residues in [0, q) unless stated otherwise.
//
// Layout conventions (match the reference device layouts):
//   RNS polynomial            : [limb][N] u64
//   ciphertext                : a then b, each [limb][N]
//   precomp D+/D- per group   : [limb][step][digit][N]
//   precomp D_final per group : [limb][digit][N]
//   precomp a per group       : [limb][N]
// Wire layouts follow src/capi.h:47-54 / :84-86 of the reference.
#include <cstdint>
#include <cstring>
#include <cmath>
#include <random>
#include <vector>
#include <array>
#include <algorithm>

namespace {

constexpr uint64_t N = 2048;
constexpr int LOGN = 11;
constexpr uint64_t Q0 = 94519297, Q1 = 95293441;
constexpr uint64_t QS[2] = {Q0, Q1};
constexpr uint64_t P = 65535;
constexpr int P_BITS = 15;
constexpr int D_GSW = 3, D_EFF = 2, BASE_LOG = 18;
constexpr uint64_t BASE = 1ull << BASE_LOG;
constexpr uint64_t AUTO_GEN = 5;
constexpr size_t SEED_BYTES = 64;
constexpr double SIGMA = 3.19;
constexpr unsigned __int128 QFULL = (unsigned __int128)Q0 * Q1;
constexpr uint64_t QF = (uint64_t)QFULL;
constexpr uint64_t DELTA = QF / P;
constexpr uint64_t HALF = N / 2;
constexpr uint64_t NSTEPS = HALF - 1;

typedef std::vector<uint64_t> Vec;

// ---------------------------------------------------------------- mod arith
template <int L> inline uint64_t mmul(uint64_t a, uint64_t b) { return (a * b) % QS[L]; }
inline uint64_t mmul_q(uint64_t a, uint64_t b, int l) { return l ? mmul<1>(a, b) : mmul<0>(a, b); }
inline uint64_t madd(uint64_t a, uint64_t b, uint64_t q) { uint64_t s = a + b; return s >= q ? s - q : s; }
inline uint64_t msub(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + q - b; }
uint64_t mpow(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1; b %= q;
    while (e) { if (e & 1) r = (unsigned __int128)r * b % q; b = (unsigned __int128)b * b % q; e >>= 1; }
    return r;
}
uint64_t minv(uint64_t a, uint64_t q) { return mpow(a, q - 2, q); }  // q prime

inline uint32_t bitrev(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; i++) if (x >> i & 1) r |= 1u << (bits - 1 - i);
    return r;
}

// ---------------------------------------------------------------- NTT tables
// psi = smallest g >= 2 with g^((q-1)/2N) of order 2N (src/ntt.cpp:53-69);
// fwd[i] = psi^BR(i), inv[i] = fwd[i]^-1, inv[0] = 1 (src/ntt.cpp:71-83);
// Shoup companions floor(w * 2^32 / q) (src/ntt.cpp:85-94).
struct Tables {
    uint64_t q, psi, inv_n, inv_n_p;
    uint64_t fwd[N], inv[N], fwdp[N], invp[N];
};
Tables make_tables(uint64_t q) {
    Tables t;
    t.q = q;
    uint64_t e = (q - 1) / (2 * N);
    for (uint64_t g = 2;; g++) {
        uint64_t psi = mpow(g, e, q);
        if (mpow(psi, N, q) == q - 1) { t.psi = psi; break; }
    }
    for (uint64_t i = 0; i < N; i++) t.fwd[i] = mpow(t.psi, bitrev((uint32_t)i, LOGN), q);
    t.inv[0] = 1;
    for (uint64_t i = 1; i < N; i++) t.inv[i] = minv(t.fwd[i], q);
    t.inv_n = minv(N, q);
    for (uint64_t i = 0; i < N; i++) { t.fwdp[i] = (t.fwd[i] << 32) / q; t.invp[i] = (t.inv[i] << 32) / q; }
    t.inv_n_p = (t.inv_n << 32) / q;
    return t;
}
const Tables& tables(int l) {
    static const Tables t0 = make_tables(Q0), t1 = make_tables(Q1);
    return l ? t1 : t0;
}

// Cooley-Tukey DIT forward, Gentleman-Sande inverse with the N^-1 scale
// (src/ntt.cpp:107-149). Output slot i holds f(psi^(2*BR(i)+1)).
template <int L> void ntt_fwd_t(uint64_t* a) {
    const Tables& t = tables(L);
    constexpr uint64_t q = QS[L];
    for (uint64_t m = 1; m < N; m <<= 1) {
        uint64_t st = N / (2 * m);
        for (uint64_t idx = 0; idx < N / 2; idx++) {
            uint64_t g = idx / st, p = idx % st, j = g * 2 * st + p, k = j + st;
            uint64_t u = a[j], v = mmul<L>(a[k], t.fwd[m + g]);
            a[j] = madd(u, v, q); a[k] = msub(u, v, q);
        }
    }
}
template <int L> void ntt_inv_t(uint64_t* a) {
    const Tables& t = tables(L);
    constexpr uint64_t q = QS[L];
    for (uint64_t m = N / 2; m >= 1; m >>= 1) {
        uint64_t st = N / (2 * m);
        for (uint64_t idx = 0; idx < N / 2; idx++) {
            uint64_t g = idx / st, p = idx % st, j = g * 2 * st + p, k = j + st;
            uint64_t u = a[j], v = a[k];
            a[j] = madd(u, v, q); a[k] = mmul<L>(msub(u, v, q), t.inv[m + g]);
        }
        if (m == 1) break;
    }
    for (uint64_t i = 0; i < N; i++) a[i] = mmul<L>(a[i], t.inv_n);
}
inline void ntt_fwd(uint64_t* a, int l) { l ? ntt_fwd_t<1>(a) : ntt_fwd_t<0>(a); }
inline void ntt_inv(uint64_t* a, int l) { l ? ntt_inv_t<1>(a) : ntt_inv_t<0>(a); }
// both limbs of an RNS polynomial [2][N]
inline void poly_ntt(uint64_t* p) { ntt_fwd(p, 0); ntt_fwd(p + N, 1); }
inline void poly_intt(uint64_t* p) { ntt_inv(p, 0); ntt_inv(p + N, 1); }

// ---------------------------------------------------------------- ring ops
// tau_k(f)(X) = f(X^k) mod X^N+1, coefficient form (src/ring.cpp:174-190).
void automorphism(const uint64_t* f, uint64_t k, uint64_t* out) {
    for (int l = 0; l < 2; l++) {
        uint64_t q = QS[l];
        for (uint64_t i = 0; i < N; i++) out[l * N + i] = 0;
        for (uint64_t j = 0; j < N; j++) {
            uint64_t d = (j * k) % (2 * N);
            if (d < N) out[l * N + d] = madd(out[l * N + d], f[l * N + j], q);
            else out[l * N + d - N] = msub(out[l * N + d - N], f[l * N + j], q);
        }
    }
}

// CRT lift of (a0 mod Q0, a1 mod Q1) to [0, Q) (src/ring.cpp:218-220).
inline uint64_t crt(uint64_t a0, uint64_t a1) {
    static const uint64_t i10 = minv(Q1 % Q0, Q0), i01 = minv(Q0 % Q1, Q1);
    unsigned __int128 x = (unsigned __int128)((unsigned __int128)a0 * i10 % Q0) * Q1
                        + (unsigned __int128)((unsigned __int128)a1 * i01 % Q1) * Q0;
    return (uint64_t)(x % QFULL);
}

// Gadget decomposition: CRT lift, center (val <= Q/2 stays positive),
// three signed base-2^18 digits by arithmetic shift, drop digit 0
// (src/ring.cpp:205-246). f coefficient form [2][N]; out [D_EFF][2][N].
void decomp(const uint64_t* f, uint64_t* out) {
    for (uint64_t i = 0; i < N; i++) {
        uint64_t val = crt(f[i], f[N + i]);
        int64_t r = (val <= QF / 2) ? (int64_t)val : (int64_t)val - (int64_t)QF;
        int64_t dg[D_GSW];
        for (int j = 0; j < D_GSW; j++) {
            int64_t d = (int64_t)((uint64_t)r << (64 - BASE_LOG)) >> (64 - BASE_LOG);
            dg[j] = d;
            r = (r - d) >> BASE_LOG;
        }
        for (int j = 0; j < D_EFF; j++) {
            int64_t d = dg[j + 1];
            for (int l = 0; l < 2; l++)
                out[(j * 2 + l) * N + i] = d >= 0 ? (uint64_t)d : QS[l] - (uint64_t)(-d);
        }
    }
}

// ---------------------------------------------------------------- SHAKE-256
// FIPS 202 Keccak-f[1600] sponge, rate 136 bytes, domain byte 0x1F.
const uint64_t RC[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
    0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
    0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};
inline uint64_t rol(uint64_t x, int s) { return s ? (x << s) | (x >> (64 - s)) : x; }
void keccak_f(uint64_t A[25]) {
    static const int rho[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43,
                                25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
    for (int rnd = 0; rnd < 24; rnd++) {
        uint64_t C[5], D[5], B[25];
        for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
        for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ rol(C[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) A[i] ^= D[i % 5];
        // B[y, 2x+3y] = rot(A[x,y], rho[x,y])
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                B[y + 5 * ((2 * x + 3 * y) % 5)] = rol(A[x + 5 * y], rho[x + 5 * y]);
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                A[x + 5 * y] = B[x + 5 * y] ^ (~B[(x + 1) % 5 + 5 * y] & B[(x + 2) % 5 + 5 * y]);
        A[0] ^= RC[rnd];
    }
}
void shake256(const uint8_t* in, size_t inlen, uint8_t* out, size_t outlen) {
    uint64_t A[25] = {0};
    const size_t rate = 136;
    uint8_t blk[rate];
    while (inlen >= rate) {
        for (size_t i = 0; i < rate / 8; i++) { uint64_t w; memcpy(&w, in + 8 * i, 8); A[i] ^= w; }
        keccak_f(A); in += rate; inlen -= rate;
    }
    memset(blk, 0, rate); memcpy(blk, in, inlen);
    blk[inlen] ^= 0x1F; blk[rate - 1] ^= 0x80;
    for (size_t i = 0; i < rate / 8; i++) { uint64_t w; memcpy(&w, blk + 8 * i, 8); A[i] ^= w; }
    keccak_f(A);
    while (outlen) {
        size_t n = outlen < rate ? outlen : rate;
        memcpy(out, A, n);  // little-endian host
        out += n; outlen -= n;
        if (outlen) keccak_f(A);
    }
}

// expand_seed = SHAKE-256(seed || domain || le64(index)) -> 2N u64 words,
// limb0[i] = w[i] % Q0, limb1[i] = w[N+i] % Q1 (src/ring.cpp:252-279).
void expand_seed(const uint8_t* seed, const char* domain, uint64_t index, uint64_t* out) {
    std::vector<uint8_t> in(seed, seed + SEED_BYTES);
    in.insert(in.end(), domain, domain + strlen(domain));
    for (int i = 0; i < 8; i++) in.push_back((uint8_t)(index >> (8 * i)));
    std::vector<uint64_t> raw(2 * N);
    shake256(in.data(), in.size(), (uint8_t*)raw.data(), 2 * N * 8);
    for (uint64_t i = 0; i < N; i++) { out[i] = raw[i] % Q0; out[N + i] = raw[N + i] % Q1; }
}

// ---------------------------------------------------------------- samplers
// Uniform ternary and CDT Gaussian over libstdc++ <random>, consuming the
// engine in the reference's order (src/ring.cpp:285-334).
void sample_ternary(std::mt19937_64& rng, uint64_t* out) {
    std::uniform_int_distribution<int> dist(-1, 1);
    for (uint64_t i = 0; i < N; i++) {
        int v = dist(rng);
        out[i] = v >= 0 ? (uint64_t)v : Q0 - 1;
        out[N + i] = v >= 0 ? (uint64_t)v : Q1 - 1;
    }
}
void sample_error(std::mt19937_64& rng, uint64_t* out) {
    static const std::vector<double> cdt = [] {
        std::vector<double> t;
        double cum = 0;
        for (int x = 0; x <= 19; x++) {
            double p = std::exp(-(double)x * x / (2.0 * SIGMA * SIGMA));
            if (x == 0) p *= 0.5;
            cum += p; t.push_back(cum);
        }
        double tot = 2.0 * cum;
        for (auto& v : t) v /= tot;
        return t;
    }();
    std::uniform_real_distribution<double> ud(0.0, 0.5);
    std::uniform_int_distribution<int> sd(0, 1);
    for (uint64_t i = 0; i < N; i++) {
        double u = ud(rng);
        int x = 0;
        for (size_t j = 0; j < cdt.size(); j++) { if (u >= cdt[j]) x = (int)j + 1; else break; }
        if (sd(rng) && x > 0) x = -x;
        out[i] = x >= 0 ? (uint64_t)x : Q0 - (uint64_t)(-x);
        out[N + i] = x >= 0 ? (uint64_t)x : Q1 - (uint64_t)(-x);
    }
}

// pointwise helpers on [2][N] polys
void p_add(uint64_t* x, const uint64_t* y) { for (int l = 0; l < 2; l++) for (uint64_t i = 0; i < N; i++) x[l*N+i] = madd(x[l*N+i], y[l*N+i], QS[l]); }
void p_sub(uint64_t* x, const uint64_t* y) { for (int l = 0; l < 2; l++) for (uint64_t i = 0; i < N; i++) x[l*N+i] = msub(x[l*N+i], y[l*N+i], QS[l]); }
void p_mul(uint64_t* x, const uint64_t* y) { for (int l = 0; l < 2; l++) for (uint64_t i = 0; i < N; i++) x[l*N+i] = mmul_q(x[l*N+i], y[l*N+i], l); }
void p_scal(uint64_t* x, uint64_t s) { for (int l = 0; l < 2; l++) for (uint64_t i = 0; i < N; i++) x[l*N+i] = mmul_q(x[l*N+i], s % QS[l], l); }

// ---------------------------------------------------------------- client
// a*s in coefficient form for a (coeff) and s_ntt
void mul_as(const uint64_t* a_coeff, const uint64_t* s_ntt, uint64_t* out) {
    memcpy(out, a_coeff, 2 * N * 8); poly_ntt(out); p_mul(out, s_ntt); poly_intt(out);
}
// KskGen_b: b_j = a_j*s + e_j + tau_k(s)*B^(j+1), coefficient form
// (src/crypto.cpp:108-145). out [D_EFF][2][N].
void ksk_gen_b(const uint64_t* s, uint64_t k, const uint8_t* seed, std::mt19937_64& rng, uint64_t* out) {
    Vec sn(s, s + 2 * N); poly_ntt(sn.data());
    Vec sa(2 * N); automorphism(s, k, sa.data());
    Vec a(2 * N), e(2 * N), b(2 * N), ss(2 * N);
    for (int j = 0; j < D_EFF; j++) {
        expand_seed(seed, "ksk", k * 100 + (j + 1), a.data());
        sample_error(rng, e.data());
        mul_as(a.data(), sn.data(), b.data());
        uint64_t Bj = 1; for (int t = 0; t <= j; t++) Bj *= BASE;
        ss = sa; p_scal(ss.data(), Bj);
        p_add(b.data(), e.data()); p_add(b.data(), ss.data());
        memcpy(out + (size_t)j * 2 * N, b.data(), 2 * N * 8);
    }
}
// LWE.Enc_b over n_r = ell/N RLWE rows (src/crypto.cpp:65-101). out [2][ell].
void lwe_enc_b(const uint64_t* s, uint64_t i_star, uint64_t ell, const uint8_t* seed,
               std::mt19937_64& rng, uint64_t* b0, uint64_t* b1) {
    Vec sn(s, s + 2 * N); poly_ntt(sn.data());
    Vec a(2 * N), e(2 * N), b(2 * N);
    for (uint64_t r = 0; r < ell / N; r++) {
        expand_seed(seed, "rlwe", r, a.data());
        sample_error(rng, e.data());
        mul_as(a.data(), sn.data(), b.data());
        p_add(b.data(), e.data());
        if (r == i_star / N) {
            uint64_t pos = i_star % N;
            b[pos] = madd(b[pos], DELTA % Q0, Q0);
            b[N + pos] = madd(b[N + pos], DELTA % Q1, Q1);
        }
        for (uint64_t j = 0; j < N; j++) { b0[r * N + j] = b[j]; b1[r * N + j] = b[N + j]; }
    }
}
// RGSW.Enc_b (src/crypto.cpp:151-213). out wire layout
// [top_a | top_b | bot_a | bot_b][digit][limb][coeff], NTT form.
void rgsw_enc_b(const uint64_t* s, const uint64_t* mu, const uint8_t* seed, std::mt19937_64& rng, uint64_t* out) {
    const size_t blk = (size_t)D_EFF * 2 * N;
    Vec sn(s, s + 2 * N); poly_ntt(sn.data());
    Vec mun(mu, mu + 2 * N); poly_ntt(mun.data());
    Vec a(2 * N), e(2 * N), b(2 * N), t(2 * N), bm(2 * N);
    for (int j = 0; j < D_EFF; j++) {
        uint64_t Bj = 1; for (int k = 0; k <= j; k++) Bj *= BASE;
        t.assign(s, s + 2 * N); p_scal(t.data(), Bj); poly_ntt(t.data()); p_mul(t.data(), mun.data()); poly_intt(t.data());
        bm.assign(mu, mu + 2 * N); p_scal(bm.data(), Bj);
        // top row: b = a*s + e - s*B^j*mu
        expand_seed(seed, "rgsw0", j + 1, a.data());
        sample_error(rng, e.data());
        mul_as(a.data(), sn.data(), b.data());
        p_add(b.data(), e.data()); p_sub(b.data(), t.data());
        poly_ntt(a.data()); poly_ntt(b.data());
        memcpy(out + 0 * blk + (size_t)j * 2 * N, a.data(), 2 * N * 8);
        memcpy(out + 1 * blk + (size_t)j * 2 * N, b.data(), 2 * N * 8);
        // bottom row: b = a*s + e + B^j*mu
        expand_seed(seed, "rgsw1", j + 1, a.data());
        sample_error(rng, e.data());
        mul_as(a.data(), sn.data(), b.data());
        p_add(b.data(), e.data()); p_add(b.data(), bm.data());
        poly_ntt(a.data()); poly_ntt(b.data());
        memcpy(out + 2 * blk + (size_t)j * 2 * N, a.data(), 2 * N * 8);
        memcpy(out + 3 * blk + (size_t)j * 2 * N, b.data(), 2 * N * 8);
    }
}

// RLWE.Dec with CRT, centering (val > Q/2 is negative) and the double
// rounding of src/crypto.cpp:19-58. ct NTT form [a|b][2][N]; s_ntt [2][N].
void rlwe_dec(const uint64_t* s_ntt, const uint64_t* ct, uint64_t* m) {
    Vec v(ct + 2 * N, ct + 4 * N), as(ct, ct + 2 * N);
    p_mul(as.data(), s_ntt); p_sub(v.data(), as.data()); poly_intt(v.data());
    for (uint64_t i = 0; i < N; i++) {
        uint64_t val = crt(v[i], v[N + i]);
        int64_t c = (val > QF / 2) ? (int64_t)(val - QF) : (int64_t)val;
        double ratio = (double)c * (double)P / (double)QF;
        int64_t r = (int64_t)std::round(ratio);
        m[i] = (uint64_t)(((r % (int64_t)P) + (int64_t)P) % (int64_t)P);
    }
}

// ---------------------------------------------------------------- ext prod
// ExtProd (src/crypto.cpp:219-257): decompose a and b (coefficient form),
// NTT the digits, res = sum_j da_j*top_j + db_j*bot_j. rgsw wire layout.
void ext_prod(const uint64_t* rgsw, const uint64_t* ct, uint64_t* res) {
    const size_t blk = (size_t)D_EFF * 2 * N;
    Vec ac(ct, ct + 2 * N), bc(ct + 2 * N, ct + 4 * N);
    poly_intt(ac.data()); poly_intt(bc.data());
    Vec da(D_EFF * 2 * N), db(D_EFF * 2 * N);
    decomp(ac.data(), da.data()); decomp(bc.data(), db.data());
    for (int j = 0; j < D_EFF; j++) { poly_ntt(&da[j * 2 * N]); poly_ntt(&db[j * 2 * N]); }
    for (int l = 0; l < 2; l++) {
        uint64_t q = QS[l];
        for (uint64_t i = 0; i < N; i++) {
            uint64_t sa = 0, sb = 0;
            for (int j = 0; j < D_EFF; j++) {
                size_t o = (size_t)j * 2 * N + l * N + i;
                sa = madd(sa, mmul_q(da[o], rgsw[0 * blk + o], l), q);
                sb = madd(sb, mmul_q(da[o], rgsw[1 * blk + o], l), q);
                sa = madd(sa, mmul_q(db[o], rgsw[2 * blk + o], l), q);
                sb = madd(sb, mmul_q(db[o], rgsw[3 * blk + o], l), q);
            }
            res[l * N + i] = sa; res[2 * N + l * N + i] = sb;
        }
    }
}

// ---------------------------------------------------------------- server
const uint32_t* br_table() {
    static const std::vector<uint32_t> t = [] {
        std::vector<uint32_t> v(N);
        for (uint32_t i = 0; i < N; i++) v[i] = bitrev(i, LOGN);
        return v;
    }();
    return t.data();
}
inline uint32_t perm_idx(uint64_t t, uint32_t i) {
    // Slot permutation induced by tau_t on the NTT domain
    // (src/gpu_preprocess.cu:116, :181-187): BR(((t*(2BR(i)+1) mod 2N) - 1)/2).
    const uint32_t* br = br_table();
    uint64_t e = (t * (2 * (uint64_t)br[i] + 1)) % (2 * N);
    return br[(e - 1) / 2];
}
// 5^k mod 2N
inline uint64_t pow5(uint64_t k) {
    static const std::vector<uint64_t> t = [] {
        std::vector<uint64_t> v(N); uint64_t x = 1;
        for (uint64_t i = 0; i < N; i++) { v[i] = x; x = x * 5 % (2 * N); }
        return v;
    }();
    return t[k % N];
}

struct Geo { uint64_t R, C, G, D, Da, T, nr; };
Geo geo_of(uint64_t R, uint64_t C, uint64_t D) {
    Geo g; g.R = R; g.C = C; g.G = C / N; g.D = D; g.Da = std::min(D, g.G);
    g.T = std::max<uint64_t>(1, g.G / D); g.nr = R / N; return g;
}

} // namespace

extern "C" {

int orc_ring_n() { return (int)N; }

// ---- primitives -----------------------------------------------------------
void orc_ntt_tables(int limb, uint64_t* fwd, uint64_t* inv, uint64_t* fwdp, uint64_t* invp, uint64_t* scal4) {
    const Tables& t = tables(limb);
    for (uint64_t i = 0; i < N; i++) { fwd[i] = t.fwd[i]; inv[i] = t.inv[i]; fwdp[i] = t.fwdp[i]; invp[i] = t.invp[i]; }
    scal4[0] = t.psi; scal4[1] = t.inv_n; scal4[2] = t.inv_n_p; scal4[3] = t.q;
}
void orc_ntt_fwd(uint64_t* a, int limb) { ntt_fwd(a, limb); }
void orc_ntt_inv(uint64_t* a, int limb) { ntt_inv(a, limb); }
void orc_automorphism(const uint64_t* f, uint64_t k, uint64_t* out) { automorphism(f, k, out); }
void orc_decomp(const uint64_t* f, uint64_t* out) { decomp(f, out); }
uint64_t orc_crt(uint64_t a0, uint64_t a1) { return crt(a0, a1); }
void orc_shake256(const uint8_t* in, size_t inlen, uint8_t* out, size_t outlen) { shake256(in, inlen, out, outlen); }
void orc_expand_seed(const uint8_t* seed, const char* domain, uint64_t index, uint64_t* out) { expand_seed(seed, domain, index, out); }
void orc_ext_prod(const uint64_t* rgsw, const uint64_t* ct, uint64_t* res) { ext_prod(rgsw, ct, res); }
void orc_rlwe_dec(const uint64_t* s_ntt, const uint64_t* ct, uint64_t* m) { rlwe_dec(s_ntt, ct, m); }
uint32_t orc_perm_idx(uint64_t t, uint32_t i) { return perm_idx(t, i); }

// Horner (src/crypto.cpp:264-279): acc = ct[L-1]; acc = ext_prod(acc)+ct[i].
void orc_horner(const uint64_t* cts, uint64_t L, const uint64_t* rgsw, uint64_t* out) {
    memcpy(out, cts + (L - 1) * 4 * N, 4 * N * 8);
    Vec t(4 * N);
    for (int64_t i = (int64_t)L - 2; i >= 0; i--) {
        ext_prod(rgsw, out, t.data());
        p_add(t.data(), cts + (size_t)i * 4 * N);
        p_add(t.data() + 2 * N, cts + (size_t)i * 4 * N + 2 * N);
        memcpy(out, t.data(), 4 * N * 8);
    }
}

// ---- geometry (src/params.cpp:23-68) ---------------------------------------
// out: [db_rows, db_cols, n_packed, D, num_cts]. Returns 0, or 1..3 for the
// three rejections in src/params.cpp:46-60.
int orc_setup(uint64_t n_entries, uint64_t w, uint64_t db_rows, uint64_t* out) {
    uint64_t cpe = (w * 8 + P_BITS - 1) / P_BITS;
    uint64_t cols = (n_entries * cpe + db_rows - 1) / db_rows;
    uint64_t C = ((cols + N - 1) / N) * N;
    if (C == 0) C = N;
    uint64_t G = C / N, D = 256, T = std::max<uint64_t>(1, G / D);
    out[0] = db_rows; out[1] = C; out[2] = G; out[3] = D; out[4] = T;
    if (cpe == 0 || N % cpe != 0) return 1;
    uint64_t da = std::min(D, G);
    if ((2 * N) % da != 0) return 2;
    if (G > D && G % D != 0) return 3;
    return 0;
}

// ---- client query / extract (src/protocol.cpp:12-71) ----------------------
// Seeded variant of query(): mt19937_64(rng_seed) replaces random_device.
// out_query: flat wire layout (src/capi.h:109-116). out_state: s coeff [2N],
// s_ntt [2N], k_star, ell_star.
void orc_query(uint64_t R, uint64_t C, uint64_t D, uint64_t w, const uint8_t* seed,
               uint64_t idx, uint64_t rng_seed, uint64_t* out_query, uint64_t* out_state) {
    Geo g = geo_of(R, C, D);
    std::mt19937_64 rng(rng_seed);
    uint64_t cpe = (w * 8 + P_BITS - 1) / P_BITS;
    uint64_t i_star = idx % R, col = (idx / R) * cpe;
    uint64_t j_star = (col / N) % g.Da, k_star = col / (N * g.Da), ell_star = col % N;
    Vec s(2 * N); sample_ternary(rng, s.data());
    const size_t blk = (size_t)D_EFF * 2 * N;
    uint64_t* w0 = out_query + 2 * R;
    ksk_gen_b(s.data(), AUTO_GEN, seed, rng, w0);
    ksk_gen_b(s.data(), 2 * N - 1, seed, rng, w0 + blk);
    lwe_enc_b(s.data(), i_star, R, seed, rng, out_query, out_query + R);
    Vec mu(2 * N, 0);
    uint64_t e = (j_star * 2 * N) / g.Da;
    if (e < N) { mu[e] = 1; mu[N + e] = 1; }
    else { mu[e - N] = Q0 - 1; mu[N + e - N] = Q1 - 1; }
    rgsw_enc_b(s.data(), mu.data(), seed, rng, w0 + 2 * blk);
    memcpy(out_state, s.data(), 2 * N * 8);
    Vec sn = s; poly_ntt(sn.data());
    memcpy(out_state + 2 * N, sn.data(), 2 * N * 8);
    out_state[4 * N] = k_star; out_state[4 * N + 1] = ell_star;
}
// extract(): decrypt resp[k_star], return slots ell*..ell*+cpe-1.
int orc_extract(uint64_t T, uint64_t w, const uint64_t* state, const uint64_t* resp, uint16_t* slots) {
    uint64_t cpe = (w * 8 + P_BITS - 1) / P_BITS;
    uint64_t k = state[4 * N], ell = state[4 * N + 1];
    if (k >= T) return 1;
    Vec m(N);
    rlwe_dec(state + 2 * N, resp + k * 4 * N, m.data());
    for (uint64_t c = 0; c < cpe && ell + c < N; c++) slots[c] = (uint16_t)(m[ell + c] % P);
    return 0;
}

// ---- preprocess ---------------------------------------------------------
// Encode: per row and response-ct block, out[m] = D^-1 sum_d x[d] X^(-2N d m/D)
// over Z_P[X]/(X^N+1) (src/gpu_preprocess.cu:931-1034, restated as the
// direct sum rather than the iterative butterflies). Applied to the listed
// rows only; in place on the row-major u16 DB.
void orc_encode_rows(uint16_t* db, uint64_t R, uint64_t C, uint64_t D, const uint64_t* rows, uint64_t nrows) {
    Geo g = geo_of(R, C, D);
    if (g.Da <= 1) return;
    uint64_t Dinv = 0;
    for (uint64_t v = 1; v < P; v++) if (v * g.Da % P == 1) { Dinv = v; break; }
    std::vector<uint64_t> out(g.Da * N);
    for (uint64_t rr = 0; rr < nrows; rr++) {
        uint16_t* row = db + rows[rr] * C;
        for (uint64_t ct = 0; ct < g.G / g.Da; ct++) {
            uint16_t* base = row + ct * g.Da * N;
            std::fill(out.begin(), out.end(), 0);
            for (uint64_t m = 0; m < g.Da; m++)
                for (uint64_t d = 0; d < g.Da; d++) {
                    // multiply x[d] by X^e, e = -(2N*d*m/Da) mod 2N
                    uint64_t e = (2 * N - (2 * N * ((d * m) % g.Da)) / g.Da) % (2 * N);
                    for (uint64_t j = 0; j < N; j++) {
                        uint64_t t = j + e, v = base[d * N + j];
                        bool neg = false;
                        while (t >= N) { t -= N; neg = !neg; }
                        uint64_t& o = out[m * N + t];
                        o = neg ? (o + P - v) % P : (o + v) % P;
                    }
                }
            for (uint64_t k = 0; k < g.Da * N; k++) base[k] = (uint16_t)(out[k] * Dinv % P);
        }
    }
}

// CRS rows: A_r = NTT(tau_{2N-1}(expand_seed(seed,"rlwe",r)))
// (src/gpu_protocol.cu:567-583). out [nr][2][N].
void orc_crs(const uint8_t* seed, uint64_t nr, uint64_t* out) {
    Vec a(2 * N);
    for (uint64_t r = 0; r < nr; r++) {
        expand_seed(seed, "rlwe", r, a.data());
        uint64_t* o = out + r * 2 * N;
        for (int l = 0; l < 2; l++) {
            o[l * N] = a[l * N];
            for (uint64_t i = 1; i < N; i++) o[l * N + i] = a[l * N + N - i] ? QS[l] - a[l * N + N - i] : 0;
        }
        poly_ntt(o);
    }
}

// a-side: acc_j = sum_r NTT(db[rN..rN+N-1][gN+j]) * A_r; a32[j][c][k] =
// acc_j[c][N-1-k] * N^-1 (src/gpu_preprocess.cu:347-410). out [N][2][N].
void orc_a32(const uint16_t* db, uint64_t R, uint64_t C, uint64_t g, const uint64_t* crs, uint64_t* out) {
    uint64_t nr = R / N;
    Vec acc(2 * N), col(2 * N);
    for (uint64_t j = 0; j < N; j++) {
        std::fill(acc.begin(), acc.end(), 0);
        for (uint64_t r = 0; r < nr; r++) {
            for (uint64_t i = 0; i < N; i++) col[i] = col[N + i] = db[(r * N + i) * C + g * N + j];
            poly_ntt(col.data());
            p_mul(col.data(), crs + r * 2 * N);
            p_add(acc.data(), col.data());
        }
        for (int l = 0; l < 2; l++)
            for (uint64_t k = 0; k < N; k++)
                out[j * 2 * N + l * N + k] = mmul_q(acc[l * N + N - 1 - k], tables(l).inv_n, l);
    }
}

// Ring embed, direct formula (src/gpu_preprocess.cu:39-92) followed by the
// per-k automorphism permutation (src/gpu_preprocess.cu:158-194):
// acc_f[k][c][i] = sum_j a32[j][c][i]*NTT(X^(j*5^((N-k)%N) mod 2N))[c][i],
// acc_c uses exponent 2N - that; a_post[k][c][i] = acc[k][c][perm(i)].
// Evaluated only for the (k, c, i) triples listed (slow: N MACs each).
void orc_ring_embed_direct(const uint64_t* a32, const uint64_t* kci, uint64_t n, uint64_t* out_fwd, uint64_t* out_conj) {
    // psi^e for all odd/even exponents per limb
    static std::vector<uint64_t> pw[2];
    if (pw[0].empty())
        for (int l = 0; l < 2; l++) {
            pw[l].resize(2 * N); uint64_t v = 1;
            for (uint64_t e = 0; e < 2 * N; e++) { pw[l][e] = v; v = mmul_q(v, tables(l).psi, l); }
        }
    for (uint64_t t = 0; t < n; t++) {
        uint64_t k = kci[3 * t], c = kci[3 * t + 1], i = kci[3 * t + 2];
        uint64_t tf = pow5(k), tc = (2 * N - tf) % (2 * N);
        uint32_t pf = perm_idx(tf, (uint32_t)i), pc = perm_idx(tc, (uint32_t)i);
        uint64_t gk = pow5((N - k) % N);
        uint64_t sf = 0, sc = 0, q = QS[c];
        for (uint64_t j = 0; j < N; j++) {
            uint64_t idx = (j * gk) % (2 * N), idc = (2 * N - idx) % (2 * N);
            // NTT(X^t)[p] = psi^((2BR(p)+1) t)
            uint64_t ef = ((2 * (uint64_t)bitrev(pf, LOGN) + 1) * idx) % (2 * N);
            uint64_t ec = ((2 * (uint64_t)bitrev(pc, LOGN) + 1) * idc) % (2 * N);
            sf = (sf + mmul_q(a32[j * 2 * N + c * N + pf], pw[c][ef], c)) % q;
            sc = (sc + mmul_q(a32[j * 2 * N + c * N + pc], pw[c][ec], c)) % q;
        }
        out_fwd[t] = sf; out_conj[t] = sc;
    }
}

// Ring embed by a length-N NTT along j (the evaluation form of the direct
// sum): F[c][p] = NTT_c(a32[.][c][p]); a_post_fwd[k][c][i] = F[c][pf][i],
// a_post_conj[k][c][i] = F[c][pc][i], pf = perm_{5^k}(i), pc = perm_{-5^k}(i).
// out [half][2][N] each.
void orc_ring_embed(const uint64_t* a32, uint64_t* a_fwd, uint64_t* a_conj) {
    Vec F(2 * N * N), v(N);
    for (int c = 0; c < 2; c++)
        for (uint64_t p = 0; p < N; p++) {
            for (uint64_t j = 0; j < N; j++) v[j] = a32[j * 2 * N + c * N + p];
            ntt_fwd(v.data(), c);
            memcpy(&F[(c * N + p) * N], v.data(), N * 8);
        }
    for (uint64_t k = 0; k < HALF; k++) {
        uint64_t tf = pow5(k), tc = (2 * N - tf) % (2 * N);
        for (int c = 0; c < 2; c++)
            for (uint32_t i = 0; i < N; i++) {
                a_fwd[k * 2 * N + c * N + i] = F[(c * N + perm_idx(tf, i)) * N + i];
                a_conj[k * 2 * N + c * N + i] = F[(c * N + perm_idx(tc, i)) * N + i];
            }
    }
}

// Expanded server-side key a-parts: ksk5a[j] = NTT(expand_seed(seed,"ksk",
// 5*100+j+1)), kneg1a[j] = NTT(expand_seed(seed,"ksk",(2N-1)*100+j+1))
// (src/gpu_protocol.cu:405-456). out [D_EFF][2][N] each (wire digit order).
void orc_ksk_a(const uint8_t* seed, uint64_t* ksk5a, uint64_t* kneg1a) {
    for (int j = 0; j < D_EFF; j++) {
        expand_seed(seed, "ksk", AUTO_GEN * 100 + j + 1, ksk5a + j * 2 * N); poly_ntt(ksk5a + j * 2 * N);
        expand_seed(seed, "ksk", (2 * N - 1) * 100 + j + 1, kneg1a + j * 2 * N); poly_ntt(kneg1a + j * 2 * N);
    }
}

// decomp of an NTT-form state: INTT, decomp, NTT digits. dig [D_EFF][2][N].
static void digits_ntt(const uint64_t* state_ntt, uint64_t* dig) {
    Vec c(state_ntt, state_ntt + 2 * N);
    poly_intt(c.data());
    decomp(c.data(), dig);
    for (int j = 0; j < D_EFF; j++) poly_ntt(dig + j * 2 * N);
}

// CollapseA for one group (src/gpu_preprocess.cu:452-881). Step s folds
// k = half-1-s with key step kskstep = n_steps-1-s, whose expanded key is
// tau_{5^kskstep} (conjugate: tau_{2N-5^kskstep}) of ksk5a.
// out_a [2][N]; Dp, Dm [2][NSTEPS][D_EFF][N]; Df [2][D_EFF][N].
void orc_collapse(const uint64_t* a_fwd, const uint64_t* a_conj, const uint64_t* ksk5a, const uint64_t* kneg1a,
                  uint64_t* out_a, uint64_t* Dp, uint64_t* Dm, uint64_t* Df) {
    Vec st(2 * N), dig(D_EFF * 2 * N), saved(2 * N);
    for (int dir = 0; dir < 2; dir++) {
        const uint64_t* ap = dir ? a_conj : a_fwd;
        uint64_t* Dout = dir ? Dm : Dp;
        memcpy(st.data(), ap + (HALF - 1) * 2 * N, 2 * N * 8);
        for (uint64_t s = 0; s < NSTEPS; s++) {
            uint64_t k = HALF - 1 - s, ks = NSTEPS - 1 - s;
            uint64_t t = pow5(ks); if (dir) t = (2 * N - t) % (2 * N);
            digits_ntt(st.data(), dig.data());
            for (int l = 0; l < 2; l++) {
                uint64_t q = QS[l];
                for (uint32_t i = 0; i < N; i++) {
                    uint32_t p = perm_idx(t, i);
                    uint64_t acc = 0;
                    for (int j = 0; j < D_EFF; j++) {
                        uint64_t dv = dig[j * 2 * N + l * N + i];
                        Dout[((l * NSTEPS + s) * D_EFF + j) * N + i] = dv;
                        acc = madd(acc, mmul_q(dv, ksk5a[j * 2 * N + l * N + p], l), q);
                    }
                    st[l * N + i] = msub(ap[(k - 1) * 2 * N + l * N + i], acc, q);
                }
            }
        }
        if (!dir) saved = st;
    }
    digits_ntt(st.data(), dig.data());
    for (int l = 0; l < 2; l++) {
        uint64_t q = QS[l];
        for (uint64_t i = 0; i < N; i++) {
            uint64_t acc = 0;
            for (int j = 0; j < D_EFF; j++) {
                uint64_t dv = dig[j * 2 * N + l * N + i];
                Df[(l * D_EFF + j) * N + i] = dv;
                acc = madd(acc, mmul_q(dv, kneg1a[j * 2 * N + l * N + i], l), q);
            }
            out_a[l * N + i] = msub(saved[l * N + i], acc, q);
        }
    }
}

// Whole preprocess of the groups in [g0, g1): encode (all rows, in place,
// caller passes the u16 DB), CRS, a32, ring embed, collapse. Output arrays
// are indexed by (g - g0). Returns 0.
int orc_preprocess_groups(uint16_t* db, uint64_t R, uint64_t C, uint64_t D, const uint8_t* seed,
                          uint64_t g0, uint64_t g1, int do_encode,
                          uint64_t* out_a, uint64_t* Dp, uint64_t* Dm, uint64_t* Df) {
    Geo g = geo_of(R, C, D);
    if (do_encode && g.Da > 1) {
        std::vector<uint64_t> rows(R);
        for (uint64_t r = 0; r < R; r++) rows[r] = r;
        orc_encode_rows(db, R, C, D, rows.data(), R);
    }
    Vec crs(g.nr * 2 * N), a32(N * 2 * N), af(HALF * 2 * N), ac(HALF * 2 * N);
    Vec k5(D_EFF * 2 * N), kn(D_EFF * 2 * N);
    orc_crs(seed, g.nr, crs.data());
    orc_ksk_a(seed, k5.data(), kn.data());
    const size_t dsz = 2 * NSTEPS * D_EFF * N, fsz = 2 * D_EFF * N;
    for (uint64_t gi = g0; gi < g1; gi++) {
        uint64_t o = gi - g0;
        orc_a32(db, R, C, gi, crs.data(), a32.data());
        orc_ring_embed(a32.data(), af.data(), ac.data());
        orc_collapse(af.data(), ac.data(), k5.data(), kn.data(),
                     out_a + o * 2 * N, Dp + o * dsz, Dm + o * dsz, Df + o * fsz);
    }
    return 0;
}

// ---- online ----------------------------------------------------------------
// Mat-vec (src/gpu_online.cu:23-41): out_c[j] = sum_i db[i][j] * b_c[i] mod q_c,
// columns [c0, c1). b given as u64 residues (reduced mod q first,
// src/gpu_protocol.cu:1044-1055). out [2][c1-c0].
void orc_matvec(const uint16_t* db, uint64_t R, uint64_t C, const uint64_t* b0, const uint64_t* b1,
                uint64_t c0, uint64_t c1, uint64_t* out) {
    uint64_t W = c1 - c0;
    std::vector<unsigned __int128> s0(W, 0), s1(W, 0);
    for (uint64_t i = 0; i < R; i++) {
        uint64_t x0 = b0[i] % Q0, x1 = b1[i] % Q1;
        const uint16_t* row = db + i * C + c0;
        for (uint64_t j = 0; j < W; j++) { s0[j] += (uint64_t)row[j] * x0; s1[j] += (uint64_t)row[j] * x1; }
    }
    for (uint64_t j = 0; j < W; j++) { out[j] = (uint64_t)(s0[j] % Q0); out[W + j] = (uint64_t)(s1[j] % Q1); }
}

// Pack one group (src/gpu_protocol.cu:1225-1283 with the kernels of
// src/gpu_online.cu:547-632): b = NTT(interm[gN..]) per limb; then
// b -= sum_s sum_j D+[s][j] * K5[j][perm_f_s]  with perm exponent
// 5^(half-2-s), D- likewise with 2N-5^(half-2-s); b -= sum_j Df[j]*Kn[j].
// K5 = NTT(ksk5_b), Kn = NTT(kskneg1_b) (client keys, coeff form, wire
// [digit][limb][coeff], reduced mod q). interm [2][N] for this group.
// out packed_b [2][N].
void orc_pack_group(const uint64_t* interm, const uint64_t* Dp, const uint64_t* Dm, const uint64_t* Df,
                    const uint64_t* ksk5_b, const uint64_t* kneg1_b, uint64_t* out_b) {
    Vec K5(D_EFF * 2 * N), Kn(D_EFF * 2 * N);
    for (uint64_t x = 0; x < D_EFF * 2 * N; x++) {
        uint64_t q = QS[(x / N) % 2];
        K5[x] = ksk5_b[x] % q; Kn[x] = kneg1_b[x] % q;
    }
    for (int j = 0; j < D_EFF; j++) { poly_ntt(&K5[j * 2 * N]); poly_ntt(&Kn[j * 2 * N]); }
    for (int l = 0; l < 2; l++) {
        uint64_t q = QS[l];
        Vec b(interm + l * N, interm + (l + 1) * N);
        ntt_fwd(b.data(), l);
        for (uint32_t i = 0; i < N; i++) {
            unsigned __int128 acc = 0;
            for (uint64_t s = 0; s < NSTEPS; s++) {
                uint64_t tf = pow5(HALF - 2 - s);  // 5^(N/2) = 1 mod 2N, so pow5 indexes mod N
                uint32_t pf = perm_idx(tf, i), pc = perm_idx((2 * N - tf) % (2 * N), i);
                for (int j = 0; j < D_EFF; j++) {
                    acc += (unsigned __int128)Dp[((l * NSTEPS + s) * D_EFF + j) * N + i] * K5[j * 2 * N + l * N + pf];
                    acc += (unsigned __int128)Dm[((l * NSTEPS + s) * D_EFF + j) * N + i] * K5[j * 2 * N + l * N + pc];
                }
            }
            for (int j = 0; j < D_EFF; j++)
                acc += (unsigned __int128)Df[(l * D_EFF + j) * N + i] * Kn[j * 2 * N + l * N + i];
            out_b[l * N + i] = msub(b[i], (uint64_t)(acc % q), q);
        }
    }
}

// Full answer for one query given the encoded DB and the precomp of all
// groups (src/gpu_protocol.cu:1121-1339). query: flat wire layout.
// out: [ct][a|b][limb][N] (src/capi.cu:115-124).
// Optional taps (may be null): interm [2][C], packed_b [G][2][N].
void orc_answer(const uint16_t* db, uint64_t R, uint64_t C, uint64_t D,
                const uint64_t* pa, const uint64_t* Dp, const uint64_t* Dm, const uint64_t* Df,
                const uint64_t* query, uint64_t* out, uint64_t* tap_interm, uint64_t* tap_packed_b) {
    Geo g = geo_of(R, C, D);
    const size_t blk = (size_t)D_EFF * 2 * N;
    const size_t dsz = 2 * NSTEPS * D_EFF * N, fsz = 2 * D_EFF * N;
    Vec interm(2 * C);
    orc_matvec(db, R, C, query, query + R, 0, C, interm.data());
    if (tap_interm) memcpy(tap_interm, interm.data(), 2 * C * 8);
    const uint64_t* k5 = query + 2 * R;
    const uint64_t* kn = k5 + blk;
    Vec rg(query + 2 * R + 2 * blk, query + 2 * R + 6 * blk);
    for (size_t x = 0; x < rg.size(); x++) rg[x] %= QS[(x / N) % 2];
    Vec packed(g.G * 4 * N), gi(2 * N);  // per group [a|b][2][N]
    for (uint64_t gg = 0; gg < g.G; gg++) {
        for (int l = 0; l < 2; l++) memcpy(&gi[l * N], &interm[l * C + gg * N], N * 8);
        memcpy(&packed[gg * 4 * N], pa + gg * 2 * N, 2 * N * 8);
        orc_pack_group(gi.data(), Dp + gg * dsz, Dm + gg * dsz, Df + gg * fsz, k5, kn, &packed[gg * 4 * N + 2 * N]);
        if (tap_packed_b) memcpy(tap_packed_b + gg * 2 * N, &packed[gg * 4 * N + 2 * N], 2 * N * 8);
    }
    for (uint64_t ct = 0; ct < g.T; ct++) {
        uint64_t start = ct * g.Da, end = std::min(start + g.Da, g.G), L = end - start;
        orc_horner(&packed[start * 4 * N], L, rg.data(), out + ct * 4 * N);
    }
}

// ---- wire codecs (src/capi_client.cpp:174-304, :310-385) -------------------
// 53-bit Garner CRT value of a residue pair.
uint64_t orc_garner(uint64_t v0, uint64_t v1) {
    static const uint64_t inv = minv(Q1 % Q0, Q0);
    uint64_t diff = (v0 + Q0 - v1 % Q0) % Q0;
    return v1 + Q1 * (uint64_t)((unsigned __int128)diff * inv % Q0);
}
// Mod-switch compression of one ct (NTT form [a|b][2][N]) to 12288 bytes:
// N x 28-bit a' then N x 20-bit b', MSB-first, flushed (src/crypto.cpp:285-313).
void orc_compress_ct(const uint64_t* ct, uint8_t* outb) {
    Vec a(ct, ct + 2 * N), b(ct + 2 * N, ct + 4 * N);
    poly_intt(a.data()); poly_intt(b.data());
    auto down = [](uint64_t v, uint64_t qp) {
        unsigned __int128 num = (unsigned __int128)v * qp + (uint64_t)(QFULL / 2);
        return (uint64_t)(num / QFULL) % qp;
    };
    uint64_t acc = 0; int nb = 0; uint8_t* p = outb;
    auto put = [&](uint64_t v, int bits) {
        acc = (acc << bits) | v; nb += bits;
        while (nb >= 8) { nb -= 8; *p++ = (uint8_t)(acc >> nb); }
        acc &= (1ull << nb) - 1;
    };
    for (uint64_t i = 0; i < N; i++) put(down(crt(a[i], a[N + i]), 1ull << 28), 28);
    for (uint64_t i = 0; i < N; i++) put(down(crt(b[i], b[N + i]), 1ull << 20), 20);
    if (nb) { *p++ = (uint8_t)(acc << (8 - nb)); }
}

} // extern "C"
