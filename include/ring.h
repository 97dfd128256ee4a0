// RNS polynomials over Z_q[X]/(X^N+1), q = Q0*Q1, and the CPU ring helpers
// the client and the server's table builders use.
//
// Drop-in for the reference's src/ring.h (RnsPoly src/ring.h:16-48, RlweCt
// :51-53, automorphism :59-62, decomp :69, expand_seed :74-75, samplers
// :80-81, twiddle export :104-118), plus src/ntt.h (ntt_forward /
// ntt_inverse / ntt_tables).
#pragma once
#include <random>
#include <vector>

#include "params.h"

namespace inspire {

struct NttTables {
    uint64_t q;
    std::array<uint64_t, N> fwd_twiddles;  // psi^BR(i)
    std::array<uint64_t, N> inv_twiddles;  // inverses, inv[0] = 1
    std::array<uint64_t, N> fwd_primes;    // floor(w * 2^32 / q)
    std::array<uint64_t, N> inv_primes;
    uint64_t inv_n;
    uint64_t inv_n_prime;
    uint64_t psi;
};
const NttTables& ntt_tables(int limb);
void ntt_forward(uint64_t* data, int limb);  // in place, N values mod q_limb
void ntt_inverse(uint64_t* data, int limb);  // in place, includes N^-1

class RnsPoly {
public:
    std::array<std::vector<uint64_t>, NUM_LIMBS> limbs;
    bool is_ntt = false;

    RnsPoly();
    static RnsPoly zero();
    void to_ntt();
    void to_coeff();
    void add_inplace(const RnsPoly& other);
    void sub_inplace(const RnsPoly& other);
    void mul_inplace(const RnsPoly& other);          // both NTT form
    void scalar_mul_inplace(uint64_t s0, uint64_t s1);
    void negate_inplace();
    void mul_sub_inplace(const RnsPoly& a, const RnsPoly& b);
    void inner_product_sub(const RnsPoly* a, const RnsPoly* b, int d);
    void set_coeff(size_t idx, uint64_t v0, uint64_t v1);
};

struct RlweCt {
    RnsPoly a, b;
};

RnsPoly automorphism(const RnsPoly& f, uint64_t k);          // coefficient form
RnsPoly automorphism_ntt(const RnsPoly& f_ntt, uint64_t k);  // NTT form
std::vector<RnsPoly> decomp(const RnsPoly& f);               // D_EFF digits, coeff form
RnsPoly expand_seed(const uint8_t seed[SEED_BYTES], const char* domain, uint64_t index);
RnsPoly sample_ternary(std::mt19937_64& rng);
RnsPoly sample_error(std::mt19937_64& rng);
uint64_t galois_power(uint64_t g, size_t k);
const std::vector<uint64_t>& gen_powers();
uint64_t mod_inv(uint64_t a, uint64_t m);

void get_ntt_twiddles(int limb, std::vector<uint32_t>& fwd, std::vector<uint32_t>& inv, uint32_t& inv_n);
void get_ntt_twiddles_and_primes(int limb, std::vector<uint32_t>& fwd, std::vector<uint32_t>& fwd_primes,
                                 std::vector<uint32_t>& inv, std::vector<uint32_t>& inv_primes,
                                 uint32_t& inv_n, uint32_t& inv_n_prime);

// SHAKE-256 (FIPS 202), used by expand_seed.
void shake256(const uint8_t* in, size_t inlen, uint8_t* out, size_t outlen);

}  // namespace inspire
