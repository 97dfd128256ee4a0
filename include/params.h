// Scheme constants and database geometry.
//
// Drop-in for the reference's src/params.h (constants src/params.h:15-42,
// PublicParams src/params.h:49-58, setup() src/params.h:63-66): same names,
// same values, same meaning, so that queries built against either library
// are interchangeable.
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>

namespace inspire {

constexpr size_t N = 2048;                  // ring degree
constexpr uint64_t Q0 = 94519297;           // RNS prime 0 (27 bits, = 1 mod 2N)
constexpr uint64_t Q1 = 95293441;           // RNS prime 1 (27 bits, = 1 mod 2N)
constexpr int NUM_LIMBS = 2;
constexpr uint64_t P = 65535;               // plaintext modulus 2^16 - 1
constexpr int P_BITS = 15;                  // raw bits carried per slot
constexpr double SIGMA = 3.19;              // error standard deviation
constexpr int D_GSW = 3;                    // gadget digits
constexpr int D_EFF = D_GSW - 1;            // digits kept (digit 0 dropped)
constexpr int BASE_LOG = 18;                // gadget base 2^18
constexpr uint64_t BASE = 1ULL << BASE_LOG;
constexpr int AUTO_GEN = 5;                 // Galois generator
constexpr size_t SEED_BYTES = 64;           // CRS seed length

constexpr uint64_t Q_PRIME_0 = 1ULL << 20;  // compressed body modulus
constexpr uint64_t Q_PRIME_1 = 1ULL << 28;  // compressed mask modulus
constexpr int BETA = 53;                    // bits of Q0*Q1

constexpr uint64_t Q_FULL = (uint64_t)((__uint128_t)Q0 * Q1);
constexpr uint64_t DELTA = Q_FULL / P;
constexpr uint64_t DELTA_MOD_Q0 = DELTA % Q0;
constexpr uint64_t DELTA_MOD_Q1 = DELTA % Q1;

struct PublicParams {
    size_t N_entries;   // number of database entries
    size_t w;           // entry size in bytes
    size_t db_rows;     // query length (multiple of N)
    size_t db_cols;     // multiple of N
    size_t n_packed;    // db_cols / N  (groups)
    size_t D;           // interpolation degree (256)
    size_t num_cts;     // response ciphertexts = max(1, n_packed / D)
    std::array<uint8_t, SEED_BYTES> seed;  // CRS seed
};

// Default geometry db_rows = 32768. Throws std::invalid_argument for shapes
// whose entries straddle N-column boundaries or whose n_packed is not a
// power of two / multiple of D.
PublicParams setup(size_t N_entries, size_t w);
PublicParams setup(size_t N_entries, size_t w, size_t db_rows_override);

// Same geometry with a caller-chosen CRS seed instead of a fresh one
// (no counterpart in the reference, which draws from std::random_device).
PublicParams setup_seeded(size_t N_entries, size_t w, size_t db_rows,
                          const uint8_t seed[SEED_BYTES]);

}  // namespace inspire
