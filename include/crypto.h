// Client-side and reference-side RLWE primitives.
//
// Drop-in for the reference's src/crypto.h (sk_gen :9, rlwe_dec :14,
// CompressedCt/compress_ct/dec_compressed :24-31, LweQuery/lwe_enc_b :36-42,
// Ksk/ksk_gen_b :47-52, RgswCt/rgsw_enc_b :57-67, ext_prod :72,
// horner_eval :77).
#pragma once
#include "ring.h"

namespace inspire {

RnsPoly sk_gen(std::mt19937_64& rng);
std::vector<uint64_t> rlwe_dec(const RnsPoly& s_ntt, const RlweCt& ct);

struct CompressedCt {
    std::vector<uint32_t> a;  // N values < 2^28
    std::vector<uint32_t> b;  // N values < 2^20
};
CompressedCt compress_ct(const RlweCt& ct);
std::vector<uint64_t> dec_compressed(const RnsPoly& s, const CompressedCt& cct);

struct LweQuery {
    std::vector<uint64_t> b_limb0;
    std::vector<uint64_t> b_limb1;
};
LweQuery lwe_enc_b(const RnsPoly& s, size_t i_star, size_t ell,
                   const uint8_t seed[SEED_BYTES], std::mt19937_64& rng);

struct Ksk {
    std::vector<RnsPoly> b_parts;  // D_EFF polynomials, coefficient form
};
Ksk ksk_gen_b(const RnsPoly& s, uint64_t k, const uint8_t seed[SEED_BYTES], std::mt19937_64& rng);

struct RgswCt {
    struct Row {
        std::vector<RnsPoly> a_parts;  // D_EFF, NTT form
        std::vector<RnsPoly> b_parts;  // D_EFF, NTT form
    };
    Row top, bottom;
};
RgswCt rgsw_enc_b(const RnsPoly& s, const RnsPoly& mu, const uint8_t seed[SEED_BYTES],
                  std::mt19937_64& rng);

RlweCt ext_prod(const RgswCt& rgsw, const RlweCt& ct);
RlweCt horner_eval(const std::vector<RlweCt>& cts, const RgswCt& rgsw, size_t D);

}  // namespace inspire
