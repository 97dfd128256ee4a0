// The direct int8-MFMA database-pass body (up to 8 queries), shared by the
// pass kernel and the streamed pass-and-collapse kernel, and the
// recombination of a lane's accumulators into residues.
#pragma once
#include "kernels.h"

namespace ipx {

namespace {

typedef int32_t i32x4 __attribute__((ext_vector_type(4)));

// Recombine one lane's accumulators of pair pr and store the canonical residue.
__device__ __forceinline__ void store_pair(const i32x4& al, const i32x4& ah, uint32_t pr, uint32_t nq, bool corr,
                                           const uint32_t* __restrict__ S, uint32_t* __restrict__ out,
                                           size_t out_qstride, uint32_t slice, uint32_t OC, uint32_t col) {
    const uint32_t q = pr >> 1;
    const int c = (int)(pr & 1);
    if (q >= nq) return;
    int64_t s = corr ? (int64_t)32896 * S[pr] : 0;
#pragma unroll
    for (int i = 0; i < 4; i++) s += ((int64_t)ah[i] * 256 + (int64_t)al[i]) * ((int64_t)1 << (8 * i));
    // |s| < 2^58.1 (|d_3| <= 8 bounds the top digit's term); add a multiple
    // of q_c above that to make it non-negative
    const uint64_t qc = c ? Q1 : Q0;
    const uint64_t off = qc * ((1ull << 59) / qc);
    out[(size_t)q * out_qstride + (size_t)slice * 2 * OC + (size_t)c * OC + col] =
        red64((uint64_t)(s + (int64_t)off), c);
}

// Up to 2*MT queries (Mp = 16*MT digit rows) over column tile xt (4 waves x
// NT 16-column tiles from column 4*NT*16*xt of db) and row slice `slice`
// (kbps 16-row blocks): the 4 waves of a block own disjoint column ranges
// and load their DB operands straight from HBM into registers, one 64-row
// step ahead; every wave reads all MT digit tiles (L2/L1-resident). Two
// waves per SIMD. Writes out[q][slice][limb][OC].
template <int MT, int NT>
__device__ __forceinline__ void dbpass_direct_body(const uint8_t* __restrict__ db, const int8_t* __restrict__ dig,
                                                   const uint32_t* __restrict__ S, uint32_t* __restrict__ out,
                                                   size_t out_qstride, uint32_t C, uint32_t OC, uint32_t nq,
                                                   uint32_t kbps, uint32_t xt, uint32_t slice) {
    constexpr uint32_t Mp = 16 * MT;
    const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const uint32_t col0 = (xt * 4 + w) * NT * 16;
    const uint32_t kb0 = slice * kbps;
    const uint32_t nks = kbps / 4;
    // lane's low-plane unit of block kb0 + l/16, column col0 + l%16; the high plane is C units later
    const i32x4* bsrc = reinterpret_cast<const i32x4*>(db) + (size_t)(kb0 + (l >> 4)) * 2 * C + col0 + (l & 15);
    const size_t bstep = (size_t)8 * C;  // 4 blocks x 2 planes, in 16-byte units
    const i32x4* asrc = reinterpret_cast<const i32x4*>(dig) + (size_t)(kb0 + (l >> 4)) * Mp + (l & 15);
    const size_t astep = (size_t)4 * Mp;
    const i32x4 flip = {(int32_t)0x80808080, (int32_t)0x80808080, (int32_t)0x80808080, (int32_t)0x80808080};
    i32x4 acc[MT][NT][2];
#pragma unroll
    for (int a = 0; a < MT; a++)
#pragma unroll
        for (int n = 0; n < NT; n++) acc[a][n][0] = acc[a][n][1] = i32x4{0, 0, 0, 0};
    i32x4 v[2][NT][2];
    i32x4 A[2][MT];
    auto load = [&](int bi, uint32_t ks) {
#pragma unroll
        for (int n = 0; n < NT; n++) {
            v[bi][n][0] = __builtin_nontemporal_load(bsrc + ks * bstep + n * 16);
            v[bi][n][1] = __builtin_nontemporal_load(bsrc + ks * bstep + n * 16 + C);
        }
#pragma unroll
        for (int a = 0; a < MT; a++) A[bi][a] = asrc[ks * astep + a * 16];
    };
    auto compute = [&](int bi) {
#pragma unroll
        for (int n = 0; n < NT; n++) {
            const i32x4 Bl = v[bi][n][0] ^ flip, Bh = v[bi][n][1] ^ flip;
#pragma unroll
            for (int a = 0; a < MT; a++) {
                acc[a][n][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[bi][a], Bl, acc[a][n][0], 0, 0, 0);
                acc[a][n][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[bi][a], Bh, acc[a][n][1], 0, 0, 0);
            }
        }
    };
    load(0, 0);
    for (uint32_t ks = 0; ks < nks; ks += 2) {  // nks is even
        load(1, ks + 1);
        compute(0);
        if (ks + 2 < nks) load(0, ks + 2);
        compute(1);
    }
#pragma unroll
    for (int a = 0; a < MT; a++)
#pragma unroll
        for (int n = 0; n < NT; n++)
            store_pair(acc[a][n][0], acc[a][n][1], (uint32_t)a * 4 + (l >> 4), nq, slice == 0, S, out, out_qstride,
                       slice, OC, col0 + n * 16 + (l & 15));
}

}  // namespace

}  // namespace ipx
