// Device-side building blocks shared by every gfx950 kernel of the server:
// ring constants, exact modular arithmetic on 27-bit residues, and a
// negacyclic NTT/INTT over one 2048-coefficient polynomial held in LDS,
// executed by a team of 256 lanes (4 wavefronts).
//
// Conventions (identical to the client's CPU NTT, so device and host
// representations are interchangeable): slot i of NTT(f) holds
// f(psi^(2*BR(i)+1)); forward is Cooley-Tukey DIT with twiddle fwd[m+g],
// inverse is Gentleman-Sande with inv[m+g] and a final N^-1 scale
// (reference src/ntt.cpp:107-149, src/gpu_common.cuh:160-238).
//
// All functions keep values canonical in [0, q) between stages.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ipx {

constexpr uint32_t N = 2048;
constexpr uint32_t LOGN = 11;
constexpr uint32_t HALF = N / 2;
constexpr uint32_t NSTEPS = HALF - 1;  // collapse steps per direction
constexpr uint32_t Q0 = 94519297u;
constexpr uint32_t Q1 = 95293441u;
constexpr uint32_t PMOD = 65535u;
constexpr int D_EFF = 2;
constexpr int BASE_LOG = 18;
constexpr uint64_t QFULL = (uint64_t)Q0 * Q1;
// (Q1 mod Q0)^-1 mod Q0 and (Q0 mod Q1)^-1 mod Q1, filled by the host
// from the same extended-Euclid computation the client uses.
struct CrtConst {
    uint32_t q1inv_q0;
    uint32_t q0inv_q1;
};

// The collapse tensors D+/D- in the reference's per-group order, u32 words
// [limb][step][digit][coeff] for each direction (RD_SZ words each); the
// preprocess chain produces them in this order and the inspection entry
// point returns them in it.
constexpr size_t RD_SZ = 2 * (size_t)NSTEPS * D_EFF * N;
// Words of one group in the reference order: [a: 2N][D+][D-][Df: 2*2*N].
constexpr size_t REF_GROUP_WORDS = 2 * (size_t)N + 2 * RD_SZ + 2 * (size_t)D_EFF * N;

// Resident per-group precomp block, u32 words: [a: 2N][Dd: DD_SZ][Df: 2*2*N].
// Dd is D+/D- in dot-product form. For slot i the 2*NSTEPS collapse steps
// of both directions gather the key at 2*NSTEPS distinct slots p, so
//   sum_s sum_j D+[s][j][i] K[j][p+(s,i)] + D-[s][j][i] K[j][p-(s,i)]
//     = sum_{j,p} Dd[j][p][i] K[j][p]
// with Dd[j][p][i] = D+/-[s][j][i] where p = p+/-(s, i), and 0 at the two
// slots no step reaches. Per limb this is a K = 2N (j, p) by N (i) matrix
// of residues (< 2^27), stored as three byte planes and a nibble plane in
// 16-row blocks, k = j*N + p:
//   plane a < 3: byte (((limb*3 + a)*(2N/16) + k/16)*N + i)*16 + k%16
//     holds byte a of the residue XOR 0x80 (read as int8: byte - 128);
//   top bits (residue >> 24, < 8): byte DD_TOP_BYTES +
//     ((limb*(2N/16) + k/16)*N + i)*8 + k%8, the low nibble for k%16 < 8
//     and the high nibble for the row 8 further on,
// so a lane's 8 nibble bytes widen (two masks, one shift) into the same
// 16-byte MFMA operand as a full plane's 16 bytes: 3.5 bytes per entry.
constexpr size_t DD_K = 2 * (size_t)N;             // K rows per limb
constexpr size_t DD_TOP_BYTES = 2 * 3 * DD_K * N;  // the full planes of both limbs
constexpr size_t DD_SZ = 7 * DD_K * N / 4;         // words (3.5 bytes per entry, both limbs)
constexpr size_t PC_A = 0;
constexpr size_t PC_DD = 2 * (size_t)N;
constexpr size_t PC_DF = PC_DD + DD_SZ;
constexpr size_t PC_GROUP = PC_DF + 2 * (size_t)D_EFF * N;  // 7174*N words = 58.77 MB

// Twiddle tables in device memory: tw[limb*4*N + {0,1,2,3}*N + i] =
// fwd, fwd Shoup companion, inv, inv Shoup companion.
struct NttConst {
    const uint32_t* tw;
    uint32_t inv_n[2];
    uint32_t inv_n_p[2];
};

__host__ __device__ __forceinline__ uint32_t qof(int limb) { return limb ? Q1 : Q0; }

__host__ __device__ __forceinline__ uint32_t add_mod(uint32_t a, uint32_t b, uint32_t q) {
    uint32_t s = a + b;
    return s >= q ? s - q : s;
}
__host__ __device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t q) {
    return a >= b ? a - b : a + q - b;
}
__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) {
    return (uint32_t)(((uint64_t)a * b) >> 32);
}
// x*w mod q for a fixed w with companion wp = floor(w*2^32/q); any x < 2^32.
__host__ __device__ __forceinline__ uint32_t mul_shoup(uint32_t x, uint32_t w, uint32_t wp, uint32_t q) {
    uint32_t r = x * w - mulhi32(x, wp) * q;  // in [0, 2q)
    return r >= q ? r - q : r;
}
// Exact x mod q for x < 2^54.2 (a sum of two products of canonical
// residues), q a 27-bit prime, mu = floor(2^55/q). The quotient estimate
// ((x>>27)*mu)>>28 undershoots by at most 2, so two conditional
// subtractions leave a canonical residue.
__host__ __device__ __forceinline__ uint32_t red54(uint64_t x, uint32_t q, uint32_t mu) {
    uint64_t t = ((x >> 27) * (uint64_t)mu) >> 28;
    uint32_t r = (uint32_t)x - (uint32_t)t * q;
    r = r >= q ? r - q : r;
    return r >= q ? r - q : r;
}
__host__ __device__ __forceinline__ uint32_t mu55(int limb) { return limb ? 378082653u : 381179274u; }
__host__ __device__ __forceinline__ uint32_t mul_mod(uint32_t a, uint32_t b, int limb) {
    return red54((uint64_t)a * b, qof(limb), mu55(limb));
}
// Exact x mod q for any u64 x: fold the high word twice through
// 2^32 mod q, landing below 2^54.03, then red54.
__host__ __device__ __forceinline__ uint32_t red64(uint64_t x, int limb) {
    const uint32_t t32 = limb ? (uint32_t)(4294967296ull % Q1) : (uint32_t)(4294967296ull % Q0);
    uint64_t y = (uint64_t)(uint32_t)(x >> 32) * t32 + (uint32_t)x;  // < 2^59.01
    y = (y >> 32) * t32 + (uint32_t)y;                                  // < 2^54.03
    return red54(y, qof(limb), mu55(limb));
}

__device__ __forceinline__ uint32_t brev11(uint32_t x) { return __builtin_bitreverse32(x) >> (32 - LOGN); }

// Slot permutation realising tau_t on the NTT domain:
// perm_t(i) = BR(((t*(2BR(i)+1) mod 2N) - 1)/2)   (t odd).
// For t' = 2N - t, perm_t'(i) = N-1-perm_t(i).
__device__ __forceinline__ uint32_t perm_slot(uint32_t t, uint32_t v /* 2BR(i)+1 */) {
    uint32_t e = (t * v) & (2 * N - 1);
    return brev11((e - 1) >> 1);
}

// ---------------------------------------------------------------- block sum
// Sum of one u64 per thread over a block of TH threads (a multiple of 64):
// a butterfly within each wave, then thread 0 adds the TH/64 wave sums
// (red: TH/64 words of LDS). The result is valid in thread 0 only.
template <int TH>
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t a, uint64_t* red) {
    static_assert(TH % 64 == 0, "whole waves");
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    uint64_t s = 0;
    if (threadIdx.x == 0)
#pragma unroll
        for (int w = 0; w < TH / 64; w++) s += red[w];
    return s;
}

// ---------------------------------------------------------------- hand-over
// Kernels that run side by side hand groups over through a counter or flag
// in device memory, from a block on one XCD to a block on another (the
// hardware keeps the XCDs' L2 caches coherent only through these fences).
// Producer: every wave waits for its own stores (s_waitcnt 0), the block
// meets at a barrier, then one lane writes the L2 back (agent-scope
// release) and waits for that write-back before it touches the flag.
// Consumer: the lane that saw the flag invalidates (agent-scope acquire:
// the CU's L1 and the XCD's L2) and waits for that, then the block meets
// at a barrier before any wave loads the group. One invalidation serves
// the whole block, whose waves share that CU and XCD; a second one per
// wave only flushes the caches again (1% of a 32-query answer). The waits
// are written out: after an explicit s_waitcnt the compiler drops the
// fence's own, which would let the flag overtake the write-back and other
// waves' loads overtake the invalidation.
__device__ __forceinline__ void fence_release_agent_done() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_s_waitcnt(0);
}
__device__ __forceinline__ void fence_acquire_agent_done() {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __builtin_amdgcn_s_waitcnt(0);
}

// ---------------------------------------------------------------- LDS NTT
// LDS slot of coefficient e: one pad word per 32 spreads the strided
// accesses of the register-blocked phases over the banks.
constexpr uint32_t POLY_LDS = N + N / 32;
__device__ __forceinline__ uint32_t ls(uint32_t e) { return e + (e >> 5); }

// Lazy butterflies (Harvey): values stay below 4q (forward) or 2q
// (inverse) between stages, which the 27-bit primes keep far below 2^32.
// x in [0, 2m) -> x mod m: x - m wraps above x exactly when x < m.
__device__ __forceinline__ uint32_t fold(uint32_t x, uint32_t m) {
    const uint32_t y = x - m;
    return y < x ? y : x;  // unsigned: a v_min_u32
}
// x*w mod q up to one extra q (result in [0, 2q)), any x < 2^32.
__device__ __forceinline__ uint32_t mul_shoup_lazy(uint32_t x, uint32_t w, uint32_t wp, uint32_t q) {
    return x * w - mulhi32(x, wp) * q;
}

// One register-blocked phase of the forward transform: the lane owns the 8
// coefficients base + S*k (k < 8), which are closed under the butterflies
// of strides 4S, 2S, S (NST = 3) or 2S, S (NST = 2). Inputs below 4q,
// outputs below 4q, canonical after the LAST phase.
template <int S, int NST, bool LAST>
__device__ __forceinline__ void fwd_phase(uint32_t* poly, uint32_t t, const uint32_t* __restrict__ w,
                                          const uint32_t* __restrict__ wp, uint32_t q) {
    const uint32_t base = (t / S) * 8 * S + (t % S);
    const uint32_t q2 = 2 * q;
    uint32_t x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = poly[ls(base + S * k)];
#pragma unroll
    for (int b = NST - 1; b >= 0; b--) {
        const uint32_t st = (uint32_t)S << b;
        const uint32_t m = N / (2 * st);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (k & (1 << b)) continue;
            const uint32_t j = base + S * k;
            const uint32_t g = j / (2 * st);
            const uint32_t u = fold(x[k], q2);                                   // [0, 2q)
            const uint32_t v = mul_shoup_lazy(x[k + (1 << b)], w[m + g], wp[m + g], q);  // [0, 2q)
            x[k] = u + v;                                                        // [0, 4q)
            x[k + (1 << b)] = u + q2 - v;                                        // (0, 4q)
        }
    }
    if (LAST) {
#pragma unroll
        for (int k = 0; k < 8; k++) x[k] = fold(fold(x[k], q2), q);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) poly[ls(base + S * k)] = x[k];
}

// Inverse phase: inputs below 2q, outputs below 2q; the scaling phase
// (by N^-1, a full Shoup product) leaves them canonical.
template <int S, int NST>
__device__ __forceinline__ void inv_phase(uint32_t* poly, uint32_t t, const uint32_t* __restrict__ w,
                                          const uint32_t* __restrict__ wp, uint32_t q, bool scale,
                                          uint32_t inv_n, uint32_t inv_n_p) {
    const uint32_t base = (t / S) * 8 * S + (t % S);
    const uint32_t q2 = 2 * q;
    uint32_t x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = poly[ls(base + S * k)];
#pragma unroll
    for (int b = 0; b < NST; b++) {
        const uint32_t st = (uint32_t)S << b;
        const uint32_t m = N / (2 * st);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (k & (1 << b)) continue;
            const uint32_t j = base + S * k;
            const uint32_t g = j / (2 * st);
            const uint32_t u = x[k], v = x[k + (1 << b)];
            x[k] = fold(u + v, q2);                                                   // [0, 2q)
            x[k + (1 << b)] = mul_shoup_lazy(u + q2 - v, w[m + g], wp[m + g], q);    // [0, 2q)
        }
    }
    if (scale) {
#pragma unroll
        for (int k = 0; k < 8; k++) x[k] = mul_shoup(x[k], inv_n, inv_n_p, q);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) poly[ls(base + S * k)] = x[k];
}

// Twiddle tables of one limb: forward, its Shoup companions, inverse, its
// Shoup companions (N words each). Either the device tables of NttConst or
// a copy staged in LDS by the kernel (stage_twiddles), which keeps the
// per-butterfly address arithmetic out of the register file.
struct Tw {
    const uint32_t* w;
    const uint32_t* wp;
    const uint32_t* iw;
    const uint32_t* iwp;
};
__device__ __forceinline__ Tw tw_global(const NttConst& nc, int limb) {
    const uint32_t* b = nc.tw + (size_t)limb * 4 * N;
    return Tw{b, b + N, b + 2 * N, b + 3 * N};
}
// Copy both limbs' tables (8*N words) into LDS; caller synchronises.
__device__ __forceinline__ void stage_twiddles(uint32_t* lds_tw, const NttConst& nc, uint32_t tid, uint32_t nthreads) {
    for (uint32_t x = tid; x < 8 * N; x += nthreads) lds_tw[x] = nc.tw[x];
}
__device__ __forceinline__ Tw tw_lds(const uint32_t* lds_tw, int limb) {
    const uint32_t* b = lds_tw + (size_t)limb * 4 * N;
    return Tw{b, b + N, b + 2 * N, b + 3 * N};
}

// Forward NTT of one LDS polynomial by a 256-lane team (t = lane in team).
// Every lane of the whole workgroup must call this (it contains workgroup
// barriers); lanes of a team with nothing to transform pass active=false
// and only take part in the barriers. The caller synchronises before
// (input visible); the routine ends with a barrier (output visible).
__device__ __forceinline__ void ntt_fwd_lds(uint32_t* poly, uint32_t t, const Tw& tw, int limb, bool active = true) {
    const uint32_t q = qof(limb);
    if (active) fwd_phase<256, 3, false>(poly, t, tw.w, tw.wp, q);
    __syncthreads();
    if (active) fwd_phase<32, 3, false>(poly, t, tw.w, tw.wp, q);
    __syncthreads();
    if (active) fwd_phase<4, 3, false>(poly, t, tw.w, tw.wp, q);
    __syncthreads();
    if (active) fwd_phase<1, 2, true>(poly, t, tw.w, tw.wp, q);
    __syncthreads();
}

__device__ __forceinline__ void ntt_inv_lds(uint32_t* poly, uint32_t t, const Tw& tw, int limb, uint32_t inv_n,
                                            uint32_t inv_n_p, bool active = true) {
    const uint32_t q = qof(limb);
    if (active) inv_phase<1, 2>(poly, t, tw.iw, tw.iwp, q, false, 0, 0);
    __syncthreads();
    if (active) inv_phase<4, 3>(poly, t, tw.iw, tw.iwp, q, false, 0, 0);
    __syncthreads();
    if (active) inv_phase<32, 3>(poly, t, tw.iw, tw.iwp, q, false, 0, 0);
    __syncthreads();
    if (active) inv_phase<256, 3>(poly, t, tw.iw, tw.iwp, q, true, inv_n, inv_n_p);
    __syncthreads();
}

// Gadget digits 1..2 of a coefficient given by its residues: CRT lift to
// [0, Q), centre (values above floor(Q/2) are negative), three signed
// base-2^18 digits by arithmetic shift, digit 0 dropped
// (reference src/ring.cpp:205-246, src/gpu_preprocess.cu:543-571).
// Outputs the two kept digits as residues mod Q0 and mod Q1.
__device__ __forceinline__ void gadget2(uint32_t r0, uint32_t r1, const CrtConst& cc, uint32_t d_q0[2],
                                        uint32_t d_q1[2]) {
    const uint32_t t0 = mul_mod(r0, cc.q1inv_q0, 0);
    const uint32_t t1 = mul_mod(r1, cc.q0inv_q1, 1);
    uint64_t v = (uint64_t)t0 * Q1 + (uint64_t)t1 * Q0;
    if (v >= QFULL) v -= QFULL;
    int64_t r = v > (QFULL >> 1) ? (int64_t)v - (int64_t)QFULL : (int64_t)v;
    const int64_t d0 = (int64_t)((uint64_t)r << (64 - BASE_LOG)) >> (64 - BASE_LOG);
    r = (r - d0) >> BASE_LOG;
    const int64_t d1 = (int64_t)((uint64_t)r << (64 - BASE_LOG)) >> (64 - BASE_LOG);
    r = (r - d1) >> BASE_LOG;
    const int64_t d2 = (int64_t)((uint64_t)r << (64 - BASE_LOG)) >> (64 - BASE_LOG);
    d_q0[0] = d1 >= 0 ? (uint32_t)d1 : (uint32_t)((int64_t)Q0 + d1);
    d_q1[0] = d1 >= 0 ? (uint32_t)d1 : (uint32_t)((int64_t)Q1 + d1);
    d_q0[1] = d2 >= 0 ? (uint32_t)d2 : (uint32_t)((int64_t)Q0 + d2);
    d_q1[1] = d2 >= 0 ? (uint32_t)d2 : (uint32_t)((int64_t)Q1 + d2);
}

}  // namespace ipx
