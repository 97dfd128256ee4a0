// The precomp-tensor ("collapse") pass in dot-product form, for gfx950.
//
// Reference semantics (src/gpu_online.cu:547-632, driven per group by
// src/gpu_protocol.cu:1225-1283): for query q, group g, limb c, slot i
//   part[q][g][c][i] = sum_s sum_j D+[s][j][i] K5[j][p+(s,i)] + D-[s][j][i] K5[j][p-(s,i)]
// where K5 = NTT(ksk5_b) of the query and p+(s,i) = perm_{5^(N/2-2-s)}(i),
// p-(s,i) = N-1-p+(s,i). Writing slot p as the odd residue u = 2BR(p)+1 mod
// 2N, p+/- correspond to u = +/-5^(N/2-2-s) * (2BR(i)+1); since 5 has order
// N/2 mod 2N and -1 is not a power of 5, the 2*(N/2-1) pairs (s, dir) hit
// distinct slots. So the pass equals sum_{j,p} Dd[j][p][i] K5[j][p] with Dd
// the D+/- entries scattered to their slots (two slots per i stay zero):
// per limb a (2N x 2N-row) by N matrix-vector product per query, which a
// batch turns into one int8 MFMA contraction that reads Dd once.
#include "kernels.h"
#include "dbpass_direct.h"

namespace ipx {

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// A lane's 16 top-bit entries of one 16-row block (8 nibble bytes) as the
// 16-byte MFMA operand of rows 0..15: low nibbles are rows 0..7, high 8..15.
__device__ __forceinline__ i32x4 widen_top(u32x2 t) {
    return i32x4{(int)(t.x & 0x0f0f0f0fu), (int)(t.y & 0x0f0f0f0fu), (int)((t.x >> 4) & 0x0f0f0f0fu),
                 (int)((t.y >> 4) & 0x0f0f0f0fu)};
}

// This lane's Dd operands of group g, limb c at the column tile starting at
// col0: full planes (16-byte units, plane stride NKB*N) and the top nibbles
// (8-byte units), both advancing by 4*N units per 64-row step.
__device__ __forceinline__ const i32x4* dd_planes(const uint32_t* precomp, uint32_t g, int c, uint32_t l,
                                                 uint32_t col0) {
    return reinterpret_cast<const i32x4*>(precomp + (size_t)g * PC_GROUP + PC_DD) +
           (size_t)c * 3 * (DD_K / 16) * N + (size_t)(l >> 4) * N + col0 + (l & 15);
}
__device__ __forceinline__ const u32x2* dd_top(const uint32_t* precomp, uint32_t g, int c, uint32_t l,
                                              uint32_t col0) {
    return reinterpret_cast<const u32x2*>(reinterpret_cast<const uint8_t*>(precomp + (size_t)g * PC_GROUP + PC_DD) +
                                          DD_TOP_BYTES) +
           (size_t)c * (DD_K / 16) * N + (size_t)(l >> 4) * N + col0 + (l & 15);
}

// e with 5^e = x mod 2N, for x = 1 mod 4 (bit by bit: 5^(2^k) = 1 + 2^(k+2)
// mod 2^(k+3), so the k-th bit is set iff x*5^-e is not 1 mod 2^(k+3)).
__device__ __forceinline__ uint32_t dlog5(uint32_t x) {
    // 5^-(2^k) mod 2^12, k = 0..9
    const uint32_t inv[10] = {3277u, 3113u, 3729u, 3617u, 65u, 129u, 257u, 513u, 1025u, 2049u};
    uint32_t e = 0;
#pragma unroll
    for (int k = 0; k < 10; k++) {
        if ((x & ((8u << k) - 1)) != 1u) {
            x = (x * inv[k]) & (2 * N - 1);
            e |= 1u << k;
        }
    }
    return e;
}

// v^-1 mod 2N for odd v (Newton: each step doubles the correct low bits).
__device__ __forceinline__ uint32_t inv_odd(uint32_t v) {
    uint32_t x = v;  // correct mod 8
#pragma unroll
    for (int k = 0; k < 3; k++) x = x * (2u - v * x);
    return x & (2 * N - 1);
}

// ------------------------------------------------------------ preprocess relayout
// Block (group, limb, digit j, 16-slot tile of i): stage D+/-[.][j][i-tile]
// (2 x 1023 steps x 16 slots) in LDS, then emit the Dd rows (j, p) of the
// tile as the three byte planes and the nibble plane. Sources in the
// reference order [limb][step][digit][coeff] at Dp/Dm + group*gstride.
__global__ __launch_bounds__(1024) void k_dd_relayout(const uint32_t* __restrict__ Dp, const uint32_t* __restrict__ Dm,
                                                      size_t gstride, uint32_t* __restrict__ pc0) {
    __shared__ uint32_t T[2][NSTEPS][16];
    const uint32_t it = blockIdx.x % (N / 16), rest = blockIdx.x / (N / 16);
    const uint32_t j = rest % D_EFF, c = (rest / D_EFF) % 2, gl = rest / (2 * D_EFF);
    const uint32_t i0 = it * 16, tid = threadIdx.x;
    for (uint32_t x = tid; x < 2 * NSTEPS * 4; x += 1024) {
        const uint32_t row = x >> 2, qd = x & 3, dir = row / NSTEPS, s = row % NSTEPS;
        const uint32_t* src = (dir ? Dm : Dp) + gl * gstride + (((size_t)c * NSTEPS + s) * D_EFF + j) * N + i0 + qd * 4;
        const u32x4 v = *reinterpret_cast<const u32x4*>(src);
        T[dir][s][qd * 4 + 0] = v[0];
        T[dir][s][qd * 4 + 1] = v[1];
        T[dir][s][qd * 4 + 2] = v[2];
        T[dir][s][qd * 4 + 3] = v[3];
    }
    __syncthreads();
    uint8_t* dd = reinterpret_cast<uint8_t*>(pc0 + (size_t)gl * PC_GROUP + PC_DD);
    for (uint32_t x = tid; x < 16 * (N / 16); x += 1024) {
        const uint32_t il = x & 15, pb = x >> 4, i = i0 + il;
        const uint32_t vinv = inv_odd(2 * brev11(i) + 1);
        uint32_t pl[3][4], top[2] = {0, 0};
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const uint32_t p = pb * 16 + e;
            const uint32_t w = ((2 * brev11(p) + 1) * vinv) & (2 * N - 1);
            const uint32_t neg = (w & 3) == 3;
            const uint32_t e5 = dlog5(neg ? 2 * N - w : w);
            const uint32_t v = e5 == HALF - 1 ? 0u : T[neg][HALF - 2 - e5][il];
            const uint32_t b[3] = {(v & 255u) ^ 128u, ((v >> 8) & 255u) ^ 128u, ((v >> 16) & 255u) ^ 128u};
#pragma unroll
            for (int a = 0; a < 3; a++) {
                if ((e & 3) == 0) pl[a][e >> 2] = 0;
                pl[a][e >> 2] |= b[a] << (8 * (e & 3));
            }
            // row e: nibble (e >= 8) of byte e % 8
            top[(e & 7) >> 2] |= (v >> 24) << (8 * (e & 3) + 4 * (e >> 3));
        }
        const size_t kb = (size_t)j * (N / 16) + pb;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            u32x4* dst = reinterpret_cast<u32x4*>(dd + ((((size_t)c * 3 + a) * (DD_K / 16) + kb) * N + i) * 16);
            *dst = u32x4{pl[a][0], pl[a][1], pl[a][2], pl[a][3]};
        }
        *reinterpret_cast<u32x2*>(dd + DD_TOP_BYTES + (((size_t)c * (DD_K / 16) + kb) * N + i) * 8) =
            u32x2{top[0], top[1]};
    }
}

// ------------------------------------------------------------ online
// Key digits: kdig[c][kb][Mp][16 bytes], row m = q*4 + b holding digit b
// (balanced base 256, as for the database pass) of K5_q[j][c][p] at
// k = j*N + p = 16*kb + byte. Padding rows (q >= nq) are zero.
__global__ __launch_bounds__(256) void k_kdigits(const uint32_t* __restrict__ keys, size_t kstride, uint32_t nq,
                                                 uint32_t Mp, int8_t* __restrict__ kdig) {
    const uint32_t nkb = DD_K / 16, nqp = Mp / 4;
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;
    if (x >= 2 * nkb * nqp) return;
    const uint32_t kb = x % nkb, q = (x / nkb) % nqp, c = x / (nkb * nqp);
    u32x4 o[4];
#pragma unroll
    for (int b = 0; b < 4; b++) o[b] = u32x4{0, 0, 0, 0};
    if (q < nq) {
        const uint32_t k0 = kb * 16, j = k0 / N, p0 = k0 % N;
        const uint32_t* K = keys + (size_t)q * kstride + ((size_t)j * 2 + c) * N + p0;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            int32_t v = (int32_t)K[e];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                int32_t d = v;
                if (b < 3) {
                    d = ((v & 255) ^ 128) - 128;
                    v = (v - d) >> 8;
                }
                o[b][e >> 2] |= ((uint32_t)d & 255u) << (8 * (e & 3));
            }
        }
    }
    u32x4* dst = reinterpret_cast<u32x4*>(kdig + (((size_t)c * nkb + kb) * Mp + (size_t)q * 4) * 16);
#pragma unroll
    for (int b = 0; b < 4; b++) dst[b] = o[b];
}

// KS[q*2 + c] = sum_{j,p} K5_q[j][c][p] mod q_c (the plane-offset correction).
__global__ __launch_bounds__(1024) void k_ksum(const uint32_t* __restrict__ keys, size_t kstride,
                                               uint32_t* __restrict__ KS) {
    __shared__ uint64_t red[16];
    const uint32_t q = blockIdx.x >> 1, c = blockIdx.x & 1;
    // DD_K = 2N entries: digit d's limb-c polynomial at keys + (2d + c) N
    const uint32_t* k = keys + (size_t)q * kstride + (size_t)c * N + threadIdx.x;
    uint64_t a = 0;
#pragma unroll
    for (int e = 0; e < 4; e++) a += k[(size_t)(e >> 1) * 2 * N + (size_t)(e & 1) * 1024];
    const uint64_t sum = block_sum_u64<1024>(a, red);
    if (threadIdx.x == 0) KS[blockIdx.x] = red64(sum, (int)c);
}

// Recombine one lane's 16 accumulators (4 Dd planes x 4 key digits) of
// query q for one slot and store the canonical residue. With Dd entry =
// sum_a 256^a x_a + 128*65793 (x_a the signed planes, x_3 the top byte) and
// K = sum_b 256^b d_b:
//   sum_k Dd_k K_k = sum_{a,b} 256^(a+b) A_ab + 128*65793*sum_k K_k,
// A_ab = sum_k x_a d_b over 4096 rows: |A_ab| <= 2^26, exact in i32; the
// 16 products fold into 7 levels weighted by 256^k mod q, |sum| < 2^58.
// Row slices (KSlice): a slice's sum over its rows, with the 128*65793 sum(K)
// term in slice 0 only, is a residue added atomically into part2 (zeroed
// before the launch): at most 8 residues, below 2^30; every reader of
// part2 reduces it with its own terms.
struct KSlice {
    uint32_t k0 = 0;       // first 64-row step of the slice
    uint32_t nks = DD_K / 64;  // its steps (even)
    bool first = true;     // adds the 128*65793 sum(K) term
    bool add = false;      // atomic add (split) or store (whole)
};
// The slice of a split launch's grid, z = group * ksplit + slice.
__device__ __forceinline__ KSlice slice_of(uint32_t ksplit) {
    KSlice k;
    const uint32_t sl = blockIdx.z % ksplit;
    k.nks = (uint32_t)(DD_K / 64) / ksplit;
    k.k0 = sl * k.nks;
    k.first = sl == 0;
    k.add = ksplit > 1;
    return k;
}
__device__ __forceinline__ void collapse_store(const i32x4 (&acc)[4], uint32_t q, int c, uint32_t g, uint32_t i,
                                               const uint32_t* __restrict__ KS, uint32_t* __restrict__ part2,
                                               size_t p2_stride, const KSlice& ks = KSlice{}) {
    const uint64_t qc = c ? Q1 : Q0;
    int64_t lv[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int pl = 0; pl < 4; pl++)
#pragma unroll
        for (int b = 0; b < 4; b++) lv[pl + b] += acc[pl][b];
    int64_t s = ks.first ? (int64_t)8421504 * KS[q * 2 + c] : 0;  // 128*65793 * sum(K)
    uint64_t wk = 1;                                // 256^k mod q_c
#pragma unroll
    for (int k = 0; k < 7; k++) {
        s += (int64_t)wk * lv[k];
        wk = (wk * 256) % qc;
    }
    const uint64_t off = qc * ((1ull << 59) / qc);
    uint32_t* dst = part2 + (size_t)q * p2_stride + (size_t)g * 2 * N + (size_t)c * N + i;
    const uint32_t r = red64((uint64_t)(s + (int64_t)off), c);
    if (ks.add)
        atomicAdd(dst, r);
    else
        *dst = r;
}

// Block (4*NT*16-slot tile of i, limb, group): 4 waves of NT 16-slot tiles,
// all M rows (MT tiles of 4 queries x 4 digits). Per 64-row K step a lane
// loads its 16-byte operand of each full Dd plane and its 8 top-nibble
// bytes straight from HBM (a step ahead). The key digits of the step (Mp x 64 bytes) are read by each wave
// from L2 (SHARED = false) or staged once per block in an LDS double
// buffer (SHARED = true, for large MT).
// (the body, for slot tile `it`, limb c and group g; the kernel's grid is
// (tile, limb, group))
template <int MT, int NT, bool SHARED>
__device__ __forceinline__ void collapse_mfma_body(const uint32_t* __restrict__ precomp,
                                                   const int8_t* __restrict__ kdig,
                                                   const uint32_t* __restrict__ KS,
                                                   uint32_t* __restrict__ part2, size_t p2_stride, uint32_t nq,
                                                   uint32_t it, int c, uint32_t g, const KSlice& ksl = KSlice{}) {
    constexpr uint32_t Mp = 16 * MT;
    constexpr uint32_t NKB = DD_K / 16;
    constexpr uint32_t AU = SHARED ? Mp * 4 / 256 : 1;  // 16-byte units of a step's digit tile per thread
    __shared__ i32x4 atile[SHARED ? 2 : 1][SHARED ? 4 * Mp : 1];
    const uint32_t tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const uint32_t col0 = (it * 4 + w) * NT * 16;
    const size_t pstride = (size_t)NKB * N;  // plane, in 16-byte units
    const size_t bstep = (size_t)4 * N;      // 4 blocks of 16 rows
    // this slice's steps [k0, k0 + nks): the step index below is relative
    const i32x4* dd = dd_planes(precomp, g, c, l, col0) + (size_t)ksl.k0 * bstep;
    const u32x2* dt = dd_top(precomp, g, c, l, col0) + (size_t)ksl.k0 * bstep;
    const i32x4* kd = reinterpret_cast<const i32x4*>(kdig) + (size_t)c * NKB * Mp +
                      (size_t)ksl.k0 * 4 * Mp;  // this limb's digits
    i32x4 acc[MT][NT][4];
#pragma unroll
    for (int a = 0; a < MT; a++)
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int pl = 0; pl < 4; pl++) acc[a][n][pl] = i32x4{0, 0, 0, 0};
    i32x4 B[2][NT][3];  // full planes; the top nibbles stay packed until their MFMA
    u32x2 Bt[2][NT];
    i32x4 A[SHARED ? 1 : 2][SHARED ? 1 : MT];
    i32x4 apre[AU];
    auto load = [&](int bi, uint32_t ks) {
#pragma unroll
        for (int n = 0; n < NT; n++) {
#pragma unroll
            for (int pl = 0; pl < 3; pl++)
                B[bi][n][pl] = __builtin_nontemporal_load(dd + pl * pstride + ks * bstep + n * 16);
            Bt[bi][n] = __builtin_nontemporal_load(dt + ks * bstep + n * 16);
        }
        if (!SHARED) {
#pragma unroll
            for (int a = 0; a < MT; a++) A[SHARED ? 0 : bi][SHARED ? 0 : a] = kd[((size_t)ks * 4 + (l >> 4)) * Mp + a * 16 + (l & 15)];
        }
    };
    auto load_a = [&](uint32_t ks) {
#pragma unroll
        for (int u = 0; u < (int)AU; u++) apre[u] = kd[(size_t)ks * 4 * Mp + tid + 256 * u];
    };
    auto store_a = [&](int abuf) {
#pragma unroll
        for (int u = 0; u < (int)AU; u++) atile[SHARED ? abuf : 0][SHARED ? tid + 256 * u : 0] = apre[u];
    };
    auto compute = [&](int bi) {
#pragma unroll
        for (int a = 0; a < MT; a++) {
            const i32x4 Aop = SHARED ? atile[SHARED ? bi : 0][SHARED ? (l >> 4) * Mp + a * 16 + (l & 15) : 0]
                                     : A[SHARED ? 0 : bi][SHARED ? 0 : a];
#pragma unroll
            for (int n = 0; n < NT; n++) {
#pragma unroll
                for (int pl = 0; pl < 3; pl++)
                    acc[a][n][pl] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Aop, B[bi][n][pl], acc[a][n][pl], 0, 0, 0);
                acc[a][n][3] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Aop, widen_top(Bt[bi][n]), acc[a][n][3], 0, 0, 0);
            }
        }
    };
    const uint32_t nks = ksl.nks;  // even
    if (SHARED) {
        load_a(0);
        store_a(0);
        load_a(1);
        __syncthreads();
    }
    load(0, 0);
    for (uint32_t ks = 0; ks < nks; ks += 2) {
        load(1, ks + 1);
        compute(0);
        if (SHARED) {
            store_a(1);  // digits of step ks+1
            __syncthreads();
            if (ks + 2 < nks) load_a(ks + 2);
        }
        if (ks + 2 < nks) load(0, ks + 2);
        compute(1);
        if (SHARED) {
            if (ks + 2 < nks) {
                store_a(0);  // digits of step ks+2
                if (ks + 3 < nks) load_a(ks + 3);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int a = 0; a < MT; a++) {
        const uint32_t q = a * 4 + (l >> 4);
        if (q >= nq) continue;
#pragma unroll
        for (int n = 0; n < NT; n++)
            collapse_store(acc[a][n], q, c, g, col0 + n * 16 + (l & 15), KS, part2, p2_stride, ksl);
    }
}

template <int MT, int NT, bool SHARED>
__global__ __launch_bounds__(256, 2) void k_collapse_mfma(const uint32_t* __restrict__ precomp,
                                                          const int8_t* __restrict__ kdig,
                                                          const uint32_t* __restrict__ KS,
                                                          uint32_t* __restrict__ part2, size_t p2_stride,
                                                          uint32_t nq, uint32_t ksplit) {
    collapse_mfma_body<MT, NT, SHARED>(precomp, kdig, KS, part2, p2_stride, nq, blockIdx.x, (int)blockIdx.y,
                                       blockIdx.z / ksplit, slice_of(ksplit));
}

// Streamed pass and collapse in one launch (up to 8 queries): per group,
// from the top, its collapse work items (CNT-wide slot tiles of both
// limbs) and then its database-pass work items (the row slices of its
// PNT-wide column tiles), all 256-lane workgroups of the same footprint, so
// the two HBM streams of a group run together and groups complete in the
// order the Horner chains consume them. Every work item, once its stores
// are complete (one agent-scope release per workgroup), adds 1 to its
// group's counter; a group is in place at 2 N/(64 CNT) + (N/(64 PNT)) x
// slices.
template <int PMT, int PNT, int CMT, int CNT>
__global__ __launch_bounds__(256, 2) void k_pass_collapse_stream(
    const uint8_t* __restrict__ db, const int8_t* __restrict__ dig, const uint32_t* __restrict__ S,
    uint32_t* __restrict__ out, size_t out_qstride, uint32_t C, uint32_t kbps, const uint32_t* __restrict__ precomp,
    const int8_t* __restrict__ kdig, const uint32_t* __restrict__ KS, uint32_t* __restrict__ part2,
    size_t p2_stride, uint32_t nq, uint32_t G, uint32_t slice_shift, uint32_t* __restrict__ group_done) {
    constexpr uint32_t CIT = N / (64 * CNT);  // collapse slot tiles per limb
    constexpr uint32_t CB = 2 * CIT;          // collapse work items per group
    constexpr uint32_t TPG = N / (64 * PNT);  // pass column tiles per group
    const uint32_t per = CB + (TPG << slice_shift);
    const uint32_t b = blockIdx.x, gi = b / per, r = b - gi * per, g = G - 1 - gi;
    if (r < CB) {
        collapse_mfma_body<CMT, CNT, false>(precomp, kdig, KS, part2, p2_stride, nq, r % CIT, (int)(r / CIT), g);
    } else {
        const uint32_t pr = r - CB;
        dbpass_direct_body<PMT, PNT>(db, dig, S, out, out_qstride, C, C, nq, kbps, g * TPG + (pr >> slice_shift),
                                     pr & ((1u << slice_shift) - 1));
    }
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are complete
    __syncthreads();
    if (threadIdx.x == 0) {
        fence_release_agent_done();
        __hip_atomic_fetch_add(group_done + g, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// 17..32 queries (8 key-digit tiles) in the load pattern of the batched
// database pass: every global load of a step (the Dd operands of the 3
// full planes and the top nibbles into a register ring of DEPTH slots, the thread's share of the
// step's key-digit tile into a ring of DEPTH-1 slots, stored to LDS one
// step ahead) is issued DEPTH-1 steps before its step, so each wait covers
// one step's loads (vector-memory loads complete in issue order); the
// block's WAVES waves share the digit tile. Past the last step the loads
// repeat it (in bounds, unused).
template <int MT, int NT, int DEPTH, int WAVES>
__global__ __launch_bounds__(64 * WAVES, 2) void k_collapse_ring(const uint32_t* __restrict__ precomp,
                                                                 const int8_t* __restrict__ kdig,
                                                                 const uint32_t* __restrict__ KS,
                                                                 uint32_t* __restrict__ part2, size_t p2_stride,
                                                                 uint32_t nq, uint32_t ksplit) {
    static_assert(DEPTH >= 3, "the digit ring needs two slots");
    constexpr int L = DEPTH - 1;
    constexpr int P = DEPTH * L;
    constexpr uint32_t Mp = 16 * MT;
    constexpr uint32_t NKB = DD_K / 16;
    constexpr uint32_t TH = 64 * WAVES;
    constexpr uint32_t AU = Mp * 4 / TH;  // 16-byte units of a step's digit tile per thread
    const KSlice ksl = slice_of(ksplit);
    const uint32_t nks = ksl.nks;
    __shared__ i32x4 atile[2][4 * Mp];
    const uint32_t tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const uint32_t g = blockIdx.z / ksplit;
    const int c = (int)blockIdx.y;
    const uint32_t col0 = (blockIdx.x * WAVES + w) * NT * 16;
    const size_t pstride = (size_t)NKB * N;  // plane, in 16-byte units
    const size_t bstep = (size_t)4 * N;      // 4 blocks of 16 rows
    // this slice's steps [k0, k0 + nks): the step index below is relative
    const i32x4* dd = dd_planes(precomp, g, c, l, col0) + (size_t)ksl.k0 * bstep;
    const u32x2* dt = dd_top(precomp, g, c, l, col0) + (size_t)ksl.k0 * bstep;
    const i32x4* kd = reinterpret_cast<const i32x4*>(kdig) + (size_t)c * NKB * Mp + (size_t)ksl.k0 * 4 * Mp + tid;
    i32x4 acc[MT][NT][4];
#pragma unroll
    for (int a = 0; a < MT; a++)
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int pl = 0; pl < 4; pl++) acc[a][n][pl] = i32x4{0, 0, 0, 0};
    i32x4 v[DEPTH][NT][3];  // full planes; the top nibbles stay packed until their MFMA
    u32x2 vt[DEPTH][NT];
    i32x4 ar[L][AU];
    auto issue = [&](int bs, int as, uint32_t step) {
        const uint32_t k = step < nks ? step : nks - 1;
#pragma unroll
        for (int u = 0; u < (int)AU; u++) ar[as][u] = kd[(size_t)k * 4 * Mp + TH * u];
#pragma unroll
        for (int n = 0; n < NT; n++) {
#pragma unroll
            for (int pl = 0; pl < 3; pl++)
                v[bs][n][pl] = __builtin_nontemporal_load(dd + pl * pstride + k * bstep + n * 16);
            vt[bs][n] = __builtin_nontemporal_load(dt + k * bstep + n * 16);
        }
    };
    auto store_a = [&](int buf, int as) {
#pragma unroll
        for (int u = 0; u < (int)AU; u++) atile[buf][tid + TH * u] = ar[as][u];
    };
    auto compute = [&](int bs, int buf) {
#pragma unroll
        for (int a = 0; a < MT; a++) {
            const i32x4 Aop = atile[buf][(l >> 4) * Mp + a * 16 + (l & 15)];
#pragma unroll
            for (int n = 0; n < NT; n++) {
#pragma unroll
                for (int pl = 0; pl < 3; pl++)
                    acc[a][n][pl] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Aop, v[bs][n][pl], acc[a][n][pl], 0, 0, 0);
                acc[a][n][3] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Aop, widen_top(vt[bs][n]), acc[a][n][3], 0, 0, 0);
            }
        }
    };
    // Step s: issue step s+L's loads, compute from slot s%DEPTH and LDS
    // buffer s&1, store step s+1's digits in buffer (s+1)&1 (its last
    // reader, step s-1, finished before the previous barrier).
#pragma unroll
    for (int j = 0; j < L; j++) issue(j, j, (uint32_t)j);
    store_a(0, 0);
    __syncthreads();
    for (uint32_t ks = 0; ks < nks; ks += P) {
#pragma unroll
        for (int j = 0; j < P; j++) {
            const uint32_t step = ks + (uint32_t)j;
            if (step >= nks) break;
            issue((j + L) % DEPTH, (j + L) % L, step + L);
            compute(j % DEPTH, (int)(step & 1));
            store_a((int)((step + 1) & 1), (j + 1) % L);
            __syncthreads();
        }
    }
#pragma unroll
    for (int a = 0; a < MT; a++) {
        const uint32_t q = a * 4 + (l >> 4);
        if (q >= nq) continue;
#pragma unroll
        for (int n = 0; n < NT; n++)
            collapse_store(acc[a][n], q, c, g, col0 + n * 16 + (l & 15), KS, part2, p2_stride, ksl);
    }
}

}  // namespace

hipError_t launch_dd_relayout(const uint32_t* Dp, const uint32_t* Dm, size_t gstride, size_t ng, uint32_t* precomp_g0,
                              hipStream_t s) {
    if (ng == 0) return hipSuccess;
    hipLaunchKernelGGL(k_dd_relayout, dim3((unsigned)(ng * 2 * D_EFF * (N / 16))), dim3(1024), 0, s, Dp, Dm, gstride,
                       precomp_g0);
    return hipGetLastError();
}

static uint32_t collapse_mt(size_t nq) { return nq <= 4 ? 1 : nq <= 8 ? 2 : nq <= 16 ? 4 : 8; }  // 4 queries per tile

hipError_t launch_collapse_keys(const uint32_t* keys, size_t kstride, size_t nq, int8_t* kdig, uint32_t* KS,
                                hipStream_t s) {
    if (nq == 0 || nq > kCollapseMaxQueries) return hipErrorInvalidValue;
    const uint32_t Mp = 16 * collapse_mt(nq);
    const size_t nthr = 2 * (DD_K / 16) * (Mp / 4);
    hipLaunchKernelGGL(k_kdigits, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, s, keys, kstride, (uint32_t)nq,
                       Mp, kdig);
    hipLaunchKernelGGL(k_ksum, dim3((unsigned)(2 * nq)), dim3(1024), 0, s, keys, kstride, KS);
    return hipGetLastError();
}

uint32_t pass_collapse_target(const DbPlan& p, size_t nq, size_t G) {
    const uint32_t cnt = collapse_mt(nq) == 1 && G < 16 ? 1 : 2;
    return (uint32_t)(2 * (N / (64 * cnt)) + (N / dbplan_cols(p)) * p.slices);
}

hipError_t launch_pass_collapse_stream(const uint8_t* db, const int8_t* dig, const uint32_t* S, size_t nq,
                                       uint32_t* out, size_t out_qstride, size_t R, size_t C, const DbPlan& p,
                                       const uint32_t* precomp, const int8_t* kdig, const uint32_t* KS,
                                       uint32_t* part2, size_t p2_stride, uint32_t* group_done, hipStream_t s) {
    const size_t G = C / N;
    if (p.kind != DbPlan::MFMA || p.param > 4 || nq == 0 || nq > 8 || nq > dbplan_max_queries(p) || C % N ||
        !group_done || p.slices == 0 || (p.slices & (p.slices - 1)) || dbplan_cols(p) != 256)
        return hipErrorInvalidValue;
    const uint32_t kbps = (uint32_t)(R / 16 / p.slices);
    if (kbps % 8 || kbps == 0) return hipErrorInvalidValue;
    uint32_t sh = 0;
    while ((1u << sh) < p.slices) sh++;
    const uint32_t cmt = collapse_mt(nq), cnt = cmt == 1 && G < 16 ? 1 : 2;
    const size_t blocks = G * pass_collapse_target(p, nq, G);
    if (blocks > 0x7fffffffu) return hipErrorInvalidValue;
    const dim3 grid((unsigned)blocks);
    const uint32_t Cc = (uint32_t)C, Q = (uint32_t)nq, Gg = (uint32_t)G;
#define PCS(PMT, CMT, CNT)                                                                                    \
    hipLaunchKernelGGL((k_pass_collapse_stream<PMT, 4, CMT, CNT>), grid, dim3(256), 0, s, db, dig, S, out, \
                       out_qstride, Cc, kbps, precomp, kdig, KS, part2, p2_stride, Q, Gg, sh, group_done)
    switch (p.param * 10 + cmt) {
        case 11: if (cnt == 1) PCS(1, 1, 1); else PCS(1, 1, 2); break;
        case 21: if (cnt == 1) PCS(2, 1, 1); else PCS(2, 1, 2); break;
        case 42: PCS(4, 2, 2); break;
        default: return hipErrorInvalidValue;
    }
#undef PCS
    return hipGetLastError();
}

namespace {
// rows x row_words words at p (rows stride words apart) set to zero
__global__ __launch_bounds__(256) void k_zero_rows(uint32_t* __restrict__ p, size_t stride, size_t row_words,
                                                   size_t total) {
    for (size_t x = (size_t)blockIdx.x * 256 + threadIdx.x; x < total; x += (size_t)gridDim.x * 256)
        p[(x / row_words) * stride + x % row_words] = 0;
}
}  // namespace

// Row slices of a collapse launch of `blocks` workgroups (2 limbs x tiles x
// groups): a grid too small to fill the chip (one group, 16-column tiles:
// 64 workgroups) splits its 4096-row contraction in up to 8 slices.
static uint32_t collapse_ksplit(size_t blocks) {
    uint32_t k = 1;
    while (blocks * k < 512 && k < 8) k *= 2;
    return k;
}

hipError_t launch_collapse(const uint32_t* precomp_g0, size_t ng, const int8_t* kdig, const uint32_t* KS, size_t nq,
                           uint32_t* part2_g0, size_t p2_stride, hipStream_t s, int variant) {
    if (nq == 0 || nq > kCollapseMaxQueries || variant < 0 || variant >= kCollapseVariants) return hipErrorInvalidValue;
    if (ng == 0) return hipSuccess;
    const uint32_t Q = (uint32_t)nq;
    // a split launch adds its slices' residues into part2, zeroed first
    auto split = [&](unsigned gx) {
        const uint32_t ks = collapse_ksplit((size_t)gx * 2 * ng);
        if (ks > 1) {
            const size_t row = ng * 2 * N, total = nq * row;
            hipLaunchKernelGGL(k_zero_rows, dim3((unsigned)((total + 255) / 256 < 1024 ? (total + 255) / 256 : 1024)),
                               dim3(256), 0, s, part2_g0, p2_stride, row, total);
        }
        return ks;
    };
#define COLLAPSE(MT, NT, SH)                                                                                          \
    do {                                                                                                              \
        const unsigned gx = N / (64 * NT);                                                                            \
        const uint32_t ks = split(gx);                                                                                \
        hipLaunchKernelGGL((k_collapse_mfma<MT, NT, SH>), dim3(gx, 2, (unsigned)(ng * ks)), dim3(256), 0, s,         \
                           precomp_g0, kdig, KS, part2_g0, p2_stride, Q, ks);                                         \
    } while (0)
#define RING(D, W)                                                                                                    \
    do {                                                                                                              \
        const unsigned gx = N / (16 * W);                                                                             \
        const uint32_t ks = split(gx);                                                                                \
        hipLaunchKernelGGL((k_collapse_ring<8, 1, D, W>), dim3(gx, 2, (unsigned)(ng * ks)), dim3(64 * W), 0, s,      \
                           precomp_g0, kdig, KS, part2_g0, p2_stride, Q, ks);                                         \
    } while (0)
    switch (collapse_mt(nq)) {
        case 1:  // few groups: 16-column tiles, twice the workgroups; many: 64
            if (ng < 16) COLLAPSE(1, 1, false);
            else if (ng < 32) COLLAPSE(1, 2, false);
            else COLLAPSE(1, 4, false);
            break;
        case 2: COLLAPSE(2, 2, false); break;
        case 4: COLLAPSE(4, 2, false); break;
        default:
            switch (variant) {
                case 1: RING(3, 4); break;
                case 2: RING(4, 4); break;
                case 3: RING(3, 8); break;
                case 4: RING(4, 8); break;
                case 5: RING(5, 8); break;
                default: COLLAPSE(8, 1, true); break;
            }
            break;
    }
#undef COLLAPSE
#undef RING
    return hipGetLastError();
}

}  // namespace ipx
