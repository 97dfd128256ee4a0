// Online (per-query) kernels for gfx950: query unmarshal, batched NTT, the
// row-slice reduction of the database pass, pack finalize, and the
// persistent Horner/external-product chain (the database and collapse
// passes live in kernels_db.hip and kernels_collapse.hip).
//
// Semantics follow the reference's online path (src/gpu_protocol.cu:
// 1121-1339): mat-vec interm_c = DB^T b_c mod q_c (src/gpu_online.cu:23-41);
// per group b = NTT(interm) minus the lazy collapse sums and the final
// D_final x ksk_neg1 term (src/gpu_online.cu:547-632); Horner
// acc = ext_prod(rgsw, acc) + packed[i] (src/gpu_online.cu:474-540). All
// arithmetic is exact, so results equal the reference's bit for bit.
#include "kernels.h"

namespace ipx {

namespace {

constexpr size_t BLK = (size_t)D_EFF * 2 * N;  // one key / RGSW part, u32 words
// s_sleep argument between two polls of a hand-over counter (units of 64
// clocks); 4, 8, 16 and 32 measured no different from 2 at the 16 GB tier
constexpr int kSpinSleep = 2;

// ------------------------------------------------------------ unmarshal
// Flat u64 wire query -> per-slot device layout, reducing every residue
// mod its limb's prime (the reference narrows ksk/RGSW words without
// reduction, src/gpu_protocol.cu:1066 and :50; canonical inputs agree).
__global__ __launch_bounds__(256) void k_unmarshal(const uint64_t* __restrict__ q, size_t total, size_t R,
                                                   double* __restrict__ bq, uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ rgsw) {
    const size_t qlen = 2 * R + 6 * BLK;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x) {
        const size_t qi = x / qlen, off = x - qi * qlen;
        const uint64_t v = q[x];
        if (off < 2 * R) {
            const int limb = off >= R;
            bq[qi * 2 * R + off] = (double)red64(v, limb);
        } else if (off < 2 * R + 2 * BLK) {
            const size_t o = off - 2 * R;
            keys[qi * 2 * BLK + o] = red64(v, (int)((o / N) & 1));
        } else {
            const size_t o = off - 2 * R - 2 * BLK;
            rgsw[qi * 4 * BLK + o] = red64(v, (int)((o / N) & 1));
        }
    }
}

// ------------------------------------------------------------ batched NTT
__global__ __launch_bounds__(256) void k_ntt_rns(uint32_t* __restrict__ polys, NttConst nc, int inverse) {
    __shared__ uint32_t buf[POLY_LDS];
    const int limb = blockIdx.x & 1;
    uint32_t* p = polys + (size_t)blockIdx.x * N;
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 8; k++) buf[ls(t + 256 * k)] = p[t + 256 * k];
    __syncthreads();
    if (inverse) ntt_inv_lds(buf, t, tw_global(nc, limb), limb, nc.inv_n[limb], nc.inv_n_p[limb]);
    else ntt_fwd_lds(buf, t, tw_global(nc, limb), limb);
#pragma unroll
    for (int k = 0; k < 8; k++) p[t + 256 * k] = buf[ls(t + 256 * k)];
}

// ------------------------------------------------------------ slice reduction
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// interm[q][limb][col] = sum over row slices of part mod q_limb, for columns
// col0 + [0, gridDim.x*256); blockIdx.y = limb, blockIdx.z = query. A
// workgroup covers 256 columns (4 per lane of a wave); its 4 waves take every
// 4th slice, so narrow passes with many slices still spread over many CUs
// and each lane has few loads in sequence. Exact 64-bit sums (< 2^32 per
// slice), combined in LDS, then one reduction.
__global__ __launch_bounds__(256) void k_reduce_slices(const uint32_t* __restrict__ part, size_t part_stride,
                                                       uint32_t S, uint32_t C, uint32_t col0, uint32_t ncols,
                                                       uint32_t* __restrict__ interm) {
    __shared__ uint64_t acc[3][64][4];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t j = blockIdx.x * 256 + lane * 4, col = col0 + j;
    const int limb = (int)blockIdx.y;
    const uint32_t q = blockIdx.z;
    const u32x4* pp = reinterpret_cast<const u32x4*>(part + q * part_stride + (size_t)limb * ncols + j);
    const size_t sstride = 2 * (size_t)ncols / 4;  // one slice, in u32x4
    uint64_t a[4] = {0, 0, 0, 0};
#pragma unroll 4
    for (uint32_t s = w; s < S; s += 4) {
        const u32x4 v = __builtin_nontemporal_load(pp + (size_t)s * sstride);
        a[0] += v[0]; a[1] += v[1]; a[2] += v[2]; a[3] += v[3];
    }
    if (w > 0)
#pragma unroll
        for (int e = 0; e < 4; e++) acc[w - 1][lane][e] = a[e];
    __syncthreads();
    if (w > 0) return;
    u32x4 o;
#pragma unroll
    for (int e = 0; e < 4; e++) o[e] = red64(a[e] + acc[0][lane][e] + acc[1][lane][e] + acc[2][lane][e], limb);
    *reinterpret_cast<u32x4*>(interm + (size_t)q * 2 * C + (size_t)limb * C + col) = o;
}

// ------------------------------------------------------------ pack finalize
// packed_b[q][g][c] = NTT(interm[q][c][gN..gN+N-1]) - part2[q][g][c]
//                     - sum_j Df[g][c][j] * Kn[j][c]  (reference
// src/gpu_protocol.cu:1225-1283, src/gpu_online.cu:547-592).
__global__ __launch_bounds__(256) void k_pack_finalize(const uint32_t* __restrict__ interm, uint32_t C,
                                                       const uint32_t* __restrict__ part2,
                                                       size_t p2_stride,
                                                       const uint32_t* __restrict__ precomp,
                                                       const uint32_t* __restrict__ keys, size_t kstride,
                                                       uint32_t* __restrict__ packed_b, size_t pstride,
                                                       uint32_t g0, NttConst nc) {
    __shared__ uint32_t buf[POLY_LDS];
    const uint32_t g = g0 + blockIdx.x, c = blockIdx.y, qi = blockIdx.z, t = threadIdx.x;
    const uint32_t q = qof((int)c);
    const uint32_t* pp = interm + (size_t)qi * 2 * C + (size_t)c * C + (size_t)g * N;
#pragma unroll
    for (int k = 0; k < 8; k++) buf[ls(t + 256 * k)] = pp[t + 256 * k];
    __syncthreads();
    ntt_fwd_lds(buf, t, tw_global(nc, (int)c), (int)c);
    const uint32_t* Df = precomp + (size_t)g * PC_GROUP + PC_DF + (size_t)c * D_EFF * N;
    const uint32_t* Kn = keys + qi * kstride + BLK + (size_t)c * N;  // key 1 = ksk_neg1, [digit][limb][N]
    const uint32_t* p2 = part2 + qi * p2_stride + (size_t)g * 2 * N + (size_t)c * N;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const uint32_t i = t + 256 * k;
        const uint64_t sub = (uint64_t)Df[i] * Kn[i] + (uint64_t)Df[N + i] * Kn[2 * N + i] + p2[i];
        packed_b[qi * pstride + (size_t)g * 2 * N + (size_t)c * N + i] = sub_mod(buf[ls(i)], red64(sub, (int)c), q);
    }
}

// The same result with one workgroup per query: four 256-lane teams take one
// (group, limb) polynomial each per round. A one-query chunk's grid of the
// kernel above (2 workgroups per group) waits for CU slots behind the
// database-pass workgroups running beside it; one workgroup starts as soon
// as one CU has room, like the Horner chain it feeds.
__global__ __launch_bounds__(1024) void k_pack_finalize_wg(const uint32_t* __restrict__ interm, uint32_t C,
                                                           const uint32_t* __restrict__ part2, size_t p2_stride,
                                                           const uint32_t* __restrict__ precomp,
                                                           const uint32_t* __restrict__ keys, size_t kstride,
                                                           uint32_t* __restrict__ packed_b, size_t pstride,
                                                           uint32_t g0, uint32_t ng, NttConst nc) {
    __shared__ uint32_t buf[4][POLY_LDS];
    __shared__ uint32_t twl[8 * N];
    const uint32_t tid = threadIdx.x, team = tid >> 8, t = tid & 255, qi = blockIdx.z;
    const uint32_t npoly = 2 * ng, rounds = (npoly + 3) / 4;
    uint32_t* b = buf[team];
    stage_twiddles(twl, nc, tid, 1024);  // the first round's barrier publishes them
    for (uint32_t r = 0; r < rounds; r++) {  // every lane runs every round (barriers inside)
        const uint32_t k = 4 * r + team;
        const bool active = k < npoly;
        const uint32_t g = g0 + (active ? k >> 1 : 0), c = k & 1;
        const uint32_t q = qof((int)c);
        if (active) {
            const uint32_t* pp = interm + (size_t)qi * 2 * C + (size_t)c * C + (size_t)g * N;
#pragma unroll
            for (int e = 0; e < 8; e++) b[ls(t + 256 * e)] = pp[t + 256 * e];
        }
        __syncthreads();
        ntt_fwd_lds(b, t, tw_lds(twl, (int)c), (int)c, active);
        if (active) {
            const uint32_t* Df = precomp + (size_t)g * PC_GROUP + PC_DF + (size_t)c * D_EFF * N;
            const uint32_t* Kn = keys + qi * kstride + BLK + (size_t)c * N;
            const uint32_t* p2 = part2 + qi * p2_stride + (size_t)g * 2 * N + (size_t)c * N;
#pragma unroll 2
            for (int e = 0; e < 8; e++) {
                const uint32_t i = t + 256 * e;
                const uint64_t sub = (uint64_t)Df[i] * Kn[i] + (uint64_t)Df[N + i] * Kn[2 * N + i] + p2[i];
                packed_b[qi * pstride + (size_t)g * 2 * N + (size_t)c * N + i] =
                    sub_mod(b[ls(i)], red64(sub, (int)c), q);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ Horner chain
// One workgroup per (response ciphertext, query); 4 teams of 256 lanes.
// acc (a0, a1, b0, b1 polys, NTT form) lives in registers between steps;
// per step: 4 INTTs, gadget split, 2 rounds of 4 NTTs, multiply-accumulate
// against the query's RGSW rows, + packed[i]. L == 1 copies packed[start].
// The RGSW rows are loaded once and kept in registers across steps; the
// step's packed parts are loaded during round 0, a phase ahead of their
// use, so HBM latency (long while a database pass streams through L2) hides
// behind the round-1 transforms. STREAMED: beside a streamed pack, each
// step first waits for its group's packed b-part (StreamWait).
template <bool STREAMED>
__global__ __launch_bounds__(1024) void k_horner(const uint32_t* __restrict__ pa_base, size_t pa_stride, uint32_t G, uint32_t Da,
                                                 uint32_t T, uint32_t* packed_b,
                                                 size_t pstride, const uint32_t* __restrict__ rgsw,
                                                 size_t rstride, uint64_t* __restrict__ out,
                                                 uint32_t* __restrict__ hstate, uint32_t g_lo, uint32_t g_hi,
                                                 NttConst nc, CrtConst cc, FusedPack fp, StreamWait sw) {
    __shared__ uint32_t buf[4][POLY_LDS];
    __shared__ uint32_t twl[8 * N];
    __shared__ uint32_t sw_abort;
    const uint32_t ct = blockIdx.x, qi = blockIdx.y, tid = threadIdx.x;
    const uint32_t team = tid >> 8, t = tid & 255;
    const uint32_t start = ct * Da;
    const uint32_t L = (G - start) < Da ? (G - start) : Da;
    const uint32_t top = start + L - 1;
    // this launch covers the chain's groups in [lo, hi), highest first
    const uint32_t lo = start > g_lo ? start : g_lo, hi = start + L < g_hi ? start + L : g_hi;
    if (lo >= hi) return;
    stage_twiddles(twl, nc, tid, 1024);
    const int tlimb = (int)(team & 1);
    const Tw tw = tw_lds(twl, tlimb);
    const uint32_t inv_n = nc.inv_n[tlimb], inv_n_p = nc.inv_n_p[tlimb];
    // Beside a streamed pack: wait until this chain's packed b-part of
    // group g is published (lane 0 spins and acquires, then the block
    // meets at a barrier; the hand-over rules in dev_common.h).
    // False: the wait exceeded its bound (every lane returns the same).
    const uint32_t* sw_done = sw.done + (size_t)(qi / sw.queries) * sw.stride;
    auto wait_packed = [&](uint32_t g) -> bool {
        if (tid == 0) {
            uint32_t spins = 0, ab = 0;
            while (sw.spin_limit == 0 ||
                   __hip_atomic_load(sw_done + g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
                if (++spins > sw.spin_limit) {
                    __hip_atomic_store(sw.status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ab = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(kSpinSleep);
            }
            fence_acquire_agent_done();
            sw_abort = ab;
        }
        __syncthreads();
        return !sw_abort;
    };
    constexpr bool streamed = STREAMED;
    uint64_t* o = out + ((size_t)qi * T + ct) * 4 * N;
    auto poison = [&]() {  // an invalid response: no residue is all-ones
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int e = 0; e < 2; e++) o[(size_t)p * N + tid + 1024 * e] = ~0ull;
    };
    if (fp.interm) {
        // this chain's packed b-parts of groups [lo, hi): team k%4 takes
        // (group lo + k/2, limb k%2), the limb of its own twiddles (as
        // k_pack_finalize_wg); written to packed_b and read back below
        // after the block's barrier
        const uint32_t npoly = 2 * (hi - lo), rounds = (npoly + 3) / 4;
        uint32_t* b = buf[team];
        for (uint32_t r = 0; r < rounds; r++) {  // every lane runs every round (barriers inside)
            const uint32_t k = 4 * r + team;
            const bool active = k < npoly;
            const uint32_t g = lo + (active ? k >> 1 : 0), c = k & 1;
            __syncthreads();  // twiddles staged, the previous round's reads done
            if (active) {
                const uint32_t* src = fp.interm + (size_t)qi * 2 * fp.C + (size_t)c * fp.C + (size_t)g * N;
#pragma unroll
                for (int e = 0; e < 8; e++) b[ls(t + 256 * e)] = src[t + 256 * e];
            }
            __syncthreads();
            ntt_fwd_lds(b, t, tw, (int)c, active);
            if (active) {
                const uint32_t q = qof((int)c);
                const uint32_t* Df = fp.precomp + (size_t)g * PC_GROUP + PC_DF + (size_t)c * D_EFF * N;
                const uint32_t* Kn = fp.keys + qi * fp.keys_stride + BLK + (size_t)c * N;
                const uint32_t* p2 = fp.part2 + qi * fp.part2_stride + (size_t)g * 2 * N + (size_t)c * N;
#pragma unroll 2
                for (int e = 0; e < 8; e++) {
                    const uint32_t i = t + 256 * e;
                    const uint64_t sub = (uint64_t)Df[i] * Kn[i] + (uint64_t)Df[N + i] * Kn[2 * N + i] + p2[i];
                    packed_b[qi * pstride + (size_t)g * 2 * N + (size_t)c * N + i] =
                        sub_mod(b[ls(i)], red64(sub, (int)c), q);
                }
            }
        }
        __syncthreads();  // the block's packed parts are in place
    }
    const uint32_t* pb = packed_b + qi * pstride;
    const uint32_t* rg = rgsw + qi * rstride;
    uint32_t* hs = hstate + ((size_t)qi * T + ct) * 4 * N;  // chain state between launches
    uint32_t acc[4][2];
    int first;
    if (top < g_hi) {  // the chain starts here: acc = packed[top]
        if constexpr (streamed)
            if (!wait_packed(top)) {
                poison();
                return;
            }
        const uint32_t* pa = pa_base + (size_t)top * pa_stride;
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t i = tid + 1024 * e;
            acc[0][e] = pa[i]; acc[1][e] = pa[N + i];
            acc[2][e] = pb[(size_t)top * 2 * N + i]; acc[3][e] = pb[(size_t)top * 2 * N + N + i];
        }
        first = (int)top - 1;
    } else {
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int e = 0; e < 2; e++) acc[p][e] = hs[(size_t)p * N + tid + 1024 * e];
        first = (int)hi - 1;
    }
    // the query's RGSW rows [round][c][e][j] (round 0: top_a / top_b, round
    // 1: bot_a / bot_b), the same for every step: loaded once, kept in
    // registers across the steps
    uint32_t rA[2][2][2][2], rB[2][2][2][2];
#pragma unroll
    for (int rd = 0; rd < 2; rd++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const uint32_t i = tid + 1024 * e;
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    rA[rd][c][e][j] = rg[(size_t)(2 * rd) * BLK + (size_t)(j * 2 + c) * N + i];
                    rB[rd][c][e][j] = rg[(size_t)(2 * rd + 1) * BLK + (size_t)(j * 2 + c) * N + i];
                }
            }
    for (int gs = first; gs >= (int)lo; gs--) {
        if constexpr (streamed)
            if (!wait_packed((uint32_t)gs)) {
                poison();
                return;
            }
        const int st = gs - (int)start;
        const uint32_t gi0 = start + (uint32_t)st;
        uint32_t pav[2][2], pbv[2][2];  // packed a / b parts [c][e], loaded a phase ahead
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int e = 0; e < 2; e++) buf[p][ls(tid + 1024 * e)] = acc[p][e];
        __syncthreads();
        ntt_inv_lds(buf[team], t, tw, tlimb, inv_n, inv_n_p);
        uint32_t da0[2][2], da1[2][2], db0[2][2], db1[2][2];  // [e][digit], residue mod Q0 / Q1
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t i = tid + 1024 * e;
            gadget2(buf[0][ls(i)], buf[1][ls(i)], cc, da0[e], da1[e]);
            gadget2(buf[2][ls(i)], buf[3][ls(i)], cc, db0[e], db1[e]);
        }
        uint64_t ra[2][2], rb[2][2];  // [limb][e]
#pragma unroll
        for (int round = 0; round < 2; round++) {
            __syncthreads();
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const uint32_t i = tid + 1024 * e;
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    buf[j * 2 + 0][ls(i)] = round ? db0[e][j] : da0[e][j];
                    buf[j * 2 + 1][ls(i)] = round ? db1[e][j] : da1[e][j];
                }
            }
            __syncthreads();
            ntt_fwd_lds(buf[team], t, tw, tlimb);
            // RGSW parts: 0 top_a, 1 top_b, 2 bot_a, 3 bot_b; [part][digit][limb][N]
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    const uint32_t i = tid + 1024 * e;
                    uint64_t sa = round ? ra[c][e] : 0, sb = round ? rb[c][e] : 0;
#pragma unroll
                    for (int j = 0; j < 2; j++) {
                        const uint64_t d = buf[j * 2 + c][ls(i)];
                        sa += d * rA[round][c][e][j];
                        sb += d * rB[round][c][e][j];
                    }
                    ra[c][e] = sa; rb[c][e] = sb;
                }
            if (round == 0) {  // the step's packed parts, issued before round 1's transforms
#pragma unroll
                for (int c = 0; c < 2; c++)
#pragma unroll
                    for (int e = 0; e < 2; e++) {
                        const uint32_t i = tid + 1024 * e;
                        pav[c][e] = pa_base[(size_t)gi0 * pa_stride + (size_t)c * N + i];
                        pbv[c][e] = pb[(size_t)gi0 * 2 * N + (size_t)c * N + i];
                    }
            }
        }
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const uint32_t q = qof(c);
                acc[c][e] = add_mod(red64(ra[c][e], c), pav[c][e], q);
                acc[2 + c][e] = add_mod(red64(rb[c][e], c), pbv[c][e], q);
            }
        __syncthreads();
    }
    if (lo > start) {  // the chain continues in a later launch
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int e = 0; e < 2; e++) hs[(size_t)p * N + tid + 1024 * e] = acc[p][e];
        return;
    }
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int e = 0; e < 2; e++) o[(size_t)p * N + tid + 1024 * e] = acc[p][e];
}

// ------------------------------------------------------------ streamed pack
// Beside a streamed database pass (launch_pack_stream): workgroup b packs
// queries [b*ps.queries, ...) group by group from the top; team k of
// round r takes polynomial 4r + k = (query, limb), the limb of its own
// twiddles. The subtrahend of round 0 (collapse sums + Df x Kn, in place
// before the launch) is loaded while lane 0 waits for the group.
__global__ __launch_bounds__(1024) void k_pack_stream(PackStream ps, const uint32_t* __restrict__ part2,
                                                      size_t p2_stride, const uint32_t* __restrict__ precomp,
                                                      const uint32_t* __restrict__ keys, size_t kstride,
                                                      uint32_t* __restrict__ packed_b, size_t pstride, uint32_t G,
                                                      uint32_t nq, NttConst nc) {
    __shared__ uint32_t buf[4][POLY_LDS];
    __shared__ uint32_t twl[8 * N];
    __shared__ uint32_t abort_flag;
    const uint32_t tid = threadIdx.x, team = tid >> 8, t = tid & 255, blk = blockIdx.x;
    const uint32_t q0 = blk * ps.queries;
    const uint32_t nqb = nq - q0 < ps.queries ? nq - q0 : ps.queries;
    const uint32_t npoly = 2 * nqb, rounds = (npoly + 3) / 4;
    const int c = (int)(team & 1);
    const uint32_t q = qof(c);
    const Tw tw = tw_lds(twl, c);
    const size_t sstride = 2 * (size_t)ps.C;  // one row slice of the pass's partials
    stage_twiddles(twl, nc, tid, 1024);        // published by the first barrier below
    for (int g = (int)G - 1; g >= 0; g--) {
        for (uint32_t r = 0; r < rounds; r++) {  // every lane runs every round (barriers inside)
            const uint32_t k = 4 * r + team;
            const bool active = k < npoly;
            const uint32_t qi = q0 + (active ? k >> 1 : 0);
            // Df x Kn (precomp and keys, in place before the launch) now;
            // the collapse sums after the wait, since a pass may produce them
            // with the group (the collapse in the pass's launch)
            uint32_t sub[8];
            if (active) {
                const uint32_t* Df = precomp + (size_t)g * PC_GROUP + PC_DF + (size_t)c * D_EFF * N;
                const uint32_t* Kn = keys + qi * kstride + BLK + (size_t)c * N;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const uint32_t x = t + 256 * e;
                    sub[e] = red64((uint64_t)Df[x] * Kn[x] + (uint64_t)Df[N + x] * Kn[2 * N + x], c);
                }
            }
            if (r == 0) {
                if (tid == 0) {
                    uint32_t spins = 0, ab = 0;
                    while (ps.spin_limit == 0 || __hip_atomic_load(ps.pass_done + g, __ATOMIC_RELAXED,
                                                                   __HIP_MEMORY_SCOPE_AGENT) < ps.pass_target) {
                        if (++spins > ps.spin_limit) {
                            __hip_atomic_store(ps.status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            ab = 1;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(kSpinSleep);
                    }
                    fence_acquire_agent_done();  // the pass's partials
                    abort_flag = ab;
                }
                __syncthreads();
                if (abort_flag) return;
            }
            if (active) {
                // partials are residues (< 2^27) and at most 16 are summed:
                // 32-bit sums, the 8 loads of a slice in flight together
                const uint32_t* src = ps.part + qi * ps.part_qstride + (size_t)c * ps.C + (size_t)g * N + t;
                const uint32_t* p2 = part2 + qi * p2_stride + (size_t)g * 2 * N + (size_t)c * N + t;
                uint32_t a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                for (uint32_t sl = 0; sl < ps.slices; sl++) {
#pragma unroll
                    for (int e = 0; e < 8; e++) a[e] += src[sl * sstride + 256 * e];
                }
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    buf[team][ls(t + 256 * e)] = red64(a[e], c);
                    sub[e] = red64((uint64_t)sub[e] + p2[256 * e], c);
                }
            }
            __syncthreads();
            ntt_fwd_lds(buf[team], t, tw, c, active);
            if (active) {
                uint32_t* dst = packed_b + qi * pstride + (size_t)g * 2 * N + (size_t)c * N;
#pragma unroll
                for (int e = 0; e < 8; e++) dst[t + 256 * e] = sub_mod(buf[team][ls(t + 256 * e)], sub[e], q);
            }
            __syncthreads();  // the buffers are free for the next round
        }
        // publish the group: every wave's stores complete, then one release
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) {
            fence_release_agent_done();
            __hip_atomic_store(ps.packed_done + (size_t)blk * G + g, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

}  // namespace

hipError_t launch_unmarshal(const uint64_t* d_q, size_t count, size_t R, double* bq, uint32_t* keys,
                            uint32_t* rgsw, hipStream_t s) {
    const size_t total = count * (2 * R + 6 * BLK);
    size_t blocks = (total + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_unmarshal, dim3((unsigned)blocks), dim3(256), 0, s, d_q, total, R, bq, keys, rgsw);
    return hipGetLastError();
}

namespace {
// Plain grid-stride copy, 16 bytes per lane: the reference point that
// scales the memory-traffic counters of the other kernels.
__global__ __launch_bounds__(256) void k_copy16(const u32x4* __restrict__ src, u32x4* __restrict__ dst, size_t n) {
    for (size_t x = (size_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (size_t)gridDim.x * 256)
        dst[x] = __builtin_nontemporal_load(src + x);
}
}  // namespace

hipError_t launch_copy16(const void* src, void* dst, size_t bytes, hipStream_t s) {
    if (bytes % 16) return hipErrorInvalidValue;
    if (bytes == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy16, dim3(8192), dim3(256), 0, s, static_cast<const u32x4*>(src), static_cast<u32x4*>(dst),
                       bytes / 16);
    return hipGetLastError();
}

hipError_t launch_ntt_rns(uint32_t* polys, size_t npoly, const NttConst& nc, bool inverse, hipStream_t s) {
    if (npoly == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ntt_rns, dim3((unsigned)(2 * npoly)), dim3(256), 0, s, polys, nc, inverse ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_reduce_slices(const uint32_t* part, size_t part_stride, uint32_t S, size_t C, uint32_t* interm,
                                size_t nq, size_t col0, size_t ncols, hipStream_t s) {
    if (ncols % 1024 || col0 % 4 || col0 + ncols > C) return hipErrorInvalidValue;
    if (nq == 0 || ncols == 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_slices, dim3((unsigned)(ncols / 256), 2, (unsigned)nq), dim3(256), 0, s, part,
                       part_stride, S, (uint32_t)C, (uint32_t)col0, (uint32_t)ncols, interm);
    return hipGetLastError();
}

hipError_t launch_pack_finalize(const uint32_t* interm, size_t C, const uint32_t* part2, size_t part2_stride,
                                const uint32_t* precomp, size_t g0, size_t ng, const uint32_t* keys,
                                size_t keys_stride, uint32_t* packed_b, size_t packed_stride, size_t nq,
                                const NttConst& nc, hipStream_t s) {
    if (ng == 0 || nq == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_finalize, dim3((unsigned)ng, 2, (unsigned)nq), dim3(256), 0, s, interm, (uint32_t)C,
                       part2, part2_stride, precomp, keys, keys_stride, packed_b, packed_stride, (uint32_t)g0, nc);
    return hipGetLastError();
}

hipError_t launch_pack_finalize_wg(const uint32_t* interm, size_t C, const uint32_t* part2, size_t part2_stride,
                                   const uint32_t* precomp, size_t g0, size_t ng, const uint32_t* keys,
                                   size_t keys_stride, uint32_t* packed_b, size_t packed_stride, size_t nq,
                                   const NttConst& nc, hipStream_t s) {
    if (ng == 0 || nq == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_finalize_wg, dim3(1, 1, (unsigned)nq), dim3(1024), 0, s, interm, (uint32_t)C, part2,
                       part2_stride, precomp, keys, keys_stride, packed_b, packed_stride, (uint32_t)g0, (uint32_t)ng,
                       nc);
    return hipGetLastError();
}

hipError_t launch_pack_stream(const PackStream& ps, const uint32_t* part2, size_t part2_stride,
                              const uint32_t* precomp, const uint32_t* keys, size_t keys_stride, uint32_t* packed_b,
                              size_t packed_stride, size_t G, size_t nq, const NttConst& nc, hipStream_t s) {
    if (nq == 0 || G == 0) return hipSuccess;
    if (!ps.part || !ps.pass_done || !ps.packed_done || !ps.status || ps.slices == 0 || ps.slices > 16 ||
        ps.C < G * N || ps.queries == 0 || ps.queries > kPackQueries)  // at most 16 partials below 2^27 keep
        return hipErrorInvalidValue;                                    // the 32-bit sums exact
    const unsigned blocks = (unsigned)((nq + ps.queries - 1) / ps.queries);
    hipLaunchKernelGGL(k_pack_stream, dim3(blocks), dim3(1024), 0, s, ps, part2, part2_stride, precomp, keys,
                       keys_stride, packed_b, packed_stride, (uint32_t)G, (uint32_t)nq, nc);
    return hipGetLastError();
}

hipError_t launch_horner(const uint32_t* pa_base, size_t pa_stride, size_t G, size_t Da, size_t T,
                         uint32_t* packed_b, size_t packed_stride, const uint32_t* rgsw, size_t rgsw_stride,
                         uint64_t* out, uint32_t* hstate, size_t g_lo, size_t g_hi, size_t nq, const NttConst& nc,
                         const CrtConst& cc, hipStream_t s, const FusedPack& pack, const StreamWait& wait) {
    if (nq == 0 || g_lo >= g_hi) return hipSuccess;
    if (wait.done && (!wait.status || wait.queries == 0)) return hipErrorInvalidValue;
    if (wait.done)
        hipLaunchKernelGGL(k_horner<true>, dim3((unsigned)T, (unsigned)nq), dim3(1024), 0, s, pa_base, pa_stride,
                           (uint32_t)G, (uint32_t)Da, (uint32_t)T, packed_b, packed_stride, rgsw, rgsw_stride, out,
                           hstate, (uint32_t)g_lo, (uint32_t)g_hi, nc, cc, pack, wait);
    else
        hipLaunchKernelGGL(k_horner<false>, dim3((unsigned)T, (unsigned)nq), dim3(1024), 0, s, pa_base, pa_stride,
                           (uint32_t)G, (uint32_t)Da, (uint32_t)T, packed_b, packed_stride, rgsw, rgsw_stride, out,
                           hstate, (uint32_t)g_lo, (uint32_t)g_hi, nc, cc, pack, wait);
    return hipGetLastError();
}

}  // namespace ipx
