// Database-pass kernels for gfx950 over the resident DB layout.
//
// Resident layout ("byte planes in 16-row blocks"): the encoded u16 value
// of (row r, col c) is split into its low and high byte; byte plane p of
// block kb = r/16 holds, for every column, the 16 bytes of rows 16kb..16kb+15:
//   plane byte of (r, c) at ((r/16)*2 + p)*C*16 + c*16 + r%16.
// A lane's 16-byte load is one column's 16 rows of one plane: exactly the
// K-slice of an int8 MFMA B operand (after a sign flip, no byte shuffles),
// and consecutive lanes read consecutive 16-byte units, so every load
// instruction of a wavefront is fully coalesced.
//
// Every kernel computes the reference's mat-vec (src/gpu_online.cu:23-41)
//   interm[q][c][col] = sum_r db[r][col] * b_{q,c}[r] mod q_c
// exactly, over a slice of the rows; slices are summed mod q afterwards.
#include <string>

#include "kernels.h"
#include "dbpass_direct.h"

namespace ipx {

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// a mod q for an integer-valued double 0 <= a < 2^53.
__device__ __forceinline__ uint32_t fmod_exact(double a, double q, double qinv) {
    const double t = floor(a * qinv);
    double r = fma(-t, q, a);  // exact: the true value is an integer below 2q in magnitude
    if (r < 0) r += q;
    if (r >= q) r -= q;
    return (uint32_t)r;
}

// ------------------------------------------------------------ relayout
// Row-major rows [0, 16*gridDim.y) of `src` -> blocks [kb0, kb0 + gridDim.y) of the planes.
__global__ __launch_bounds__(256) void k_to_planes(const uint16_t* __restrict__ src, uint8_t* __restrict__ dst,
                                                   uint32_t C, uint32_t kb0) {
    const uint32_t col = blockIdx.x * 256 + threadIdx.x;
    const uint32_t kbl = blockIdx.y;
    uint32_t lo[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const uint32_t v = src[((size_t)kbl * 16 + e) * C + col];
        lo[e >> 2] |= (v & 255u) << (8 * (e & 3));
        hi[e >> 2] |= (v >> 8) << (8 * (e & 3));
    }
    const size_t kb = (size_t)kb0 + kbl;
    *reinterpret_cast<u32x4*>(dst + ((kb * 2 + 0) * C + col) * 16) = u32x4{lo[0], lo[1], lo[2], lo[3]};
    *reinterpret_cast<u32x4*>(dst + ((kb * 2 + 1) * C + col) * 16) = u32x4{hi[0], hi[1], hi[2], hi[3]};
}

// ------------------------------------------------------------ plain pass (fp64)
// One query. Thread = one column; per 16-row block it loads the column's low
// and high plane (16 bytes each) and uses wave-uniform query words.
// Products db*b < 2^43 accumulate exactly in fp64 FMA for 1024 rows
// (< 2^53), in four interleaved accumulators per limb; every 1024 rows the
// sums are reduced mod q and folded into a canonical u32.
// out[limb][OC] for this slice at out + slice*2*OC.
template <int U>
__global__ __launch_bounds__(256) void k_dbpass_f64(const uint8_t* __restrict__ db, const double* __restrict__ bq,
                                                    uint32_t* __restrict__ out, uint32_t R, uint32_t C, uint32_t OC,
                                                    uint32_t kbps) {
    const uint32_t col = blockIdx.x * 256 + threadIdx.x;
    const uint32_t kb0 = blockIdx.y * kbps;
    const u32x4* p = reinterpret_cast<const u32x4*>(db) + (size_t)kb0 * 2 * C + col;
    const size_t kstep = (size_t)2 * C;  // one block (both planes), in 16-byte units
    const double* b0 = bq + (size_t)kb0 * 16;
    const double* b1 = b0 + R;
    const double dq0 = (double)Q0, dq1 = (double)Q1, iq0 = 1.0 / dq0, iq1 = 1.0 / dq1;
    uint32_t m0 = 0, m1 = 0;
    double a0[4] = {0, 0, 0, 0}, a1[4] = {0, 0, 0, 0};
    for (uint32_t kb = 0; kb < kbps; kb += U) {
        u32x4 lo[U], hi[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            lo[u] = __builtin_nontemporal_load(p + (size_t)(kb + u) * kstep);
            hi[u] = __builtin_nontemporal_load(p + (size_t)(kb + u) * kstep + C);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const double* x0 = b0 + (size_t)(kb + u) * 16;
            const double* x1 = b1 + (size_t)(kb + u) * 16;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                // rows 4j..4j+3: interleave the low and high bytes into two u16 pairs
                const uint32_t w01 = __builtin_amdgcn_perm(hi[u][j], lo[u][j], 0x05010400u);
                const uint32_t w23 = __builtin_amdgcn_perm(hi[u][j], lo[u][j], 0x07030602u);
                const double d[4] = {(double)(w01 & 0xffffu), (double)(w01 >> 16), (double)(w23 & 0xffffu),
                                     (double)(w23 >> 16)};
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    a0[e] = fma(d[e], x0[4 * j + e], a0[e]);
                    a1[e] = fma(d[e], x1[4 * j + e], a1[e]);
                }
            }
        }
        if (((kb + U) & 63) == 0 || kb + U >= kbps) {  // every 64 blocks = 1024 rows
            m0 = add_mod(m0, fmod_exact((a0[0] + a0[1]) + (a0[2] + a0[3]), dq0, iq0), Q0);
            m1 = add_mod(m1, fmod_exact((a1[0] + a1[1]) + (a1[2] + a1[3]), dq1, iq1), Q1);
#pragma unroll
            for (int e = 0; e < 4; e++) a0[e] = a1[e] = 0.0;
        }
    }
    uint32_t* o = out + (size_t)blockIdx.y * 2 * OC;
    o[col] = m0;
    o[OC + col] = m1;
}

// ------------------------------------------------------------ batched pass (int8 MFMA)
// Query side: b = sum_k d_k 256^k with balanced base-256 digits d_0..d_2 in
// [-128, 127] and d_3 in [0, 8] (b < 2^27). DB side: v = 256*h' + l' + 32896
// with h' = hi byte - 128, l' = lo byte - 128 (a byte XOR 0x80 read as int8).
// So   sum_r v_r b_r = sum_k 256^k (256*H_k + L_k) + 32896 * sum_r b_r,
// with H_k = sum_r h'_r d_{k,r} and L_k = sum_r l'_r d_{k,r}: int8 x int8 dot
// products; with at most 2^16 rows per slice |H_k|, |L_k| <= 2^30, exact in
// i32.
//
// A operand rows m = (q*2 + c)*4 + k (pair pr = q*2 + c, digit k), stored as
// dig[R/16][Mp][16 bytes] (byte e of row m in block kb = digit of row 16kb+e).
// MFMA 16x16x64 i8 lane mapping (checked on the device): A/B lane l holds 16
// K-bytes of M row / N col l%16 at K offset 16*(l/16); the i32 output holds
// rows 4*(l/16)+i, col l%16 in element i. A lane thus owns all four digits
// of one pair for one column and recombines them locally.

// dig for query pairs [0, 2*nq) and zero digits for padding pairs < Mp/4.
__global__ __launch_bounds__(256) void k_digits(const double* __restrict__ bq, uint32_t R, uint32_t nq,
                                                uint32_t Mp, int8_t* __restrict__ dig) {
    const uint32_t nkb = R / 16;
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;
    if (x >= nkb * (Mp / 4)) return;
    const uint32_t pr = x / nkb, kb = x % nkb;
    u32x4 o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = u32x4{0, 0, 0, 0};
    if (pr < 2 * nq) {
        const double* b = bq + (size_t)pr * R + (size_t)kb * 16;  // bq [q][c][R] = [pr][R]
#pragma unroll
        for (int e = 0; e < 16; e++) {
            int32_t v = (int32_t)(uint32_t)b[e];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                int32_t d = v;
                if (k < 3) {
                    d = ((v & 255) ^ 128) - 128;  // balanced digit in [-128, 127]
                    v = (v - d) >> 8;
                }
                o[k][e >> 2] |= ((uint32_t)d & 255u) << (8 * (e & 3));
            }
        }
    }
    u32x4* dst = reinterpret_cast<u32x4*>(dig + ((size_t)kb * Mp + (size_t)pr * 4) * 16);
#pragma unroll
    for (int k = 0; k < 4; k++) dst[k] = o[k];
}

// 7-digit rows (17..32 queries): both limbs of a query share one digit
// expansion, that of the CRT value B_r in [0, Q0*Q1) with B_r = b_{q,c}[r]
// mod q_c. sum_r v_r B_r is congruent to either limb's sum mod q_c, and
// Q0*Q1 < 2^53 needs 6 balanced digits in [-128, 127] and a top digit in
// [0, 32]: 7 rows per query instead of 2 x 4. Rows m = 7q + k.
constexpr uint32_t pow_mod_c(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1;
    b %= m;
    for (; e; e >>= 1, b = b * b % m)
        if (e & 1) r = r * b % m;
    return (uint32_t)r;
}
constexpr uint32_t kQ1InvQ0 = pow_mod_c(Q1, Q0 - 2, Q0);  // Q0 is prime

// pmajor: rows (Mp/7) k + q (the power-row form) instead of 7q + k.
__global__ __launch_bounds__(256) void k_digits7(const double* __restrict__ bq, uint32_t R, uint32_t nq,
                                                 uint32_t Mp, int8_t* __restrict__ dig, uint32_t pmajor) {
    const uint32_t nkb = R / 16;
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;
    if (x >= nkb * (Mp / 7)) return;
    const uint32_t q = x / nkb, kb = x % nkb;
    u32x4 o[7];
#pragma unroll
    for (int k = 0; k < 7; k++) o[k] = u32x4{0, 0, 0, 0};
    if (q < nq) {
        const double* b0 = bq + (size_t)q * 2 * R + (size_t)kb * 16;  // bq [q][c][R]
        const double* b1 = b0 + R;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const uint32_t r0 = (uint32_t)b0[e], r1 = (uint32_t)b1[e];
            // Garner: B = r1 + Q1 * ((r0 - r1) * Q1^-1 mod Q0)
            const uint32_t t = mul_mod(sub_mod(r0, r1 % Q0, Q0), kQ1InvQ0, 0);
            int64_t v = (int64_t)r1 + (int64_t)Q1 * t;
#pragma unroll
            for (int k = 0; k < 7; k++) {
                int64_t d = v;
                if (k < 6) {
                    d = ((v & 255) ^ 128) - 128;
                    v = (v - d) >> 8;
                }
                o[k][e >> 2] |= ((uint32_t)d & 255u) << (8 * (e & 3));
            }
        }
    }
    u32x4* dst = reinterpret_cast<u32x4*>(dig + (size_t)kb * Mp * 16);
#pragma unroll
    for (int k = 0; k < 7; k++) dst[pmajor ? (Mp / 7) * k + q : 7 * q + k] = o[k];
}

// Residue mod q_c of one output of the 7-digit form from the dot products
// t_k = 256*H_k + L_k (|t_k| < 2^38.01, at most 2^16 rows per slice):
//   sum_k 256^k t_k + 32896*S_c.
// The terms k < 3 are summed exactly (< 2^54.02 in magnitude, lifted by a
// multiple of q_c); the others are reduced first and multiplied by
// 256^k mod q_c (each product < 2^54); the sum stays below 2^57.1.
__device__ __forceinline__ uint32_t crt7_residue(const int64_t (&t)[7], uint32_t Sc, int c) {
    const uint64_t qc = c ? Q1 : Q0;
    constexpr uint32_t w0[4] = {pow_mod_c(256, 3, Q0), pow_mod_c(256, 4, Q0), pow_mod_c(256, 5, Q0),
                                pow_mod_c(256, 6, Q0)};
    constexpr uint32_t w1[4] = {pow_mod_c(256, 3, Q1), pow_mod_c(256, 4, Q1), pow_mod_c(256, 5, Q1),
                                pow_mod_c(256, 6, Q1)};
    const int64_t lo = t[0] + t[1] * 256 + t[2] * 65536;
    uint64_t s = (uint64_t)(lo + (int64_t)(qc * ((1ull << 55) / qc))) + (uint64_t)32896 * Sc;
    const int64_t off40 = (int64_t)(qc * ((1ull << 40) / qc));
#pragma unroll
    for (int k = 3; k < 7; k++)
        s += (uint64_t)red64((uint64_t)(t[k] + off40), c) * (c ? w1[k - 3] : w0[k - 3]);
    return red64(s, c);
}

// Residue mod q_c of one output of the power-row form from p_0..p_7
// (|p_k| <= 2^30): sum_k 256^k p_k + 32896*S_c. The terms k < 4 are summed
// exactly (< 2^54.01 in magnitude, lifted by a multiple of q_c); the others
// are reduced first and multiplied by 256^k mod q_c (each product < 2^54);
// the sum stays below 2^57.
__device__ __forceinline__ uint32_t pow_residue(const int32_t (&p)[8], uint32_t Sc, int c) {
    const uint64_t qc = c ? Q1 : Q0;
    constexpr uint32_t w0[4] = {pow_mod_c(256, 4, Q0), pow_mod_c(256, 5, Q0), pow_mod_c(256, 6, Q0),
                                pow_mod_c(256, 7, Q0)};
    constexpr uint32_t w1[4] = {pow_mod_c(256, 4, Q1), pow_mod_c(256, 5, Q1), pow_mod_c(256, 6, Q1),
                                pow_mod_c(256, 7, Q1)};
    const int64_t lo = (int64_t)p[0] + (int64_t)p[1] * 256 + (int64_t)p[2] * 65536 + (int64_t)p[3] * 16777216;
    uint64_t s = (uint64_t)(lo + (int64_t)(qc * ((1ull << 55) / qc))) + (uint64_t)32896 * Sc;
    const int64_t off32 = (int64_t)(qc * ((1ull << 32) / qc));
#pragma unroll
    for (int k = 4; k < 8; k++)
        s += (uint64_t)red64((uint64_t)(p[k] + off32), c) * (c ? w1[k - 4] : w0[k - 4]);
    return red64(s, c);
}

// S[pr] = sum_r b_pr[r] mod q_c, the DB-offset correction. One block of
// 1024 lanes per pair, R a multiple of 2048: R/2048 16-byte loads per lane,
// all in flight together (a one-query answer waits for this kernel).
__global__ __launch_bounds__(1024) void k_bsum(const double* __restrict__ bq, uint32_t R, uint32_t* __restrict__ S) {
    __shared__ uint64_t red[16];
    const uint32_t pr = blockIdx.x;
    const double2* b = reinterpret_cast<const double2*>(bq + (size_t)pr * R) + threadIdx.x;
    uint64_t a = 0;
#pragma unroll 16
    for (uint32_t i = 0; i < R / 2048; i++) {
        const double2 v = b[(size_t)i * 1024];
        a += (uint64_t)v.x + (uint64_t)v.y;
    }
    const uint64_t sum = block_sum_u64<1024>(a, red);
    if (threadIdx.x == 0) S[pr] = red64(sum, (int)(pr & 1));
}


// Column tile and row slice of this workgroup: the 2-D grid's, or in a
// streamed pass the 1-D index read as (tile counted from the top, slice).
__device__ __forceinline__ void pass_coords(const PassStream& ps, uint32_t& xt, uint32_t& slice) {
    if (ps.group_done) {
        slice = blockIdx.x & ((1u << ps.slice_shift) - 1);
        xt = ps.tiles - 1 - (blockIdx.x >> ps.slice_shift);
    } else {
        xt = blockIdx.x;
        slice = blockIdx.y;
    }
}

// Streamed pass, after the workgroup's stores of column tile xt: count the
// workgroup into its group. One agent-scope release per workgroup (it
// writes back the L2): every wave first waits for its own stores, the
// barrier orders them before lane 0's release.
__device__ __forceinline__ void pass_handover(const PassStream& ps, uint32_t xt) {
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are complete
    __syncthreads();
    if (threadIdx.x == 0) {
        fence_release_agent_done();
        __hip_atomic_fetch_add(ps.group_done + ps.group_base + xt / ps.tiles_per_group, 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Up to 2*MT queries: dbpass_direct_body over the grid's (tile, slice).
template <int MT, int NT>
__global__ __launch_bounds__(256, 2) void k_dbpass_mfma(const uint8_t* __restrict__ db,
                                                        const int8_t* __restrict__ dig,
                                                        const uint32_t* __restrict__ S,
                                                        uint32_t* __restrict__ out, size_t out_qstride,
                                                        uint32_t C, uint32_t OC, uint32_t nq, uint32_t kbps,
                                                        PassStream ps) {
    uint32_t xt, slice;
    pass_coords(ps, xt, slice);
    dbpass_direct_body<MT, NT>(db, dig, S, out, out_qstride, C, OC, nq, kbps, xt, slice);
    if (ps.group_done) pass_handover(ps, xt);
}

// Many queries (MT = 8 or 16 tiles of per-limb digits, up to 16 or 32
// queries; MT = 7 or 14 tiles of CRT digits, below): as above, but the
// block's WAVES waves share the step's digit tile (Mp x 64 bytes) through an
// LDS double buffer (one barrier per step). Every block reads the whole
// digit tile each step, twice the bytes of its 64 columns' DB operands at 4
// waves; 8 waves (both waves of every SIMD of a CU) halve that ratio, and
// the digit loads are what the vector-memory path spends most on. Every
// global load is issued L = DEPTH-1 steps before its step: the DB operands
// into a register ring of DEPTH slots, the thread's share of the digit tile
// into a ring of L slots (stored to LDS one step ahead). Vector-memory loads complete in issue
// order, so a load issued closer to its use (a digit prefetch one step ahead)
// would make the wait for it also wait for every DB operand issued before it;
// with one lead for all loads each wait covers exactly one step's loads. Past
// the slice end the loads repeat the last step (in bounds, never used), so
// every step issues the same instructions.
//
// MT = 14 is the 7-digit form (rows 7q + k, 32 queries in 224 rows): a
// query's digits straddle the lanes' 4-row output groups, so the block
// regroups its accumulators through LDS (the digit ring's memory, free after
// the last step), half of the queries at a time, and each lane recombines
// whole queries.
//
// POW = 1 (MT = 7 QT, QT = 1 or 2 tiles of 16 queries, digit rows
// 16 QT k + q: digit k of query q) accumulates by power of 256 instead of
// by (digit, plane):
//   sum_r v_r B_r = sum_{k=0..7} 256^k p_k + 32896 sum_r B_r,
//   p_k = sum_r h'_r d_{k-1,r} + l'_r d_{k,r}   (d_{-1} = d_7 = 0),
// so the high-plane MFMA of digit tile a adds into output tile a + QT (the
// next power of the same 16 queries) and the low-plane one into tile a:
// 8 QT output tiles of one accumulator each instead of 7 QT of two. The
// freed registers hold a second column tile (NT = 2), which halves the
// digit-tile bytes loaded and read from LDS per DB byte. A lane's output
// rows are powers k = a/QT of queries 16(a%QT) + 4(l/16) + i: every lane
// recombines its own queries from registers. |p_k| <= 2 * 2^14 per row: at
// most 2^15 rows per slice keep it within 2^30.
template <int MT, int NT, int OCC, int DEPTH, int WAVES = 4, int POW = 0>
__global__ __launch_bounds__(64 * WAVES, OCC) void k_dbpass_mfma_sa(const uint8_t* __restrict__ db,
                                                           const int8_t* __restrict__ dig,
                                                           const uint32_t* __restrict__ S,
                                                           uint32_t* __restrict__ out, size_t out_qstride,
                                                           uint32_t C, uint32_t OC, uint32_t nq, uint32_t kbps,
                                                           PassStream ps) {
    static_assert(DEPTH >= 3, "the digit ring needs two slots");
    constexpr int L = DEPTH - 1;            // lead of every load, in steps
    constexpr int P = DEPTH * L;            // period of both rings' slot pattern
    static_assert(!POW || MT == 7 || MT == 14, "power rows come from the 7-digit expansion");
    constexpr bool CRT7 = MT == 14 && !POW;
    static_assert(!CRT7 || NT == 1, "the 7-digit epilogue covers one column tile per wave");
    constexpr int QT = POW ? MT / 7 : 1;                      // power rows: tiles per power
    constexpr int MA = POW ? MT + QT : MT, PL = POW ? 1 : 2;  // output tiles, accumulators per tile
    constexpr uint32_t Mp = 16 * MT;
    constexpr uint32_t TH = 64 * WAVES;              // threads per block
    constexpr uint32_t AU = (Mp * 4 + TH - 1) / TH;  // 16-byte units of the step's digit tile per thread
    constexpr bool APART = AU * TH != Mp * 4;        // the last unit exists for some threads only
    // 7-digit epilogue: per wave E[plane][col][row] int32, 112 rows (16
    // queries) of a half, rows padded to 116 words (spreads the column reads over banks)
    constexpr uint32_t EPS = 116, EPW = 2 * 16 * EPS;
    constexpr uint32_t AT_UNITS = 2 * 4 * Mp, EP_UNITS = CRT7 ? WAVES * EPW / 4 : 0;
    __shared__ i32x4 smem[AT_UNITS > EP_UNITS ? AT_UNITS : EP_UNITS];
    i32x4(*atile)[4 * Mp] = reinterpret_cast<i32x4(*)[4 * Mp]>(smem);  // [buffer][block of the step][row]
    const uint32_t tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    uint32_t xt, slice;
    pass_coords(ps, xt, slice);
    const uint32_t col0 = (xt * WAVES + w) * NT * 16;
    const uint32_t kb0 = slice * kbps;
    const uint32_t nks = kbps / 4;
    const i32x4* bsrc = reinterpret_cast<const i32x4*>(db) + (size_t)(kb0 + (l >> 4)) * 2 * C + col0 + (l & 15);
    const size_t bstep = (size_t)8 * C;
    // the step's digit tile is 4*Mp contiguous 16-byte units starting at block 4*ks
    const i32x4* atile_src = reinterpret_cast<const i32x4*>(dig) + (size_t)kb0 * Mp;
    const i32x4 flip = {(int32_t)0x80808080, (int32_t)0x80808080, (int32_t)0x80808080, (int32_t)0x80808080};
    i32x4 acc[MA][NT][PL];
#pragma unroll
    for (int a = 0; a < MA; a++)
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int pl = 0; pl < PL; pl++) acc[a][n][pl] = i32x4{0, 0, 0, 0};
    i32x4 v[DEPTH][NT][2];
    i32x4 ar[L][AU];
    auto issue = [&](int bs, int as, uint32_t step) {
        const uint32_t k = step < nks ? step : nks - 1;
#pragma unroll
        for (int u = 0; u < (int)AU; u++) {
            // a thread past the tile's end reloads its last unit (in bounds, not stored)
            const uint32_t x = !APART || tid + TH * u < 4 * Mp ? tid + TH * u : 4 * Mp - 1;
            ar[as][u] = atile_src[(size_t)k * 4 * Mp + x];
        }
#pragma unroll
        for (int n = 0; n < NT; n++) {
            v[bs][n][0] = __builtin_nontemporal_load(bsrc + k * bstep + n * 16);
            v[bs][n][1] = __builtin_nontemporal_load(bsrc + k * bstep + n * 16 + C);
        }
    };
    auto store_a = [&](int buf, int as) {
#pragma unroll
        for (int u = 0; u < (int)AU; u++)
            if (!APART || tid + TH * u < 4 * Mp) atile[buf][tid + TH * u] = ar[as][u];
    };
    auto compute_pow = [&](int bs, int buf) {
        i32x4 (&B)[NT][2] = v[bs];  // flipped in place: the slot is reloaded before its next use
#pragma unroll
        for (int n = 0; n < NT; n++) {
            B[n][0] ^= flip;
            B[n][1] ^= flip;
        }
        constexpr int AW = 4;  // digit tiles read ahead of their MFMAs
        i32x4 Aw[AW];
#pragma unroll
        for (int a = 0; a < AW; a++) Aw[a] = atile[buf][(l >> 4) * Mp + a * 16 + (l & 15)];
#pragma unroll
        for (int a = 0; a < MT; a++) {
            const i32x4 A = Aw[a % AW];
            if (a + AW < MT) Aw[a % AW] = atile[buf][(l >> 4) * Mp + (a + AW) * 16 + (l & 15)];
#pragma unroll
            for (int n = 0; n < NT; n++) {
                acc[a][n][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, B[n][0], acc[a][n][0], 0, 0, 0);
                acc[a + QT][n][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, B[n][1], acc[a + QT][n][0], 0, 0, 0);
            }
        }
        __builtin_amdgcn_sched_group_barrier(0x100, AW, 0);
#pragma unroll
        for (int a = 0; a + AW < MT; a++) {
            __builtin_amdgcn_sched_group_barrier(0x008, 2 * NT, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 2 * NT * AW, 0);
    };
    auto compute = [&](int bs, int buf) {
        if constexpr (POW) {
            compute_pow(bs, buf);
            return;
        }
#pragma unroll
        for (int n = 0; n < NT; n++) {
            const i32x4 Bl = v[bs][n][0] ^ flip, Bh = v[bs][n][1] ^ flip;
            constexpr int AW = 4;  // digit tiles read ahead of their MFMAs
            i32x4 Aw[AW];
#pragma unroll
            for (int a = 0; a < AW; a++) Aw[a] = atile[buf][(l >> 4) * Mp + a * 16 + (l & 15)];
#pragma unroll
            for (int a = 0; a < MT; a++) {
                const i32x4 A = Aw[a % AW];
                if (a + AW < MT) Aw[a % AW] = atile[buf][(l >> 4) * Mp + (a + AW) * 16 + (l & 15)];
                acc[a][n][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, Bl, acc[a][n][0], 0, 0, 0);
                acc[a][n][PL - 1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, Bh, acc[a][n][PL - 1], 0, 0, 0);
            }
            // keep the digit-tile reads AW tiles ahead of their MFMAs
            __builtin_amdgcn_sched_group_barrier(0x100, AW, 0);
#pragma unroll
            for (int a = 0; a + AW < MT; a++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
            }
            __builtin_amdgcn_sched_group_barrier(0x008, 2 * AW, 0);
        }
    };
    // Step s: issue the loads of step s+L (B slot (s+L)%DEPTH held step s-1's
    // operands, digit slot (s+L)%L held step s's digits, stored last step),
    // compute from B slot s%DEPTH and LDS buffer s&1, then store step s+1's
    // digits in buffer (s+1)&1 (every wave finished step s-1, its last
    // reader, before the previous barrier). The step count is uniform over
    // the block, so every barrier is reached by all waves.
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
    if constexpr (POW) {
        const uint32_t g = l >> 4, cl = l & 15;
#pragma unroll
        for (int h = 0; h < QT; h++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t q = 16 * h + 4 * g + i;
#pragma unroll
                for (int n = 0; n < NT; n++) {
                    int32_t pk[8];
#pragma unroll
                    for (int k = 0; k < 8; k++) pk[k] = acc[QT * k + h][n][0][i];
                    uint32_t* o = out + (size_t)q * out_qstride + (size_t)slice * 2 * OC + col0 + n * 16 + cl;
                    if (q < nq) {
                        const uint32_t r0 = pow_residue(pk, slice == 0 ? S[2 * q] : 0, 0);
                        const uint32_t r1 = pow_residue(pk, slice == 0 ? S[2 * q + 1] : 0, 1);
                        o[0] = r0;
                        o[OC] = r1;
                    }
                }
            }
    } else if constexpr (CRT7) {
        // after the last barrier no wave reads the digit ring again
        int32_t* ep = reinterpret_cast<int32_t*>(smem) + w * EPW;
        const uint32_t g = l >> 4, cl = l & 15;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            // tiles 7h..7h+6 = rows 112h.. = queries 16h..16h+15; lane (g, cl)
            // holds rows 16a + 4g + i of column cl
#pragma unroll
            for (int a = 0; a < 7; a++)
#pragma unroll
                for (int pl = 0; pl < 2; pl++)
                    *reinterpret_cast<i32x4*>(ep + (pl * 16 + cl) * EPS + 16 * a + 4 * g) = acc[7 * h + a][0][pl];
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t ql = g + 4 * j, q = 16 * h + ql;
                int64_t t[7];
#pragma unroll
                for (int k = 0; k < 7; k++)
                    t[k] = (int64_t)ep[(16 + cl) * EPS + 7 * ql + k] * 256 + ep[cl * EPS + 7 * ql + k];
                if (q < nq) {
                    uint32_t* o = out + (size_t)q * out_qstride + (size_t)slice * 2 * OC + col0 + cl;
#pragma unroll
                    for (int c = 0; c < 2; c++) o[(size_t)c * OC] = crt7_residue(t, slice == 0 ? S[2 * q + c] : 0, c);
                }
            }
            __syncthreads();  // the second half reuses the region
        }
    } else {
#pragma unroll
        for (int a = 0; a < MT; a++)
#pragma unroll
            for (int n = 0; n < NT; n++)
                store_pair(acc[a][n][0], acc[a][n][1], (uint32_t)a * 4 + (l >> 4), nq, slice == 0, S, out,
                           out_qstride, slice, OC, col0 + n * 16 + (l & 15));
    }
    if (ps.group_done) pass_handover(ps, xt);
}

// 17..32 queries with the DB tile shared instead: each of the block's 4
// waves owns MT/4 digit tiles (8 queries) for all of the block's NT*16
// columns, reads its digit operands straight from memory (they are
// L2-resident: every block reads the same R*Mp bytes), and takes its DB
// operands from a per-step tile the 4 waves load once into an LDS double
// buffer (flipped to signed bytes on the way in). Per MFMA this moves half
// the LDS bytes of the shared-digit form, whose every wave reads the whole
// Mp-row digit tile for only 16 columns. Loads follow the same one-lead
// rule: the digits of step s+L go to a register ring of DEPTH slots, the
// thread's share of the DB tile of step s+L to a ring of L slots (stored to
// LDS one step ahead).
template <int MT, int NT, int DEPTH>
__global__ __launch_bounds__(256, 2) void k_dbpass_mfma_bs(const uint8_t* __restrict__ db,
                                                           const int8_t* __restrict__ dig,
                                                           const uint32_t* __restrict__ S,
                                                           uint32_t* __restrict__ out, size_t out_qstride,
                                                           uint32_t C, uint32_t OC, uint32_t nq, uint32_t kbps,
                                                           PassStream) {  // never streamed
    static_assert(DEPTH >= 3 && MT % 4 == 0, "shape");
    constexpr int L = DEPTH - 1;
    constexpr int P = DEPTH * L;
    constexpr int MW = MT / 4;                  // digit tiles per wave
    constexpr uint32_t Mp = 16 * MT;
    constexpr uint32_t TC = 16 * NT;            // columns per block
    constexpr uint32_t BU = 4 * 2 * TC / 256;   // 16-byte units of the step's DB tile per thread
    __shared__ i32x4 btile[2][4][2][TC];        // [buffer][block of the step][plane][column]
    const uint32_t tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const uint32_t slice = blockIdx.y;
    const uint32_t bcol0 = blockIdx.x * TC;
    const uint32_t kb0 = slice * kbps;
    const uint32_t nks = kbps / 4;
    // this thread's DB units: unit u = tid + 256*i -> (block u/(2*TC), plane (u/TC)%2, column u%TC)
    const i32x4* bsrc = reinterpret_cast<const i32x4*>(db) + (size_t)kb0 * 2 * C + bcol0;
    // digit operand of lane l for tile a: row 16*(MW*w + a) + l%16 of block 4*ks + l/16
    const i32x4* asrc = reinterpret_cast<const i32x4*>(dig) + ((size_t)kb0 + (l >> 4)) * Mp + 16 * MW * w + (l & 15);
    const i32x4 flip = {(int32_t)0x80808080, (int32_t)0x80808080, (int32_t)0x80808080, (int32_t)0x80808080};
    i32x4 acc[MW][NT][2];
#pragma unroll
    for (int a = 0; a < MW; a++)
#pragma unroll
        for (int n = 0; n < NT; n++) acc[a][n][0] = acc[a][n][1] = i32x4{0, 0, 0, 0};
    i32x4 av[DEPTH][MW];
    i32x4 br[L][BU];
    auto issue = [&](int as, int bs, uint32_t step) {
        const uint32_t k = step < nks ? step : nks - 1;
#pragma unroll
        for (int a = 0; a < MW; a++) av[as][a] = asrc[(size_t)k * 4 * Mp + 16 * a];
#pragma unroll
        for (int i = 0; i < (int)BU; i++) {
            const uint32_t u = tid + 256 * i, kb = u / (2 * TC), pl = (u / TC) % 2, c = u % TC;
            br[bs][i] = __builtin_nontemporal_load(bsrc + ((size_t)(4 * k + kb) * 2 + pl) * C + c);
        }
    };
    auto store_b = [&](int buf, int bs) {
#pragma unroll
        for (int i = 0; i < (int)BU; i++) {
            const uint32_t u = tid + 256 * i, kb = u / (2 * TC), pl = (u / TC) % 2, c = u % TC;
            btile[buf][kb][pl][c] = br[bs][i] ^ flip;
        }
    };
    auto compute = [&](int as, int buf) {
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int pl = 0; pl < 2; pl++) {
                const i32x4 B = btile[buf][l >> 4][pl][n * 16 + (l & 15)];
#pragma unroll
                for (int a = 0; a < MW; a++)
                    acc[a][n][pl] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[as][a], B, acc[a][n][pl], 0, 0, 0);
            }
    };
    // Step s: issue step s+L's loads (digit slot (s+L)%DEPTH held step s-1's
    // operands; DB slot (s+L)%L held step s's share, stored last step),
    // compute from digit slot s%DEPTH and LDS buffer s&1, then store step
    // s+1's share in buffer (s+1)&1 (its last reader, step s-1, finished
    // before the previous barrier). The step count is uniform over the
    // block, so every barrier is reached by all waves.
#pragma unroll
    for (int j = 0; j < L; j++) issue(j, j, (uint32_t)j);
    store_b(0, 0);
    __syncthreads();
    for (uint32_t ks = 0; ks < nks; ks += P) {
#pragma unroll
        for (int j = 0; j < P; j++) {
            const uint32_t step = ks + (uint32_t)j;
            if (step >= nks) break;
            issue((j + L) % DEPTH, (j + L) % L, step + L);
            compute(j % DEPTH, (int)(step & 1));
            store_b((int)((step + 1) & 1), (j + 1) % L);
            __syncthreads();
        }
    }
#pragma unroll
    for (int a = 0; a < MW; a++)
#pragma unroll
        for (int n = 0; n < NT; n++)
            store_pair(acc[a][n][0], acc[a][n][1], (uint32_t)(MW * w + a) * 4 + (l >> 4), nq, slice == 0, S, out,
                       out_qstride, slice, OC, bcol0 + n * 16 + (l & 15));
}

// Largest power-of-two slice count s with blocks*s <= target that keeps
// every slice a whole number of `unit`-block chunks.
uint32_t pick_slices(size_t nkb, size_t blocks, size_t target, size_t unit) {
    uint32_t s = 1;
    while (blocks * s * 2 <= target && nkb % (s * 2 * unit) == 0 && nkb / (s * 2) >= unit) s *= 2;
    return s;
}

// 16-column tiles per wave for MT digit tiles (accumulators MT*NT*8 registers).
uint32_t mfma_nt(uint32_t mt) { return mt <= 4 ? 4 : mt == 8 ? 2 : 1; }

}  // namespace

hipError_t launch_to_planes(const uint16_t* src, uint8_t* dst, size_t C, size_t kb0, size_t nkb, hipStream_t s) {
    if (C % 256) return hipErrorInvalidValue;
    if (nkb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_to_planes, dim3((unsigned)(C / 256), (unsigned)nkb), dim3(256), 0, s, src, dst, (uint32_t)C,
                       (uint32_t)kb0);
    return hipGetLastError();
}

DbPlan dbplan_f64(size_t R, size_t ncols, int unroll, size_t target_blocks) {
    DbPlan p;
    p.kind = DbPlan::F64;
    p.param = (uint32_t)unroll;
    p.slices = pick_slices(R / 16, ncols / 256, target_blocks, (size_t)(unroll > 4 ? unroll : 4));
    return p;
}

DbPlan dbplan_mfma(size_t R, size_t ncols, size_t nq, size_t target_blocks, uint32_t depth, bool dbtile,
                   uint32_t waves, bool crt, bool pow) {
    DbPlan p;
    p.kind = DbPlan::MFMA;
    p.param = nq <= 2 ? 1 : nq <= 4 ? 2 : nq <= 8 ? 4 : nq <= 16 ? 8 : 16;  // 2 queries per 16-row digit tile
    p.dbtile = dbtile && p.param == 16;
    // 7 rows per query: 32 queries in 14 tiles (16 in 7 with power rows)
    p.crt = !p.dbtile && ((crt && p.param == 16) || (pow && p.param >= 8));
    p.pow = pow && p.crt;
    if (p.crt) p.param = p.param == 16 ? 14 : 7;
    p.waves = p.param >= 8 && !p.dbtile && waves == 8 ? 8 : 4;
    if (p.pow) p.waves = 8;
    // ring depths the kernels are built with (what fits 256 VGPRs): 14 or 16
    // tiles with shared digits 3..7 (8-wave blocks) or 3..5, power rows
    // (two column tiles) 3..5, otherwise 3..4
    const uint32_t maxd = p.pow ? 5 : p.param >= 14 && !p.dbtile ? (p.waves == 8 ? 7 : 5) : 4;
    p.depth = p.param >= 7 ? (depth < 3 ? 3 : depth > maxd ? maxd : depth) : 3;
    const size_t cols = dbplan_cols(p);
    // slices: whole number of 128-row (8-block) chunks so the step count is even
    p.slices = pick_slices(R / 16, ncols / cols, target_blocks, 8);
    // at most 2^16 rows per slice keeps |H_k|, |L_k| <= 2^30 (i32-exact) and
    // the recombined sum below 2^58.1; the power rows' sums of two products
    // need 2^15
    while (R / p.slices > (p.pow ? 32768u : 65536u)) p.slices *= 2;
    return p;
}

DbPlan dbplan_form(size_t R, size_t ncols, size_t nq, size_t target_blocks, uint32_t depth, int form) {
    return dbplan_mfma(R, ncols, nq, target_blocks, depth, form == 1, form == 2 || form == 3 || form == 5 ? 8 : 4,
                       form == 3 || form == 4, form == 5);
}

size_t dbplan_mp(const DbPlan& p) { return p.kind == DbPlan::MFMA ? 16 * p.param : 0; }

size_t dbplan_max_queries(const DbPlan& p) {
    return p.kind != DbPlan::MFMA ? 1 : p.crt ? dbplan_mp(p) / 7 : dbplan_mp(p) / 8;
}

// 16-column tiles per wave of an MFMA plan
static uint32_t plan_nt(const DbPlan& p) { return p.pow ? 2 : p.nt ? p.nt : mfma_nt(p.param); }

size_t dbplan_cols(const DbPlan& p) {
    return p.kind == DbPlan::MFMA ? 16 * plan_nt(p) * p.waves : 256;  // power rows: 2 tiles
}

std::string dbplan_kernel(const DbPlan& p) {
    if (p.kind == DbPlan::F64) return "k_dbpass_f64<" + std::to_string(p.param) + ">";
    const std::string mt_nt = std::to_string(p.param) + ", " + std::to_string(plan_nt(p));
    if (p.param <= 4) return "k_dbpass_mfma<" + mt_nt + ">";  // 1, 2 or 4 tiles: the direct kernel
    if (p.dbtile) return "k_dbpass_mfma_bs<16, 4, " + std::to_string(p.depth) + ">";
    // every template argument, defaults included (as profilers demangle it)
    return "k_dbpass_mfma_sa<" + (p.pow ? std::to_string(p.param) + ", 2" : mt_nt) + ", 2, " + std::to_string(p.depth) + ", " +
           std::to_string(p.waves) + ", " + (p.pow ? "1>" : "0>");
}

hipError_t launch_dbpass_f64(const uint8_t* db, const double* bq, uint32_t* out, size_t ostride, size_t R, size_t C,
                             size_t col0, size_t ncols, const DbPlan& p, hipStream_t s) {
    if (p.kind != DbPlan::F64 || ncols % 256 || col0 % 256 || col0 + ncols > C || ostride < ncols)
        return hipErrorInvalidValue;
    const uint32_t kbps = (uint32_t)(R / 16 / p.slices);
    const dim3 grid((unsigned)(ncols / 256), p.slices);
    const uint8_t* d = db + col0 * 16;
    const uint32_t Cc = (uint32_t)C, OC = (uint32_t)ostride;
#define F64(U) hipLaunchKernelGGL(k_dbpass_f64<U>, grid, dim3(256), 0, s, d, bq, out, (uint32_t)R, Cc, OC, kbps)
    switch (p.param) {
        case 2: F64(2); break;
        case 8: F64(8); break;
        default: F64(4); break;
    }
#undef F64
    return hipGetLastError();
}

hipError_t launch_mfma_digits(const double* bq, size_t nq, size_t R, const DbPlan& p, int8_t* dig, uint32_t* S,
                              hipStream_t s) {
    if (p.kind != DbPlan::MFMA || nq == 0 || nq > kMfmaMaxQueries || R % 2048) return hipErrorInvalidValue;
    const uint32_t Mp = (uint32_t)dbplan_mp(p);
    if (nq > dbplan_max_queries(p)) return hipErrorInvalidValue;
    const size_t nthr = (R / 16) * (p.crt ? Mp / 7 : Mp / 4);
    const dim3 grid((unsigned)((nthr + 255) / 256));
    if (p.crt)
        hipLaunchKernelGGL(k_digits7, grid, dim3(256), 0, s, bq, (uint32_t)R, (uint32_t)nq, Mp, dig,
                           (uint32_t)p.pow);
    else
        hipLaunchKernelGGL(k_digits, grid, dim3(256), 0, s, bq, (uint32_t)R, (uint32_t)nq, Mp, dig);
    hipLaunchKernelGGL(k_bsum, dim3((unsigned)(2 * nq)), dim3(1024), 0, s, bq, (uint32_t)R, S);
    return hipGetLastError();
}

bool dbplan_streams(const DbPlan& p, size_t C) {
    const size_t cols = dbplan_cols(p);
    return p.kind == DbPlan::MFMA && !p.dbtile && p.slices > 0 && (p.slices & (p.slices - 1)) == 0 &&
           N % cols == 0 && C % N == 0;
}

hipError_t launch_dbpass_mfma(const uint8_t* db, const int8_t* dig, const uint32_t* S, size_t nq, uint32_t* out,
                              size_t out_qstride, size_t ostride, size_t R, size_t C, size_t col0, size_t ncols,
                              const DbPlan& p, hipStream_t s, const PassStream* ps) {
    if (p.kind != DbPlan::MFMA || nq == 0 || nq > kMfmaMaxQueries) return hipErrorInvalidValue;
    const uint32_t Mp = (uint32_t)dbplan_mp(p);
    const uint32_t kbps = (uint32_t)(R / 16) / p.slices;
    const size_t cols = dbplan_cols(p);
    if (nq > dbplan_max_queries(p) || (p.crt && p.param != 14 && !(p.pow && p.param == 7)) || kbps % 8 ||
        (p.nt && !(p.nt == 8 && p.param == 1)) ||
        ncols % cols || col0 % cols ||
        col0 + ncols > C || ostride < ncols)
        return hipErrorInvalidValue;
    PassStream st{};
    dim3 grid((unsigned)(ncols / cols), p.slices);
    if (ps) {
        // the whole width, one 1-D grid (slices of a tile in a row, the highest tiles first)
        st = *ps;
        uint32_t sh = 0;
        while ((1u << sh) < p.slices) sh++;
        st.slice_shift = sh;
        st.tiles = (uint32_t)(ncols / cols);
        st.tiles_per_group = (uint32_t)(N / cols);
        st.group_base = (uint32_t)(col0 / N);
        if (!dbplan_streams(p, C) || col0 % N || ncols % N || ncols == 0 || !st.group_done || ostride != C ||
            (size_t)st.tiles * p.slices > 0x7fffffffu)
            return hipErrorInvalidValue;
        grid = dim3(st.tiles * p.slices, 1);
    }
    const uint32_t Cc = (uint32_t)C, OC = (uint32_t)ostride, Q = (uint32_t)nq;
    const uint8_t* d = db + col0 * 16;
#define MFMA(K, ...) \
    hipLaunchKernelGGL((K<__VA_ARGS__>), grid, dim3(256), 0, s, d, dig, S, out, out_qstride, Cc, OC, Q, kbps, st)
#define MFMA8(K, ...) \
    hipLaunchKernelGGL((K<__VA_ARGS__, 8>), grid, dim3(512), 0, s, d, dig, S, out, out_qstride, Cc, OC, Q, kbps, st)
#define MFMA8P(MT, D)                                                                                          \
    hipLaunchKernelGGL((k_dbpass_mfma_sa<MT, 2, 2, D, 8, 1>), grid, dim3(512), 0, s, d, dig, S, out, out_qstride, \
                       Cc, OC, Q, kbps, st)
    switch (p.param) {
        case 1:
            if (p.nt == 8)
                MFMA(k_dbpass_mfma, 1, 8);
            else
                MFMA(k_dbpass_mfma, 1, 4);
            break;
        case 2: MFMA(k_dbpass_mfma, 2, 4); break;
        case 4: MFMA(k_dbpass_mfma, 4, 4); break;
        case 8:
            switch (p.depth * 10 + p.waves) {
                case 34: MFMA(k_dbpass_mfma_sa, 8, 2, 2, 3); break;
                case 44: MFMA(k_dbpass_mfma_sa, 8, 2, 2, 4); break;
                case 38: MFMA8(k_dbpass_mfma_sa, 8, 2, 2, 3); break;
                case 48: MFMA8(k_dbpass_mfma_sa, 8, 2, 2, 4); break;
                default: return hipErrorInvalidValue;
            }
            break;
        case 14:
            if (p.pow) {
                switch (p.depth) {
                    case 3: MFMA8P(14, 3); break;
                    case 4: MFMA8P(14, 4); break;
                    case 5: MFMA8P(14, 5); break;
                    default: return hipErrorInvalidValue;
                }
                break;
            }
            switch (p.depth * 10 + p.waves) {
                case 34: MFMA(k_dbpass_mfma_sa, 14, 1, 2, 3); break;
                case 44: MFMA(k_dbpass_mfma_sa, 14, 1, 2, 4); break;
                case 54: MFMA(k_dbpass_mfma_sa, 14, 1, 2, 5); break;
                case 38: MFMA8(k_dbpass_mfma_sa, 14, 1, 2, 3); break;
                case 48: MFMA8(k_dbpass_mfma_sa, 14, 1, 2, 4); break;
                case 58: MFMA8(k_dbpass_mfma_sa, 14, 1, 2, 5); break;
                case 68: MFMA8(k_dbpass_mfma_sa, 14, 1, 2, 6); break;
                case 78: MFMA8(k_dbpass_mfma_sa, 14, 1, 2, 7); break;
                default: return hipErrorInvalidValue;
            }
            break;
        case 7:
            if (!p.pow) return hipErrorInvalidValue;
            switch (p.depth) {
                case 3: MFMA8P(7, 3); break;
                case 4: MFMA8P(7, 4); break;
                case 5: MFMA8P(7, 5); break;
                default: return hipErrorInvalidValue;
            }
            break;
        case 16:
            if (p.dbtile) {
                switch (p.depth) {
                    case 3: MFMA(k_dbpass_mfma_bs, 16, 4, 3); break;
                    case 4: MFMA(k_dbpass_mfma_bs, 16, 4, 4); break;
                    default: return hipErrorInvalidValue;
                }
                break;
            }
            switch (p.depth * 10 + p.waves) {
                case 34: MFMA(k_dbpass_mfma_sa, 16, 1, 2, 3); break;
                case 44: MFMA(k_dbpass_mfma_sa, 16, 1, 2, 4); break;
                case 54: MFMA(k_dbpass_mfma_sa, 16, 1, 2, 5); break;
                case 38: MFMA8(k_dbpass_mfma_sa, 16, 1, 2, 3); break;
                case 48: MFMA8(k_dbpass_mfma_sa, 16, 1, 2, 4); break;
                case 58: MFMA8(k_dbpass_mfma_sa, 16, 1, 2, 5); break;
                case 68: MFMA8(k_dbpass_mfma_sa, 16, 1, 2, 6); break;
                case 78: MFMA8(k_dbpass_mfma_sa, 16, 1, 2, 7); break;
                default: return hipErrorInvalidValue;
            }
            break;
        default: return hipErrorInvalidValue;
    }
#undef MFMA
#undef MFMA8
#undef MFMA8P
    return hipGetLastError();
}

}  // namespace ipx
