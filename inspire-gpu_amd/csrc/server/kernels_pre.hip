// One-time preprocessing kernels for gfx950: in-place encode (inverse DFT
// over Z_P[X]/(X^N+1)), the a-side matmul, the ring embedding and the
// CollapseA chains that produce the per-group precomp tensors.
//
// Semantics follow the reference preprocess (src/gpu_protocol.cu:535-817):
// encode (src/gpu_preprocess.cu:931-1034), a32 = N^-1 * reversed
// sum_r NTT(db column slice) * NTT(tau_{2N-1}(a_r))
// (src/gpu_preprocess.cu:347-410), ring embed + automorphism
// (src/gpu_preprocess.cu:39-92, :158-194), collapse (src/gpu_preprocess.cu:
// 452-881). Two re-formulations replace the reference's kernels while
// producing the same values:
//  * ring embed: the 2^34-MAC sum per group is the evaluation of
//    P_{c,p}(Y) = sum_j a32[j][c][p] Y^j at the odd powers of psi, i.e. a
//    length-N negacyclic NTT along j; combined with the automorphism
//    permutation, a_post[k][c][i] = F[c][perm_{+-5^k}(i)][i] with
//    F[c][p] = NTT(a32[.][c][p]).
//  * encode: the transform factors by coefficient residue k mod (2N/D) into
//    independent D x D/2 negacyclic blocks, each done in LDS in one pass
//    over the DB instead of log2(D)+2 global passes.
#include "kernels.h"

namespace ipx {

namespace {

__device__ __forceinline__ uint32_t modp(uint32_t x) {  // x mod 65535, any u32
    uint32_t r = (x & 0xffffu) + (x >> 16);
    r = (r & 0xffffu) + (r >> 16);
    return r >= PMOD ? r - PMOD : r;
}

// ------------------------------------------------------------ encode
// Block (row, ct, k0-chunk). Sub-problem for residue k0 of the coefficient
// index mod S = 2N/Da: y[d][k1] = x[d][k0 + S*k1], k1 < Da/2, on which
// X^S acts as a negacyclic shift. out[m] = Da^-1 sum_d y[d] * X^(-2N d m/Da)
// computed as a radix-2 DIT FFT (bit-reversed load, natural-order output).
__global__ __launch_bounds__(1024) void k_encode(uint16_t* __restrict__ db, uint32_t C, uint32_t Da, uint32_t logDa,
                                                 uint32_t KC, uint32_t nchunk, uint32_t Dinv) {
    __shared__ uint16_t y[32768];
    const uint32_t row = blockIdx.x, ct = blockIdx.y / nchunk, chunk = blockIdx.y % nchunk;
    const uint32_t S = 2 * N / Da, L1 = Da / 2, E = KC * Da * L1;
    uint16_t* base = db + (size_t)row * C + (size_t)ct * Da * N + (size_t)chunk * KC;
    for (uint32_t x = threadIdx.x; x < E; x += 1024) {
        const uint32_t k0 = x % KC, rest = x / KC, k1 = rest % L1, d = rest / L1;
        const uint32_t dr = __builtin_bitreverse32(d) >> (32 - logDa);
        y[(k0 * Da + dr) * L1 + k1] = base[(size_t)d * N + (size_t)S * k1 + k0];
    }
    __syncthreads();
    const uint32_t npairs = E / 2;
    for (uint32_t lv = 0; lv < logDa; lv++) {
        const uint32_t half = 1u << lv, size = half << 1;
        uint16_t sum[16], dif[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const uint32_t z = threadIdx.x + 1024 * r;
            if (z < npairs) {
                const uint32_t k1 = z % L1, rest = z / L1, k0 = rest % KC, pidx = rest / KC;
                const uint32_t de = (pidx / half) * size + pidx % half, dodd = de + half;
                const uint32_t sh = (Da / size) * (pidx % half);
                const uint32_t e = y[(k0 * Da + de) * L1 + k1];
                const uint32_t src = k1 + sh;
                uint32_t o = src < L1 ? y[(k0 * Da + dodd) * L1 + src] : y[(k0 * Da + dodd) * L1 + src - L1];
                if (src >= L1 && o) o = PMOD - o;
                uint32_t a = e + o;
                if (a >= PMOD) a -= PMOD;
                uint32_t b = e + PMOD - o;
                if (b >= PMOD) b -= PMOD;
                sum[r] = (uint16_t)a;
                dif[r] = (uint16_t)b;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const uint32_t z = threadIdx.x + 1024 * r;
            if (z < npairs) {
                const uint32_t k1 = z % L1, rest = z / L1, k0 = rest % KC, pidx = rest / KC;
                const uint32_t de = (pidx / half) * size + pidx % half, dodd = de + half;
                y[(k0 * Da + de) * L1 + k1] = sum[r];
                y[(k0 * Da + dodd) * L1 + k1] = dif[r];
            }
        }
        __syncthreads();
    }
    for (uint32_t x = threadIdx.x; x < E; x += 1024) {
        const uint32_t k0 = x % KC, rest = x / KC, k1 = rest % L1, d = rest / L1;
        base[(size_t)d * N + (size_t)S * k1 + k0] = (uint16_t)modp((uint32_t)y[(k0 * Da + d) * L1 + k1] * Dinv);
    }
}

// ------------------------------------------------------------ a-side
// Block (group, 4 columns) over the resident DB (byte planes in 16-row
// blocks): acc[col][c] = sum_r NTT(db[rN..rN+N-1][col]) * crs[r][c]; a32T[g][c][k][col] = N^-1 * acc[col][c][N-1-k].
// Polynomial p = col*2 + c; team tm transforms p = tm and p = tm + 4.
__global__ __launch_bounds__(1024) void k_aside(const uint8_t* __restrict__ db, uint32_t C, uint32_t nr, uint32_t g0,
                                                const uint32_t* __restrict__ crs, uint32_t* __restrict__ a32T,
                                                NttConst nc) {
    __shared__ uint32_t buf[8][POLY_LDS];
    __shared__ uint32_t twl[8 * N];
    const uint32_t gl = blockIdx.x / (N / 4), j0 = (blockIdx.x % (N / 4)) * 4;
    const uint32_t tid = threadIdx.x, team = tid >> 8, t = tid & 255;
    stage_twiddles(twl, nc, tid, 1024);
    const int c = (int)(team & 1);  // p = team + 4*rr has limb team&1
    const Tw tw = tw_lds(twl, c);
    const size_t col0 = (size_t)(g0 + gl) * N + j0;
    uint32_t acc[2][8];
#pragma unroll
    for (int rr = 0; rr < 2; rr++)
#pragma unroll
        for (int k = 0; k < 8; k++) acc[rr][k] = 0;
    for (uint32_t r = 0; r < nr; r++) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t i = tid + 1024 * h;
            const size_t row = (size_t)r * N + i;  // plane p of (row, col) at ((row/16)*2 + p)*C*16 + col*16 + row%16
            const uint8_t* src = db + ((row >> 4) * 2 * C + col0) * 16 + (row & 15);
#pragma unroll
            for (int col = 0; col < 4; col++) {
                const uint32_t x = (uint32_t)src[col * 16] | ((uint32_t)src[(size_t)C * 16 + col * 16] << 8);
                buf[col * 2][ls(i)] = x;
                buf[col * 2 + 1][ls(i)] = x;
            }
        }
        __syncthreads();
        ntt_fwd_lds(buf[team], t, tw, c);
        ntt_fwd_lds(buf[team + 4], t, tw, c);
#pragma unroll
        for (int rr = 0; rr < 2; rr++) {
            const uint32_t p = team + 4 * rr;
            const uint32_t* A = crs + ((size_t)r * 2 + c) * N;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const uint32_t e = t + 256 * k;
                acc[rr][k] = add_mod(acc[rr][k], mul_mod(buf[p][ls(e)], A[e], c), qof(c));
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int rr = 0; rr < 2; rr++) {
        const uint32_t p = team + 4 * rr;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t e = t + 256 * k;
            buf[p][ls(e)] = mul_shoup(acc[rr][k], nc.inv_n[c], nc.inv_n_p[c], qof(c));
        }
    }
    __syncthreads();
    for (uint32_t x = tid; x < 2 * N; x += 1024) {
        const uint32_t cc = x / N, k = x % N, e = N - 1 - k;
        uint4 o;
        o.x = buf[0 * 2 + cc][ls(e)];
        o.y = buf[1 * 2 + cc][ls(e)];
        o.z = buf[2 * 2 + cc][ls(e)];
        o.w = buf[3 * 2 + cc][ls(e)];
        *reinterpret_cast<uint4*>(a32T + (((size_t)gl * 2 + cc) * N + k) * N + j0) = o;
    }
}

// ------------------------------------------------------------ ring embed
// In-place forward NTT of a32T rows; row x = (g*2 + c)*N + p has limb c.
__global__ __launch_bounds__(256) void k_ntt_rows(uint32_t* __restrict__ rows, NttConst nc) {
    __shared__ uint32_t buf[POLY_LDS];
    const int limb = (int)((blockIdx.x >> LOGN) & 1);
    uint32_t* p = rows + (size_t)blockIdx.x * N;
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 8; k++) buf[ls(t + 256 * k)] = p[t + 256 * k];
    __syncthreads();
    ntt_fwd_lds(buf, t, tw_global(nc, limb), limb);
#pragma unroll
    for (int k = 0; k < 8; k++) p[t + 256 * k] = buf[ls(t + 256 * k)];
}

// ------------------------------------------------------------ collapse
// One block per (group, direction): 1023 dependent steps. Step s folds
// k = N/2-1-s: state -> INTT -> gadget digits 1,2 -> NTT -> D[s]; state =
// a_post[k-1] - sum_j D[s][j] * K[j][perm]. Both the expanded key of step s
// (tau_{+-5^(k-1)} of ksk5a, reference kskstep = n_steps-1-s) and a_post[k-1]
// are gathered through the same slot permutation p.
__global__ __launch_bounds__(1024) void k_collapse_chain(const uint32_t* __restrict__ F,
                                                         const uint32_t* __restrict__ ksk5a,
                                                         uint32_t* __restrict__ dtmp,
                                                         uint32_t* __restrict__ states, NttConst nc, CrtConst cc) {
    __shared__ uint32_t buf[4][POLY_LDS];
    __shared__ uint32_t K[4][N];  // [digit*2 + limb][N]
    __shared__ uint32_t twl[8 * N];
    const uint32_t gl = blockIdx.x, dir = blockIdx.y, tid = threadIdx.x, team = tid >> 8, t = tid & 255;
    for (uint32_t x = tid; x < 4 * N; x += 1024) K[x / N][x % N] = ksk5a[x];
    stage_twiddles(twl, nc, tid, 1024);
    const int tlimb = (int)(team & 1);
    const Tw tw = tw_lds(twl, tlimb);
    const uint32_t inv_n = nc.inv_n[tlimb], inv_n_p = nc.inv_n_p[tlimb];
    const uint32_t* Fg = F + (size_t)gl * 2 * N * N;
    uint32_t* Dout = dtmp + ((size_t)gl * 2 + dir) * RD_SZ;
    uint32_t v[2], st[2][2];  // [e], [limb][e]
#pragma unroll
    for (int e = 0; e < 2; e++) v[e] = 2 * brev11(tid + 1024 * e) + 1;
    // t = 5^(N/2-1) mod 2N, then 5^(N/2-2-s) for step s
    uint32_t tk = 1;
    for (uint32_t ex = HALF - 1, b = 5; ex; ex >>= 1, b = (b * b) & (2 * N - 1))
        if (ex & 1) tk = (tk * b) & (2 * N - 1);
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const uint32_t pf = perm_slot(tk, v[e]);
        const uint32_t p = dir ? N - 1 - pf : pf;
        const uint32_t i = tid + 1024 * e;
        st[0][e] = Fg[(size_t)p * N + i];
        st[1][e] = Fg[((size_t)N + p) * N + i];
    }
    uint32_t ts = (tk * 3277u) & (2 * N - 1);
    for (uint32_t s = 0; s < NSTEPS; s++) {
        uint32_t pidx[2], nxt[2][2];
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t pf = perm_slot(ts, v[e]);
            pidx[e] = dir ? N - 1 - pf : pf;
            const uint32_t i = tid + 1024 * e;
            nxt[0][e] = Fg[(size_t)pidx[e] * N + i];
            nxt[1][e] = Fg[((size_t)N + pidx[e]) * N + i];
        }
#pragma unroll
        for (int e = 0; e < 2; e++) {
            buf[0][ls(tid + 1024 * e)] = st[0][e];
            buf[1][ls(tid + 1024 * e)] = st[1][e];
        }
        __syncthreads();
        ntt_inv_lds(buf[team], t, tw, tlimb, inv_n, inv_n_p, team < 2);
        uint32_t d0[2][2], d1[2][2];
#pragma unroll
        for (int e = 0; e < 2; e++) gadget2(buf[0][ls(tid + 1024 * e)], buf[1][ls(tid + 1024 * e)], cc, d0[e], d1[e]);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 2; e++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                buf[j * 2 + 0][ls(tid + 1024 * e)] = d0[e][j];
                buf[j * 2 + 1][ls(tid + 1024 * e)] = d1[e][j];
            }
        __syncthreads();
        ntt_fwd_lds(buf[team], t, tw, tlimb);
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const uint32_t i = tid + 1024 * e;
                uint64_t acc = 0;
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const uint32_t dv = buf[j * 2 + c][ls(i)];
                    Dout[(((size_t)c * NSTEPS + s) * D_EFF + j) * N + i] = dv;
                    acc += (uint64_t)dv * K[j * 2 + c][pidx[e]];
                }
                st[c][e] = sub_mod(nxt[c][e], red64(acc, c), qof(c));
            }
        ts = (ts * 3277u) & (2 * N - 1);
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < 2; e++)
            states[(((size_t)gl * 2 + dir) * 2 + c) * N + tid + 1024 * e] = st[c][e];
}

// Final CollapseOne: digits of the conjugate chain's last state -> D_final;
// a = (forward chain's last state) - sum_j D_final[j] * NTT(ksk_neg1 a-part)[j].
__global__ __launch_bounds__(1024) void k_collapse_final(const uint32_t* __restrict__ states,
                                                         const uint32_t* __restrict__ kneg1a,
                                                         uint32_t* __restrict__ precomp_g0, NttConst nc,
                                                         CrtConst cc) {
    __shared__ uint32_t buf[4][POLY_LDS];
    __shared__ uint32_t twl[8 * N];
    const uint32_t gl = blockIdx.x, tid = threadIdx.x, team = tid >> 8, t = tid & 255;
    stage_twiddles(twl, nc, tid, 1024);
    const int tlimb = (int)(team & 1);
    const Tw tw = tw_lds(twl, tlimb);
    const uint32_t* sf = states + (size_t)gl * 4 * N;  // dir 0
    const uint32_t* sc = sf + 2 * N;                   // dir 1
    uint32_t* pc = precomp_g0 + (size_t)gl * PC_GROUP;
#pragma unroll
    for (int e = 0; e < 2; e++) {
        buf[0][ls(tid + 1024 * e)] = sc[tid + 1024 * e];
        buf[1][ls(tid + 1024 * e)] = sc[N + tid + 1024 * e];
    }
    __syncthreads();
    ntt_inv_lds(buf[team], t, tw, tlimb, nc.inv_n[tlimb], nc.inv_n_p[tlimb], team < 2);
    uint32_t d0[2][2], d1[2][2];
#pragma unroll
    for (int e = 0; e < 2; e++) gadget2(buf[0][ls(tid + 1024 * e)], buf[1][ls(tid + 1024 * e)], cc, d0[e], d1[e]);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 2; e++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            buf[j * 2 + 0][ls(tid + 1024 * e)] = d0[e][j];
            buf[j * 2 + 1][ls(tid + 1024 * e)] = d1[e][j];
        }
    __syncthreads();
    ntt_fwd_lds(buf[team], t, tw, tlimb);
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t i = tid + 1024 * e;
            uint64_t acc = 0;
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const uint32_t dv = buf[j * 2 + c][ls(i)];
                pc[PC_DF + ((size_t)c * D_EFF + j) * N + i] = dv;
                acc += (uint64_t)dv * kneg1a[((size_t)j * 2 + c) * N + i];
            }
            pc[PC_A + (size_t)c * N + i] = sub_mod(sf[(size_t)c * N + i], red64(acc, c), qof(c));
        }
}

}  // namespace

hipError_t launch_encode(uint16_t* db, size_t R, size_t C, size_t Da, hipStream_t s) {
    if (Da <= 1) return hipSuccess;
    uint32_t logDa = 0;
    while ((1u << logDa) < Da) logDa++;
    const uint32_t S = (uint32_t)(2 * N / Da), L1 = (uint32_t)(Da / 2);
    uint32_t KC = 32768u / ((uint32_t)Da * L1);
    if (KC > S) KC = S;
    const uint32_t nchunk = S / KC;
    uint32_t Dinv = 0;  // Da^-1 mod 65535 (Da a power of two, odd modulus)
    for (uint32_t x = 1; x < PMOD; x++)
        if ((uint64_t)x * Da % PMOD == 1) { Dinv = x; break; }
    const size_t T = C / (Da * N);
    hipLaunchKernelGGL(k_encode, dim3((unsigned)R, (unsigned)(T * nchunk)), dim3(1024), 0, s, db, (uint32_t)C,
                       (uint32_t)Da, logDa, KC, nchunk, Dinv);
    return hipGetLastError();
}

hipError_t launch_aside(const uint8_t* db, size_t R, size_t C, size_t g0, size_t ng, const uint32_t* crs,
                        uint32_t* a32T, const NttConst& nc, hipStream_t s) {
    hipLaunchKernelGGL(k_aside, dim3((unsigned)(ng * (N / 4))), dim3(1024), 0, s, db, (uint32_t)C,
                       (uint32_t)(R / N), (uint32_t)g0, crs, a32T, nc);
    return hipGetLastError();
}

hipError_t launch_ring_embed(uint32_t* a32T, size_t ng, const NttConst& nc, hipStream_t s) {
    hipLaunchKernelGGL(k_ntt_rows, dim3((unsigned)(ng * 2 * N)), dim3(256), 0, s, a32T, nc);
    return hipGetLastError();
}

hipError_t launch_collapse_chain(const uint32_t* F, size_t ng, const uint32_t* ksk5a, uint32_t* dtmp,
                                 uint32_t* states, const NttConst& nc, const CrtConst& cc, hipStream_t s) {
    hipLaunchKernelGGL(k_collapse_chain, dim3((unsigned)ng, 2), dim3(1024), 0, s, F, ksk5a, dtmp, states, nc, cc);
    return hipGetLastError();
}

hipError_t launch_collapse_final(const uint32_t* states, size_t ng, const uint32_t* kneg1a, uint32_t* precomp_g0,
                                 const NttConst& nc, const CrtConst& cc, hipStream_t s) {
    hipLaunchKernelGGL(k_collapse_final, dim3((unsigned)ng), dim3(1024), 0, s, states, kneg1a, precomp_g0, nc, cc);
    return hipGetLastError();
}

}  // namespace ipx
