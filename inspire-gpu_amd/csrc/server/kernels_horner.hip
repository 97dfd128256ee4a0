// Horner chain split over two workgroups per chain, one per RNS limb.
//
// The step acc <- ext_prod(rgsw, acc) + packed[i] (reference
// src/gpu_online.cu:474-540, src/crypto.cpp:219-279) costs 4 INTTs, two
// gadget decompositions and 8 NTTs. Only the decomposition couples the
// limbs (its CRT lift needs both residues of a coefficient), so each limb's
// workgroup inverts its own two polynomials, swaps the coefficient-form
// values with its partner through global memory, decomposes every
// coefficient itself and transforms the four digit polynomials mod its own
// prime: 6 transforms per workgroup and one exchange per step instead of 12
// transforms in one workgroup.
//
// Pairing: block b serves chain (b/16)*8 + b%8, limb (b/8)%2, so partners
// are b and b^8 (consecutive workgroups are spread over the XCDs
// round-robin, which puts partners on the same XCD and L2). The exchange
// is a release store of the step number and a bounded wait for the
// partner's (acquire at agent scope, so correctness does not depend on the
// placement); a wait that exceeds the bound sets *status and ends both
// workgroups, so every wave always reaches the end (a bound of zero takes
// that path at the first exchange).
#include "kernels.h"

namespace ipx {

namespace {

constexpr size_t BLK = (size_t)D_EFF * 2 * N;

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    for (size_t x = (size_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (size_t)gridDim.x * 256) p[x] = v;
}

__global__ __launch_bounds__(1024) void k_horner_pair(const uint32_t* __restrict__ pa_base, size_t pa_stride,
                                                      uint32_t G, uint32_t Da, uint32_t T,
                                                      const uint32_t* __restrict__ packed_b, size_t pstride,
                                                      const uint32_t* __restrict__ rgsw, size_t rstride,
                                                      uint64_t* __restrict__ out, uint32_t* __restrict__ hstate,
                                                      uint32_t g_lo, uint32_t g_hi, uint32_t* __restrict__ xbuf,
                                                      uint32_t* __restrict__ flags, uint32_t* __restrict__ status,
                                                      uint32_t nchains, uint32_t spin_limit, NttConst nc,
                                                      CrtConst cc) {
    __shared__ uint32_t buf[4][POLY_LDS];
    __shared__ uint32_t twl[4 * N];
    __shared__ uint32_t rgl[4][2][N];  // this limb's RGSW rows [part][digit][coeff]
    __shared__ uint32_t abort_flag;
    const uint32_t b = blockIdx.x, chain = (b >> 4) * 8 + (b & 7);
    const int c = (int)((b >> 3) & 1);
    if (chain >= nchains) return;  // both partners of a padding chain leave
    const uint32_t ct = chain % T, qi = chain / T;
    const uint32_t start = ct * Da;
    const uint32_t L = (G - start) < Da ? (G - start) : Da;
    const uint32_t top = start + L - 1;
    // this launch covers the chain's groups in [lo, hi), highest first
    const uint32_t lo = start > g_lo ? start : g_lo, hi = start + L < g_hi ? start + L : g_hi;
    if (lo >= hi) return;  // both partners leave
    const uint32_t tid = threadIdx.x, team = tid >> 8, t = tid & 255;
    const uint32_t q = qof(c);
    for (uint32_t x = tid; x < 4 * N; x += 1024) twl[x] = nc.tw[(size_t)c * 4 * N + x];
    if (tid == 0) abort_flag = 0;
    const Tw tw{twl, twl + N, twl + 2 * N, twl + 3 * N};
    const uint32_t* pb = packed_b + qi * pstride;
    const uint32_t* rg = rgsw + qi * rstride;
    for (uint32_t x = tid; x < 8 * N; x += 1024) {
        const uint32_t part = x / (2 * N), j = (x / N) & 1, i = x % N;
        rgl[part][j][i] = rg[(size_t)part * BLK + ((size_t)j * 2 + c) * N + i];
    }
    uint32_t* hs = hstate + ((size_t)qi * T + ct) * 4 * N;  // chain state between launches, [a|b][limb][N]
    uint32_t acc_a[2], acc_b[2];
    int first;
    if (top < g_hi) {  // the chain starts here: acc = packed[top]
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t i = tid + 1024 * e;
            acc_a[e] = pa_base[(size_t)top * pa_stride + (size_t)c * N + i];
            acc_b[e] = pb[(size_t)top * 2 * N + (size_t)c * N + i];
        }
        first = (int)top - 1;
    } else {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t i = tid + 1024 * e;
            acc_a[e] = hs[(size_t)c * N + i];
            acc_b[e] = hs[2 * (size_t)N + (size_t)c * N + i];
        }
        first = (int)hi - 1;
    }
    uint32_t* xb = xbuf + (size_t)chain * 2 * 2 * 2 * N;  // [parity][limb][a|b][N]
    uint32_t* my_flag = flags + (size_t)chain * 2 + c;
    uint32_t* other_flag = flags + (size_t)chain * 2 + (1 - c);
    __syncthreads();
    uint32_t k = 0;
    for (int gs = first; gs >= (int)lo; gs--) {
        k++;
#pragma unroll
        for (int e = 0; e < 2; e++) {
            buf[0][ls(tid + 1024 * e)] = acc_a[e];
            buf[1][ls(tid + 1024 * e)] = acc_b[e];
        }
        __syncthreads();
        ntt_inv_lds(buf[team & 1], t, tw, c, nc.inv_n[c], nc.inv_n_p[c], team < 2);
        uint32_t own_a[2], own_b[2];
        uint32_t* xw = xb + (size_t)(k & 1) * 4 * N + (size_t)c * 2 * N;
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t i = tid + 1024 * e;
            own_a[e] = buf[0][ls(i)];
            own_b[e] = buf[1][ls(i)];
            xw[i] = own_a[e];
            xw[N + i] = own_b[e];
        }
        // the hand-over rules in dev_common.h: every wave's stores complete,
        // one release, the flag; the partner's flag, one acquire, the barrier
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) {
            fence_release_agent_done();
            __hip_atomic_store(my_flag, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // a zero bound reports at the first exchange whatever the partner
            // did (a deterministic way into this failure path)
            uint32_t spins = 0;
            while (spin_limit == 0 || __hip_atomic_load(other_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < k) {
                if (++spins > spin_limit) {
                    __hip_atomic_store(status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    abort_flag = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            fence_acquire_agent_done();  // the partner's values, for every wave after the barrier
        }
        __syncthreads();
        if (abort_flag) {
            // let the partner stop waiting for this step too
            if (tid == 0) __hip_atomic_store(my_flag, 0xffffffffu, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        const uint32_t* xr = xb + (size_t)(k & 1) * 4 * N + (size_t)(1 - c) * 2 * N;
        uint32_t dA[2][2], dB[2][2];  // [e][digit], residues mod this limb's prime
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t i = tid + 1024 * e;
            const uint32_t oa = xr[i], ob = xr[N + i];
            uint32_t d0[2], d1[2];
            gadget2(c ? oa : own_a[e], c ? own_a[e] : oa, cc, d0, d1);
            dA[e][0] = c ? d1[0] : d0[0];
            dA[e][1] = c ? d1[1] : d0[1];
            gadget2(c ? ob : own_b[e], c ? own_b[e] : ob, cc, d0, d1);
            dB[e][0] = c ? d1[0] : d0[0];
            dB[e][1] = c ? d1[1] : d0[1];
        }
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t i = tid + 1024 * e;
            buf[0][ls(i)] = dA[e][0];
            buf[1][ls(i)] = dA[e][1];
            buf[2][ls(i)] = dB[e][0];
            buf[3][ls(i)] = dB[e][1];
        }
        __syncthreads();
        ntt_fwd_lds(buf[team], t, tw, c);
        // RGSW parts: 0 top_a, 1 top_b, 2 bot_a, 3 bot_b (a-digits use top, b-digits bottom)
        const uint32_t gi = (uint32_t)gs;
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t i = tid + 1024 * e;
            uint64_t sa = 0, sb = 0;
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const uint64_t xa = buf[j][ls(i)], xbv = buf[2 + j][ls(i)];
                sa += xa * rgl[0][j][i] + xbv * rgl[2][j][i];
                sb += xa * rgl[1][j][i] + xbv * rgl[3][j][i];
            }
            acc_a[e] = add_mod(red64(sa, c), pa_base[(size_t)gi * pa_stride + (size_t)c * N + i], q);
            acc_b[e] = add_mod(red64(sb, c), pb[(size_t)gi * 2 * N + (size_t)c * N + i], q);
        }
        __syncthreads();
    }
    if (lo > start) {  // the chain continues in a later launch
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t i = tid + 1024 * e;
            hs[(size_t)c * N + i] = acc_a[e];
            hs[2 * (size_t)N + (size_t)c * N + i] = acc_b[e];
        }
        return;
    }
    uint64_t* o = out + ((size_t)qi * T + ct) * 4 * N;
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const uint32_t i = tid + 1024 * e;
        o[(size_t)c * N + i] = acc_a[e];
        o[2 * (size_t)N + (size_t)c * N + i] = acc_b[e];
    }
}

}  // namespace

size_t horner_pair_xbuf_words(size_t nchains) { return nchains * 2 * 2 * 2 * N; }

hipError_t launch_horner_pair(const uint32_t* pa_base, size_t pa_stride, size_t G, size_t Da, size_t T,
                              const uint32_t* packed_b, size_t packed_stride, const uint32_t* rgsw,
                              size_t rgsw_stride, uint64_t* out, uint32_t* hstate, size_t g_lo, size_t g_hi,
                              size_t nq, uint32_t* xbuf, uint32_t* flags, uint32_t* status, uint32_t spin_limit,
                              const NttConst& nc, const CrtConst& cc, hipStream_t s) {
    const size_t nchains = nq * T;
    if (nchains == 0 || g_lo >= g_hi) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, flags, 2 * nchains, 0u);
    const size_t blocks = 16 * ((nchains + 7) / 8);
    hipLaunchKernelGGL(k_horner_pair, dim3((unsigned)blocks), dim3(1024), 0, s, pa_base, pa_stride, (uint32_t)G,
                       (uint32_t)Da, (uint32_t)T, packed_b, packed_stride, rgsw, rgsw_stride, out, hstate,
                       (uint32_t)g_lo, (uint32_t)g_hi, xbuf, flags, status, (uint32_t)nchains, spin_limit, nc, cc);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // namespace ipx
