// On-device wire codecs. The byte formats are the client library's
// (reference src/capi_client.cpp:174-304 for the 53-bit CRT packing,
// :310-385 and src/crypto.cpp:285-313 for the response compression); running
// them next to the answer pipeline keeps the host-device crossings at the
// packed sizes (a query's 57344 values as 53-bit fields instead of 2 x 8-byte
// residues, a response ct as 12288 bytes instead of 32 KiB of residues).
#include "kernels.h"

namespace ipx {

namespace {

constexpr int BETA = 53;  // bits of one CRT-packed value (Q0*Q1 < 2^53)

// One value of the MSB-first packed stream per thread: fields start at bit
// v*53 of the query's bytes; the value's residues go where the flat query
// keeps them (lwe limbs at v and R+v; part j of the 6*D_EFF key/RGSW parts
// at 2R + 2jN + i and + N).
__global__ __launch_bounds__(256) void k_query_unpack(const uint8_t* __restrict__ packed, size_t pbytes, uint32_t R,
                                                      uint64_t* __restrict__ flat, size_t qlen) {
    const uint32_t v = blockIdx.x * 256 + threadIdx.x;
    const uint32_t V = R + 6 * D_EFF * N;
    if (v >= V) return;
    const uint8_t* src = packed + (size_t)blockIdx.y * pbytes;
    const uint64_t bit0 = (uint64_t)v * BETA;
    const uint32_t skip = (uint32_t)(bit0 & 7);
    const uint32_t nbytes = (skip + BETA + 7) >> 3;  // <= 8: the bytes holding bits [bit0, bit0 + 53)
    const uint8_t* b = src + (bit0 >> 3);
    uint64_t acc = 0;
    for (uint32_t i = 0; i < nbytes; i++) acc = (acc << 8) | b[i];
    const uint64_t val = (acc >> (nbytes * 8 - skip - BETA)) & ((1ull << BETA) - 1);
    size_t i0, i1;
    if (v < R) {
        i0 = v;
        i1 = (size_t)R + v;
    } else {
        const uint32_t j = (v - R) / N, i = (v - R) % N;
        i0 = 2 * (size_t)R + (size_t)j * 2 * N + i;
        i1 = i0 + N;
    }
    uint64_t* dst = flat + (size_t)blockIdx.y * qlen;
    dst[i0] = val % Q0;
    dst[i1] = val % Q1;
}

__global__ __launch_bounds__(256) void k_narrow_u32(const uint64_t* __restrict__ src, size_t n,
                                                    uint32_t* __restrict__ dst) {
    for (size_t x = (size_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (size_t)gridDim.x * 256)
        dst[x] = (uint32_t)src[x];
}

// The value in [0, Q0*Q1) with residues r0, r1 (Garner).
__device__ __forceinline__ uint64_t crt_value(uint32_t r0, uint32_t r1, uint32_t q1inv_q0) {
    const uint64_t d = ((uint64_t)r0 + Q0 - r1 % Q0) % Q0;
    return r1 + (uint64_t)Q1 * (d * q1inv_q0 % Q0);
}

// floor((v * 2^k + floor(Q/2)) / Q) mod 2^k for v < Q = Q0*Q1: a double
// estimate (every input below 2^81 is represented to 2^-52 relative, so
// the estimate is within one of the quotient, which is below 2^29), then
// one exact correction step each way in 128-bit integers.
__device__ __forceinline__ uint32_t mod_switch(uint64_t v, int k) {
    const uint64_t Qh = QFULL / 2;
    const unsigned __int128 num = ((unsigned __int128)v << k) + Qh;
    const double est = ((double)v * (double)(1ull << k) + (double)Qh) / (double)QFULL;
    uint64_t q = est > 0.0 ? (uint64_t)est : 0;
    __int128 r = (__int128)num - (__int128)((unsigned __int128)q * QFULL);
    if (r < 0) {
        q--;
        r += (__int128)QFULL;
    }
    if (r >= (__int128)QFULL) {
        q++;
        r -= (__int128)QFULL;
    }
    return (uint32_t)(q & ((1ull << k) - 1));
}

// One coefficient pair of one ct per thread: a' (28 bits) of coefficients
// 2p, 2p+1 fill bytes [7p, 7p+7) of the ct's 12288, b' (20 bits) bytes
// 7168 + [5p, 5p+5) (MSB first; both halves end on a byte boundary).
__global__ __launch_bounds__(256) void k_compress(const uint32_t* __restrict__ coeff, uint32_t nct,
                                                  uint8_t* __restrict__ out, CrtConst cc) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    const uint32_t ct = blockIdx.y;
    if (p >= N / 2 || ct >= nct) return;
    const uint32_t* a0 = coeff + (size_t)ct * 4 * N;
    const uint32_t *a1 = a0 + N, *b0 = a0 + 2 * N, *b1 = a0 + 3 * N;
    uint8_t* o = out + (size_t)ct * 12288;
    const uint64_t wa = ((uint64_t)mod_switch(crt_value(a0[2 * p], a1[2 * p], cc.q1inv_q0), 28) << 28) |
                        mod_switch(crt_value(a0[2 * p + 1], a1[2 * p + 1], cc.q1inv_q0), 28);
    const uint64_t wb = ((uint64_t)mod_switch(crt_value(b0[2 * p], b1[2 * p], cc.q1inv_q0), 20) << 20) |
                        mod_switch(crt_value(b0[2 * p + 1], b1[2 * p + 1], cc.q1inv_q0), 20);
#pragma unroll
    for (int k = 0; k < 7; k++) o[7 * p + k] = (uint8_t)(wa >> (48 - 8 * k));
#pragma unroll
    for (int k = 0; k < 5; k++) o[7168 + 5 * p + k] = (uint8_t)(wb >> (32 - 8 * k));
}

}  // namespace

size_t query_packed_bytes(size_t R) { return ((R + 6 * (size_t)D_EFF * N) * BETA + 7) / 8; }

hipError_t launch_query_unpack(const uint8_t* packed, size_t count, size_t R, uint64_t* flat, size_t qlen,
                               hipStream_t s) {
    if (count == 0) return hipSuccess;
    if (qlen < 2 * R + 12 * (size_t)D_EFF * N) return hipErrorInvalidValue;
    const size_t V = R + 6 * (size_t)D_EFF * N;
    hipLaunchKernelGGL(k_query_unpack, dim3((unsigned)((V + 255) / 256), (unsigned)count), dim3(256), 0, s, packed,
                       query_packed_bytes(R), (uint32_t)R, flat, qlen);
    return hipGetLastError();
}

hipError_t launch_compress(const uint64_t* resp, size_t nct, uint32_t* scratch, uint8_t* out, const NttConst& nc,
                           const CrtConst& cc, hipStream_t s) {
    if (nct == 0) return hipSuccess;
    const size_t words = nct * 4 * N;
    hipLaunchKernelGGL(k_narrow_u32, dim3((unsigned)((words + 255) / 256 < 4096 ? (words + 255) / 256 : 4096)),
                       dim3(256), 0, s, resp, words, scratch);
    hipError_t e = launch_ntt_rns(scratch, nct * 2, nc, true, s);  // [ct][a|b][limb][N] -> coefficient form
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_compress, dim3(N / 2 / 256, (unsigned)nct), dim3(256), 0, s, scratch, (uint32_t)nct, out,
                       cc);
    return hipGetLastError();
}

}  // namespace ipx
