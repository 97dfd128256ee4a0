// Horner evaluation over device buffers that allocates nothing: the
// variant of ipir_dbg_horner (include/inspire_debug.h) for callers that
// cannot allocate device memory (code under HIP graph capture, tight
// serving loops). No counterpart in the reference, whose gpu_horner_eval
// (src/gpu_online.cu:474-540) allocates its scratch per call.
//
// Its device memory beyond the caller's inputs and outputs is one
// file-scope array per GPU (`scratch`, kScratchWords u32 = 8.3 MiB, held
// until the process ends):
//   [twiddles 8N][packed a-parts kMaxL x 2N][packed b-parts kMaxL x 2N]
//   [RGSW rows 4 x 2 x 2N][chain state 4N]
// The twiddles are built in place by a kernel at a device's first call
// (and again in any call made while the stream is capturing, so a captured
// graph carries them too); every other region is overwritten by each
// call before it is read, so nothing must start at zero. Calls that
// overlap on one GPU share the array and race: callers order them (one
// stream, or events).
#include <mutex>

#include "inspire_debug.h"
#include "kernels.h"
#include "server_ctx.h"

using namespace ipx;

namespace {

constexpr size_t kMaxL = 256;  // the longest Horner chain of any geometry (D = 256)
constexpr size_t kBlk = (size_t)D_EFF * 2 * N;
constexpr size_t kTwOff = 0, kTwWords = 8 * (size_t)N;
constexpr size_t kPaOff = kTwOff + kTwWords;
constexpr size_t kPbOff = kPaOff + kMaxL * 2 * N;
constexpr size_t kRgOff = kPbOff + kMaxL * 2 * N;
constexpr size_t kHsOff = kRgOff + 4 * kBlk;
constexpr size_t kScratchWords = kHsOff + 4 * (size_t)N;

static __device__ __align__(256) uint32_t scratch[kScratchWords];

// psi^e mod q by square and multiply (q < 2^27, so products fit in u64)
__device__ uint32_t pow_mod(uint64_t b, uint32_t e, uint32_t q) {
    uint64_t r = 1;
    for (b %= q; e; e >>= 1, b = b * b % q)
        if (e & 1) r = r * b % q;
    return (uint32_t)r;
}

// The NTT twiddle tables of both limbs in NttConst's layout: forward
// psi^BR(i), its Shoup companion floor(w 2^32 / q), inverse psi^(2N-BR(i))
// (= (psi^BR(i))^-1, 1 at i = 0), its companion; the same values as the
// host tables every other kernel uploads.
__global__ __launch_bounds__(256) void k_build_twiddles(uint32_t* __restrict__ tw, uint32_t psi0, uint32_t psi1) {
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;  // [limb][i]
    if (x >= 2 * N) return;
    const uint32_t limb = x / N, i = x % N;
    const uint32_t q = limb ? Q1 : Q0, psi = limb ? psi1 : psi0;
    const uint32_t e = brev11(i);
    const uint32_t w = pow_mod(psi, e, q), iw = pow_mod(psi, (2 * N - e) % (2 * N), q);
    uint32_t* t = tw + (size_t)limb * 4 * N;
    t[i] = w;
    t[N + i] = (uint32_t)(((uint64_t)w << 32) / q);
    t[2 * N + i] = iw;
    t[3 * N + i] = (uint32_t)(((uint64_t)iw << 32) / q);
}

// Inputs as the dbg entry point takes them, reduced to u32 residues:
// cts [L][a|b][limb][N] u64 -> pa [L][limb][N], pb [L][limb][N];
// rgsw [part][digit][limb][N] u64 -> the same order in u32.
__global__ __launch_bounds__(256) void k_narrow_inputs(const uint64_t* __restrict__ cts, size_t L,
                                                       const uint64_t* __restrict__ rgsw, uint32_t* __restrict__ pa,
                                                       uint32_t* __restrict__ pb, uint32_t* __restrict__ rg) {
    const size_t nct = L * 4 * N, total = nct + 4 * kBlk;
    for (size_t x = (size_t)blockIdx.x * 256 + threadIdx.x; x < total; x += (size_t)gridDim.x * 256) {
        if (x < nct) {
            const size_t g = x / (4 * N), r = x % (4 * N);  // r: [a|b][limb][N]
            const uint32_t limb = (uint32_t)((r / N) & 1);
            const uint32_t v = (uint32_t)(cts[x] % (limb ? Q1 : Q0));
            (r < 2 * N ? pa : pb)[g * 2 * N + r % (2 * N)] = v;
        } else {
            const size_t y = x - nct;
            rg[y] = (uint32_t)(rgsw[y] % ((y / N) & 1 ? Q1 : Q0));
        }
    }
}

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t* g_scratch[kMaxDevices] = {};  // the array's device address, per device
bool g_tables[kMaxDevices] = {};        // twiddles built (outside a capture)

}  // namespace

extern "C" int ipir_horner_noalloc(const uint64_t* d_cts, size_t L, const uint64_t* d_rgsw, uint64_t* d_out,
                                   void* hip_stream) {
    if (!d_cts || !d_rgsw || !d_out || L == 0) return 1;
    if (L > kMaxL) return 2;
    const hipStream_t s = static_cast<hipStream_t>(hip_stream);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 4;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) != hipSuccess) return 4;
    uint32_t* base = nullptr;
    bool build = false;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_scratch[dev]) {
            void* p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return 4;
            g_scratch[dev] = static_cast<uint32_t*>(p);
        }
        base = g_scratch[dev];
        build = !g_tables[dev] || cap != hipStreamCaptureStatusNone;
        if (cap == hipStreamCaptureStatusNone) g_tables[dev] = true;
    }
    NttConst nc{};
    CrtConst cc{};
    for (int l = 0; l < 2; l++) {
        const inspire::NttTables& t = inspire::ntt_tables(l);
        nc.inv_n[l] = (uint32_t)t.inv_n;
        nc.inv_n_p[l] = (uint32_t)t.inv_n_prime;
    }
    nc.tw = base + kTwOff;
    cc.q1inv_q0 = (uint32_t)inspire::mod_inv(inspire::Q1 % inspire::Q0, inspire::Q0);
    cc.q0inv_q1 = (uint32_t)inspire::mod_inv(inspire::Q0 % inspire::Q1, inspire::Q1);
    if (build) {
        hipLaunchKernelGGL(k_build_twiddles, dim3(2 * N / 256), dim3(256), 0, s, base + kTwOff,
                           (uint32_t)inspire::ntt_tables(0).psi, (uint32_t)inspire::ntt_tables(1).psi);
        if (hipGetLastError() != hipSuccess) return 4;
    }
    uint32_t *pa = base + kPaOff, *pb = base + kPbOff, *rg = base + kRgOff, *hs = base + kHsOff;
    hipLaunchKernelGGL(k_narrow_inputs, dim3(512), dim3(256), 0, s, d_cts, L, d_rgsw, pa, pb, rg);
    if (hipGetLastError() != hipSuccess) return 4;
    // one chain over groups [0, L): acc = packed[L-1], then ext_prod and add down to 0
    if (launch_horner(pa, 2 * N, L, L, 1, pb, L * 2 * N, rg, 4 * kBlk, d_out, hs, 0, L, 1, nc, cc, s) != hipSuccess)
        return 4;
    return 0;
}
