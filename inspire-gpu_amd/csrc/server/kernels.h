// Host-side launchers of the gfx950 kernels. Every launcher enqueues on the
// given stream and returns the launch status; none allocates or
// synchronises.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

#include "dev_common.h"

namespace ipx {

// ---------------------------------------------------------------- online
// Per-query device layout produced by unmarshal (one "slot"):
//   bq  : double [2][R]            (b residues as exact doubles, DB pass)
//   keys: u32 [2 keys][digit][limb][N]  (ksk5, kskneg1; coeff -> NTT in place)
//   rgsw: u32 [4 parts][digit][limb][N]  (NTT form)
struct SlotLayout {
    size_t R, C, G;
    size_t keys_words() const { return 2 * (size_t)D_EFF * 2 * N; }
    size_t rgsw_words() const { return 4 * (size_t)D_EFF * 2 * N; }
    size_t query_u64s() const { return 2 * R + 6 * (size_t)D_EFF * 2 * N; }
};

hipError_t launch_unmarshal(const uint64_t* d_q, size_t count, size_t R, double* bq, uint32_t* keys,
                            uint32_t* rgsw, hipStream_t s);
// In-place forward NTT of `npoly` RNS polynomials laid out [npoly][2][N].
hipError_t launch_ntt_rns(uint32_t* polys, size_t npoly, const NttConst& nc, bool inverse, hipStream_t s);
// dst <- src, bytes a multiple of 16 (grid-stride, 16 bytes per lane).
hipError_t launch_copy16(const void* src, void* dst, size_t bytes, hipStream_t s);

// Wire codecs on the device (the client library's byte formats).
// Bytes of one query in the 53-bit CRT-packed form.
size_t query_packed_bytes(size_t R);
// count packed queries (query_packed_bytes(R) bytes each, back to back) ->
// flat u64 queries (qlen u64s apart).
hipError_t launch_query_unpack(const uint8_t* packed, size_t count, size_t R, uint64_t* flat, size_t qlen,
                               hipStream_t s);
// nct response cts (flat u64, NTT form, [ct][a|b][limb][N]) -> 12288
// compressed bytes each; scratch holds nct*4*N words.
hipError_t launch_compress(const uint64_t* resp, size_t nct, uint32_t* scratch, uint8_t* out, const NttConst& nc,
                           const CrtConst& cc, hipStream_t s);

// Database pass over the resident DB, byte planes in 16-row blocks: the
// low (p = 0) and high (p = 1) byte of (r, c) at ((r/16)*2 + p)*C*16 +
// c*16 + r%16. out + q*out_qstride + slice*2C holds [limb][C] = sum over the
// slice's rows of db[r][c]*b_{q,limb}[r] mod q_limb.
struct DbPlan {
    enum Kind { F64, MFMA } kind;
    uint32_t param;      // F64: 16-row blocks per iteration (2, 4, 8); MFMA: 16-row digit tiles (1, 2, 4, 7, 8, 14, 16)
    uint32_t slices;     // row slices (power of two)
    uint32_t depth = 3;  // 8 or 16 digit tiles: register-ring slots of DB operands (kMfmaDepths)
    bool dbtile = false; // 16 digit tiles: the waves share the DB tile (k_dbpass_mfma_bs), not the digit tile
    uint32_t waves = 4;  // 8 or 16 digit tiles, shared digits: waves per block (4, or 8 = one block per CU)
    bool crt = false;    // shared digits: 7 digit rows per query (CRT value), 14 tiles (7 for power rows, 9..16 queries)
    bool pow = false;    // crt digits accumulated by power of 256 (8-wave blocks, two column tiles per wave)
    uint32_t nt = 0;     // 1 digit tile: 16-column tiles per wave, 8 (streamed passes) or 0 = 4
};
// Forms of the 9..32-query pass: 0 shared digits in 4-wave blocks, 1 shared
// DB tile (17..32 queries), 2 shared digits in 8-wave blocks, 3 and 4 the
// 8- and 4-wave shared-digit forms with 7 digit rows per query (17..32
// queries), 5 those digits accumulated by power (9..32 queries).
constexpr int kMfmaForms = 6;
// Ring depths of the 9..32-query MFMA passes; plans cap them at what each
// kernel is built with (7 for 14 or 16 shared-digit tiles in 8-wave blocks,
// 5 in 4-wave blocks, 4 otherwise).
constexpr uint32_t kMfmaDepths[] = {3, 4, 5, 6, 7};
// Row-major staging rows (16*nkb of them) -> blocks [kb0, kb0 + nkb) of the planes.
hipError_t launch_to_planes(const uint16_t* src, uint8_t* dst, size_t C, size_t kb0, size_t nkb, hipStream_t s);
// Plans for a pass over ncols columns (the slice count fills the device).
DbPlan dbplan_f64(size_t R, size_t ncols, int unroll, size_t target_blocks);
// Queries one batched (int8 MFMA) database pass serves.
constexpr size_t kMfmaMaxQueries = 32;
// int8-MFMA plan for nq (1..kMfmaMaxQueries) queries.
DbPlan dbplan_mfma(size_t R, size_t ncols, size_t nq, size_t target_blocks, uint32_t depth = 4,
                   bool dbtile = false, uint32_t waves = 4, bool crt = false, bool pow = false);
// The same for a form (0 .. kMfmaForms-1) of the 9..32-query pass.
DbPlan dbplan_form(size_t R, size_t ncols, size_t nq, size_t target_blocks, uint32_t depth, int form);
// Queries an MFMA plan's digit tiles hold.
size_t dbplan_max_queries(const DbPlan& p);
// Digit rows the MFMA plan stages per 16-row block (int8 scratch: R*Mp bytes, Mp <= 256).
size_t dbplan_mp(const DbPlan& p);
// Column granularity of a plan (one block's columns): col0 and ncols are multiples of it.
size_t dbplan_cols(const DbPlan& p);
// Name of the kernel a plan launches (as profilers report it, without arguments).
std::string dbplan_kernel(const DbPlan& p);
// Output of the database passes: out points at the first column's residue
// (limb 0, row slice 0); limb 1 sits ostride words after limb 0 and row
// slice k at k*2*ostride (ostride = C into interm, ncols for partials kept
// relative to col0).
// One query, fp64 FMA on the vector ALUs, over columns [col0, col0 + ncols).
hipError_t launch_dbpass_f64(const uint8_t* db, const double* bq, uint32_t* out, size_t ostride, size_t R, size_t C,
                             size_t col0, size_t ncols, const DbPlan& p, hipStream_t s);
// Query digits and offset sums of nq queries (bq [q][limb][R] doubles) for
// an MFMA plan: dig (R*Mp bytes) and S (2*nq words).
hipError_t launch_mfma_digits(const double* bq, size_t nq, size_t R, const DbPlan& p, int8_t* dig, uint32_t* S,
                              hipStream_t s);
// The int8-MFMA pass of nq queries over columns [col0, col0 + ncols), from
// launch_mfma_digits' output for a plan with the same digit tile count.
//
// A streamed pass (ps set; full width, direct kernels and shared-digit
// forms only) hands its groups over while it runs: its 1-D grid takes the
// column tiles from the highest down (the order the Horner chains consume
// groups), each tile's row slices in a row; every workgroup, once its
// stores are complete at agent scope (one release per workgroup), adds 1 to
// ps.group_done[column / N]. A group's columns are in place (out: one
// slice's final values, or [q][slice][limb][OC] partials) when its word
// reaches (N / dbplan_cols(p)) * p.slices. group_done must be zero before
// the pass. The pass itself never waits for anything, so whatever waits
// beside it for its groups cannot stall it beyond taking CUs.
// A streamed launch may cover a whole number of groups [col0 / N, ...)
// (a pass split into launches): its counts go to group_done[col0 / N + ...].
struct PassStream {
    uint32_t* group_done = nullptr;
    uint32_t group_base = 0;   // first group of the launch
    uint32_t tiles = 0;        // column tiles of the launch (ncols / dbplan_cols(p))
    uint32_t slice_shift = 0;  // log2 of the plan's row slices
    uint32_t tiles_per_group = 0;
};
hipError_t launch_dbpass_mfma(const uint8_t* db, const int8_t* dig, const uint32_t* S, size_t nq, uint32_t* out,
                              size_t out_qstride, size_t ostride, size_t R, size_t C, size_t col0, size_t ncols,
                              const DbPlan& p, hipStream_t s, const PassStream* ps = nullptr);
// Whether launch_dbpass_mfma streams plan p over C columns.
bool dbplan_streams(const DbPlan& p, size_t C);
// n words at p set to v (a kernel, not a memset: graph-capture safe).
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
// interm[q][limb][col0 + j] = sum over the S row slices of part (per query
// stride part_stride; [slice][limb][j], j < ncols), for j < ncols.
hipError_t launch_reduce_slices(const uint32_t* part, size_t part_stride, uint32_t S, size_t C, uint32_t* interm,
                                size_t nq, size_t col0, size_t ncols, hipStream_t s);

// Queries one collapse pass serves.
constexpr size_t kCollapseMaxQueries = 32;
// Key digits and key sums of nq queries whose NTT-form keys sit at
// keys + q*kstride, for launch_collapse: kdig (1 MiB) and KS (2*nq words).
hipError_t launch_collapse_keys(const uint32_t* keys, size_t kstride, size_t nq, int8_t* kdig, uint32_t* KS,
                                hipStream_t s);
// Precomp-tensor pass in dot-product form (int8 MFMA, Dd read once) of ng
// groups whose precomp blocks start at precomp_g0, for nq queries:
// part2_g0 + q*p2_stride + g*2N holds [limb][N] of group g (relative).
// A launch too small to fill the chip splits its contraction into row
// slices whose residues add up in part2 (zeroed first on s): the words are
// then sums of up to 8 residues (< 2^30), which every reader reduces.
// 17..32 queries come in kCollapseVariants forms: 0 the two-step
// prefetch kernel, 1..5 the ring kernel (depth 3, 4 in 4-wave blocks; 3, 4,
// 5 in 8-wave blocks); fewer queries ignore the variant.
constexpr int kCollapseVariants = 6;
hipError_t launch_collapse(const uint32_t* precomp_g0, size_t ng, const int8_t* kdig, const uint32_t* KS, size_t nq,
                           uint32_t* part2_g0, size_t p2_stride, hipStream_t s, int variant = 0);
// A streamed pass of up to 8 queries with the collapse in the same launch:
// per group from the top, the collapse (all groups' precomp, part2 as in
// launch_collapse) and the database pass of plan p (a direct-kernel plan,
// 256-column tiles; out as a streamed pass's) as 256-lane work items;
// group_done[g] counts them (zero before the launch) and reaches
// pass_collapse_target(p, nq, G) when both halves of group g are in place.
// The key digits (launch_collapse_keys) and query digits must be ready.
uint32_t pass_collapse_target(const DbPlan& p, size_t nq, size_t G);
hipError_t launch_pass_collapse_stream(const uint8_t* db, const int8_t* dig, const uint32_t* S, size_t nq,
                                       uint32_t* out, size_t out_qstride, size_t R, size_t C, const DbPlan& p,
                                       const uint32_t* precomp, const int8_t* kdig, const uint32_t* KS,
                                       uint32_t* part2, size_t p2_stride, uint32_t* group_done, hipStream_t s);
// packed_b[q][g] = NTT(interm[q] group slice) - collapse sums - Df*Kn, for
// groups [g0, g0 + ng) (absolute indexes into every buffer).
hipError_t launch_pack_finalize(const uint32_t* interm, size_t C, const uint32_t* part2, size_t part2_stride,
                                const uint32_t* precomp, size_t g0, size_t ng, const uint32_t* keys,
                                size_t keys_stride, uint32_t* packed_b, size_t packed_stride, size_t nq,
                                const NttConst& nc, hipStream_t s);
// The same output from one 1024-lane workgroup per query (for one-query
// chunks beside database passes, where it starts without waiting for slots).
hipError_t launch_pack_finalize_wg(const uint32_t* interm, size_t C, const uint32_t* part2, size_t part2_stride,
                                   const uint32_t* precomp, size_t g0, size_t ng, const uint32_t* keys,
                                   size_t keys_stride, uint32_t* packed_b, size_t packed_stride, size_t nq,
                                   const NttConst& nc, hipStream_t s);
// Horner over each response ciphertext, restricted to the groups in
// [g_lo, g_hi): a chain whose top group lies in the range starts there,
// one that started in an earlier launch resumes from hstate
// ([q][ct][a|b][limb][N] u32), one that continues below g_lo saves its state
// there, and a finished chain writes the flat u64 response
// [q][ct][a|b][limb][N] (launches go from the highest group range down).
// Group g's packed a-part (shared by all queries) is at pa_base +
// g*pa_stride; its b-part at packed_b + q*packed_stride + g*2N.
// With pack.interm set, each chain first computes the packed b-parts of
// its own groups in [g_lo, g_hi) (launch_pack_finalize's output, written
// to packed_b) in its own workgroup, so a chunk's pack needs no grid of its
// own beside the database passes.
struct FusedPack {
    const uint32_t* interm = nullptr;  // [q][limb][C] (query 0 of the launch)
    size_t C = 0;
    const uint32_t* part2 = nullptr;   // collapse sums, [q][g][limb][N]
    size_t part2_stride = 0;
    const uint32_t* precomp = nullptr; // arena base (Df of every group)
    const uint32_t* keys = nullptr;    // [q][key][digit][limb][N]
    size_t keys_stride = 0;
};
// With wait.done set (beside a streamed pack), chain q waits before the
// step of group g until done[(q / wait.queries) * wait.stride + g] is
// nonzero (the group's packed b-parts are published). A wait longer than
// wait.spin_limit sleeps (0: report at once) sets *wait.status to 1, writes
// all-ones response words (no residue has that value) and ends the chain.
struct StreamWait {
    const uint32_t* done = nullptr;
    uint32_t queries = 1, stride = 0;
    uint32_t spin_limit = 0;
    uint32_t* status = nullptr;
};
// The streamed pack beside a streamed database pass: one 1024-lane
// workgroup per ps.queries queries (1, 2 or kPackQueries; its 4 teams
// take 4 of the queries' (query, limb) polynomials per round) takes the
// groups from the top down;
// for each it waits until pass_done[g] reaches pass_target, then packs the
// group's b-parts of its queries (the pass's row-slice partials summed, NTT,
// minus the collapse sums and Df x Kn: launch_pack_finalize's output) into
// packed_b and publishes them: packed_done[block * G + g] = 1 after an
// agent-scope release. packed_done must be zero before the launch. A wait
// longer than spin_limit sleeps sets *status to 1 and ends the workgroup
// (its chains then give up their own waits).
constexpr uint32_t kPackQueries = 4;
struct PackStream {
    const uint32_t* part = nullptr;  // the pass's output: [q][slice][limb][C] partials
    size_t part_qstride = 0;         //   (one slice: its values, [q][limb][C])
    uint32_t slices = 1;
    uint32_t C = 0;
    const uint32_t* pass_done = nullptr;
    uint32_t pass_target = 0;
    uint32_t* packed_done = nullptr;
    uint32_t spin_limit = 0;
    uint32_t* status = nullptr;
    uint32_t queries = kPackQueries;  // queries per workgroup
};
hipError_t launch_pack_stream(const PackStream& ps, const uint32_t* part2, size_t part2_stride,
                              const uint32_t* precomp, const uint32_t* keys, size_t keys_stride, uint32_t* packed_b,
                              size_t packed_stride, size_t G, size_t nq, const NttConst& nc, hipStream_t s);
hipError_t launch_horner(const uint32_t* pa_base, size_t pa_stride, size_t G, size_t Da, size_t T,
                         uint32_t* packed_b, size_t packed_stride, const uint32_t* rgsw, size_t rgsw_stride,
                         uint64_t* out, uint32_t* hstate, size_t g_lo, size_t g_hi, size_t nq, const NttConst& nc,
                         const CrtConst& cc, hipStream_t s, const FusedPack& pack = FusedPack{},
                         const StreamWait& wait = StreamWait{});

// The same chains with two workgroups per chain (one per limb) exchanging
// coefficient-form values once per step. xbuf holds
// horner_pair_xbuf_words(nq*T) words and flags 2*nq*T words (reset by the
// launch); a partner wait longer than spin_limit sleeps sets *status to 1
// and the chain's output is then undefined. spin_limit 0 sets it at the
// first exchange.
size_t horner_pair_xbuf_words(size_t nchains);
hipError_t launch_horner_pair(const uint32_t* pa_base, size_t pa_stride, size_t G, size_t Da, size_t T,
                              const uint32_t* packed_b, size_t packed_stride, const uint32_t* rgsw,
                              size_t rgsw_stride, uint64_t* out, uint32_t* hstate, size_t g_lo, size_t g_hi,
                              size_t nq, uint32_t* xbuf, uint32_t* flags, uint32_t* status, uint32_t spin_limit,
                              const NttConst& nc, const CrtConst& cc, hipStream_t s);

// ---------------------------------------------------------------- preprocess
hipError_t launch_encode(uint16_t* db, size_t R, size_t C, size_t Da, hipStream_t s);
// a32T[g][limb][p][j] for groups [g0, g0+ng): transposed a-side matmul
// output. crs [nr][2][N] NTT form.
hipError_t launch_aside(const uint8_t* db, size_t R, size_t C, size_t g0, size_t ng, const uint32_t* crs,
                        uint32_t* a32T, const NttConst& nc, hipStream_t s);
// F = NTT along j of a32T (in place), per (g, limb, p).
hipError_t launch_ring_embed(uint32_t* a32T, size_t ng, const NttConst& nc, hipStream_t s);
// CollapseA chains (forward and conjugate) for ng groups; writes D+/D- in
// the reference order to dtmp [ng][dir][RD_SZ] and the final chain states
// into states [ng][2 dir][2][N].
hipError_t launch_collapse_chain(const uint32_t* F, size_t ng, const uint32_t* ksk5a, uint32_t* dtmp,
                                 uint32_t* states, const NttConst& nc, const CrtConst& cc, hipStream_t s);
// D+/D- (reference order, group g at Dp/Dm + g*gstride) -> the Dd planes of
// precomp blocks [0, ng) at precomp_g0.
hipError_t launch_dd_relayout(const uint32_t* Dp, const uint32_t* Dm, size_t gstride, size_t ng, uint32_t* precomp_g0,
                              hipStream_t s);
hipError_t launch_collapse_final(const uint32_t* states, size_t ng, const uint32_t* kneg1a,
                                 uint32_t* precomp_g0, const NttConst& nc, const CrtConst& cc, hipStream_t s);

}  // namespace ipx
