"""GPU parity: every device stage against the oracle, bit for bit, through
the C ABI (and the inspection entry points of include/inspire_debug.h).

Covers the reference's GPU test classes (tests/test_gpu_ntt.cu,
test_gpu_extprod.cu / test_gpu_horner.cu, test_e2e_gpu.cu,
test_gpu_batch.cu, test_capi.cu) plus direct checks of the stages the
reference never pins (encode, precomp tensors, collapse).
"""
import ctypes as C
import sys

import numpy as np
import pytest

import inspire_amd as ia
import oracle_lib as ol
from oracle_lib import BLK, N, NSTEPS, P, Q0, Q1, QS

pytestmark = pytest.mark.gpu
SEED = np.array([0x42] + [0] * 63, np.uint8)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ---------------------------------------------------------------- kernels
def test_ntt_kernel_matches_oracle():
    rng = np.random.default_rng(11)
    npoly = 16
    x = ol.rand_poly(rng, (npoly,))
    lib = ia.load("server")
    for inv in (0, 1):
        xin = x.astype(np.uint32)
        out = np.zeros_like(xin)
        assert lib.ipir_dbg_ntt(ia.ptr(xin), ia.ptr(out), npoly, inv) == 0
        want = x.copy()
        for p in range(npoly):
            for limb in (0, 1):
                v = want[p, limb].copy()
                (ol.oracle().orc_ntt_inv if inv else ol.oracle().orc_ntt_fwd)(v, limb)
                want[p, limb] = v
        assert np.array_equal(out.astype(np.uint64), want), f"inverse={inv}"


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("L", [1, 2, 5, 9])
def test_horner_kernel_matches_oracle(L, variant):
    rng = np.random.default_rng(L)
    rc, geo = ol.setup(262144, 120, 2048)
    q, _ = ol.make_query(geo, SEED, 1000, 3)
    rg = np.ascontiguousarray(q[2 * geo["R"] + 2 * BLK:])
    cts = ol.rand_poly(rng, (L, 2)).reshape(-1)
    out = np.zeros(4 * N, np.uint64)
    say(f"horner L={L} variant={variant}")
    assert ia.load("server").ipir_dbg_horner_variant(ia.ptr(cts), L, ia.ptr(rg), ia.ptr(out), variant) == 0
    want = np.zeros(4 * N, np.uint64)
    ol.oracle().orc_horner(cts, L, rg, want)
    assert np.array_equal(out, want)


# ---------------------------------------------------------------- servers
GEOS = {
    "small": (256, 120, 2048, [0, 42, 100, 255]),             # G=1, no encode, no Horner
    "medium": (262144, 120, 2048, [0, 42, 1000, 100000, 262143]),  # G=4, encode D=4, Horner
}


@pytest.fixture(scope="module", params=sorted(GEOS))
def served(request):
    n, w, rows, idxs = GEOS[request.param]
    rc, geo = ol.setup(n, w, rows)
    db = np.random.default_rng(0).integers(0, P, size=(geo["R"], geo["C"]), dtype=np.uint16)
    say(f"[{request.param}] oracle preprocess G={geo['G']}")
    pc = ol.OraclePrecomp(db, geo, SEED)
    say(f"[{request.param}] GPU create")
    srv = ia.Server(db, n, w, rows, max_batch=4, crs_seed=bytes(SEED))
    yield request.param, geo, db, pc, srv, idxs
    srv.close()


def test_encoded_db_matches_oracle(served):
    name, geo, db, pc, srv, _ = served
    got = np.zeros((geo["R"], geo["C"]), np.uint16)
    assert srv.lib.ipir_dbg_db_rows(srv.h, 0, geo["R"], ia.ptr(got)) == 0
    assert np.array_equal(got, pc.db)


def test_precomp_tensors_match_oracle(served):
    name, geo, db, pc, srv, _ = served
    words = srv.lib.ipir_dbg_precomp_group_words()
    for g in range(geo["G"]):
        blk = np.zeros(words, np.uint32)
        assert srv.lib.ipir_dbg_precomp_group(srv.h, g, ia.ptr(blk)) == 0
        a = blk[:2 * N].astype(np.uint64)
        dsz = 2 * NSTEPS * 2 * N
        Dp = blk[2 * N:2 * N + dsz].astype(np.uint64)
        Dm = blk[2 * N + dsz:2 * N + 2 * dsz].astype(np.uint64)
        Df = blk[2 * N + 2 * dsz:].astype(np.uint64)
        assert np.array_equal(Dp, pc.Dp[g].reshape(-1)), f"D+ group {g}"
        assert np.array_equal(Dm, pc.Dm[g].reshape(-1)), f"D- group {g}"
        assert np.array_equal(Df, pc.Df[g].reshape(-1)), f"D_final group {g}"
        assert np.array_equal(a, pc.a[g].reshape(-1)), f"a group {g}"


def test_answer_stages_match_oracle_and_extract(served):
    name, geo, db, pc, srv, idxs = served
    for t, idx in enumerate(idxs):
        say(f"[{name}] idx={idx}")
        q, st = ol.make_query(geo, SEED, idx, 500 + t)
        interm = np.zeros(2 * geo["C"], np.uint32)
        packed = np.zeros(geo["G"] * 2 * N, np.uint32)
        resp = np.zeros(geo["T"] * 4 * N, np.uint64)
        assert srv.lib.ipir_dbg_answer_taps(srv.h, ia.ptr(q), ia.ptr(interm), ia.ptr(packed), ia.ptr(resp)) == 0
        want, w_interm, w_packed = pc.answer(q, taps=True)
        assert np.array_equal(interm.astype(np.uint64), w_interm), "DB pass"
        assert np.array_equal(packed.astype(np.uint64), w_packed), "collapse / packed b"
        assert np.array_equal(resp, want.reshape(-1)), "response"
        assert np.array_equal(ol.extract(geo, st, resp.reshape(geo["T"], -1)), ol.expected_slots(geo, db, idx))


def test_noncanonical_inputs_are_reduced(served):
    """Residues >= q on the wire are reduced mod q (out of the reference's
    contract; this library defines them)."""
    name, geo, db, pc, srv, idxs = served
    q, st = ol.make_query(geo, SEED, idxs[-1], 99)
    R = geo["R"]
    q2 = q.copy()
    q2[:R] += np.uint64(Q0)            # lwe limb0
    q2[R:2 * R] += np.uint64(5 * Q1)  # lwe limb1
    body = q2[2 * R:].reshape(6 * 2, 2, N)
    body[:, 0, :] += np.uint64(Q0)
    body[:, 1, :] += np.uint64(2 * Q1)
    a = srv.answer_batch([q])[0]
    b = srv.answer_batch([q2])[0]
    assert np.array_equal(a, b)


# ---------------------------------------------------------------- batch contract
MB8 = 32  # max_batch of the medium8 server (covers the 17..32-query MFMA passes at full width)


@pytest.fixture(scope="module")
def medium8():
    n, w, rows = 524288, 120, 2048  # G=8, D=8 (reference tests/test_gpu_batch.cu:26-37)
    rc, geo = ol.setup(n, w, rows)
    db = np.random.default_rng(42).integers(0, P, size=(geo["R"], geo["C"]), dtype=np.uint16)
    say("[medium8] GPU create")
    srv = ia.Server(db, n, w, rows, max_batch=MB8, crs_seed=bytes(SEED))
    yield geo, db, srv
    srv.close()


@pytest.mark.parametrize("path", [0, 1, -1])
def test_db_pass_paths_match_oracle(medium8, path):
    """Both database-pass kernels (fp64 per query, int8 MFMA batched) and the
    tuned dispatch, on the resident (encoded) DB, against the oracle's
    128-bit mat-vec; extreme residues q-1 and 0 in the query vectors."""
    geo, db, srv = medium8
    R, Cc = geo["R"], geo["C"]
    enc = np.zeros((R, Cc), np.uint16)
    assert srv.lib.ipir_dbg_db_rows(srv.h, 0, R, ia.ptr(enc)) == 0
    flat = enc.reshape(-1)
    rng = np.random.default_rng(100 + path)
    for nq in (1, 2, 3, 5, 8, 10, 17, 24, 32):
        b = np.zeros((nq, 2, R), np.uint64)
        b[:, 0] = rng.integers(0, Q0, (nq, R), dtype=np.uint64)
        b[:, 1] = rng.integers(0, Q1, (nq, R), dtype=np.uint64)
        b[0, 0, :R // 4] = Q0 - 1
        b[0, 1, :R // 4] = Q1 - 1
        b[-1, :, R // 2:R // 2 + 64] = 0
        got = np.zeros((nq, 2, Cc), np.uint32)
        assert srv.lib.ipir_dbg_db_pass(srv.h, ia.ptr(b), nq, path, ia.ptr(got)) == 0
        for k in range(nq):
            want = np.zeros(2 * Cc, np.uint64)
            ol.oracle().orc_matvec(flat, R, Cc, np.ascontiguousarray(b[k, 0]), np.ascontiguousarray(b[k, 1]),
                                   0, Cc, want)
            assert np.array_equal(got[k].reshape(-1).astype(np.uint64), want), f"path {path} nq {nq} query {k}"
    bad = np.full((1, 2, R), Q0, np.uint64)  # non-canonical residues are rejected
    assert srv.lib.ipir_dbg_db_pass(srv.h, ia.ptr(bad), 1, path, ia.ptr(np.zeros((1, 2, Cc), np.uint32))) == 1


@pytest.mark.parametrize("depth,form", [(3, 0), (4, 0), (5, 0), (3, 1), (4, 1), (3, 2), (5, 2), (6, 2), (7, 2),
                                        (3, 3), (4, 3), (5, 3), (6, 3), (7, 3), (3, 4), (4, 4), (5, 4), (3, 5), (4, 5),
                                        (5, 5)])
def test_many_query_ring_depths_match_oracle(medium8, depth, form):
    """Every register-ring depth and form of the batched MFMA passes against
    the oracle's mat-vec (12 queries: 8 digit tiles, depth capped at 4; 17,
    24 and 32 queries: 16 tiles, or 14 in the forms with 7 digit rows per
    query, accumulated per (digit, plane) or per power of 256; the waves of
    4- or 8-wave blocks sharing the digit tile, or the DB tile)."""
    geo, db, srv = medium8
    R, Cc = geo["R"], geo["C"]
    enc = np.zeros((R, Cc), np.uint16)
    assert srv.lib.ipir_dbg_db_rows(srv.h, 0, R, ia.ptr(enc)) == 0
    flat = enc.reshape(-1)
    rng = np.random.default_rng(400 + depth)
    name = C.create_string_buffer(128)
    assert srv.lib.ipir_dbg_dbpass_kernel(srv.h, MB8, name, 128) == 0
    # k_dbpass_mfma_sa<tiles, col tiles, 2, depth, waves, power rows> or k_dbpass_mfma_bs<16, 4, depth>
    kname = name.value.decode()
    args = [int(x) for x in kname.split("<")[1].split(">")[0].split(",")]
    if "_bs<" in kname:
        tuned = (args[2], 1)
    elif args[5]:
        tuned = (args[3], 5)
    else:
        tuned = (args[3], (3 if args[0] == 14 else 2) if args[4] == 8 else (4 if args[0] == 14 else 0))
    assert srv.lib.ipir_dbg_set_mfma_depth(srv.h, 2, 0) == 1
    assert srv.lib.ipir_dbg_set_mfma_depth(srv.h, 6, 0) == 1
    assert srv.lib.ipir_dbg_set_mfma_depth(srv.h, 5, 1) == 1
    assert srv.lib.ipir_dbg_set_mfma_depth(srv.h, 8, 2) == 1
    assert srv.lib.ipir_dbg_set_mfma_depth(srv.h, 8, 3) == 1
    assert srv.lib.ipir_dbg_set_mfma_depth(srv.h, 6, 4) == 1
    assert srv.lib.ipir_dbg_set_mfma_depth(srv.h, 6, 5) == 1
    assert srv.lib.ipir_dbg_set_mfma_depth(srv.h, 3, 6) == 1
    assert srv.lib.ipir_dbg_set_mfma_depth(srv.h, depth, form) == 0
    try:
        for nq in (12, 17, 24, MB8):
            _ring_case(srv, flat, R, Cc, rng, nq, depth)
    finally:
        assert srv.lib.ipir_dbg_set_mfma_depth(srv.h, *tuned) == 0


def _ring_case(srv, flat, R, Cc, rng, nq, depth):
    b = np.zeros((nq, 2, R), np.uint64)
    b[:, 0] = rng.integers(0, Q0, (nq, R), dtype=np.uint64)
    b[:, 1] = rng.integers(0, Q1, (nq, R), dtype=np.uint64)
    b[-1, :, :R // 8] = np.array([Q0 - 1, Q1 - 1], np.uint64)[:, None]
    got = np.zeros((nq, 2, Cc), np.uint32)
    assert srv.lib.ipir_dbg_db_pass(srv.h, ia.ptr(b), nq, 1, ia.ptr(got)) == 0
    for k in range(nq):
        want = np.zeros(2 * Cc, np.uint64)
        ol.oracle().orc_matvec(flat, R, Cc, np.ascontiguousarray(b[k, 0]), np.ascontiguousarray(b[k, 1]),
                               0, Cc, want)
        assert np.array_equal(got[k].reshape(-1).astype(np.uint64), want), f"depth {depth} nq {nq} query {k}"


def test_batch_equals_singles_and_extracts(medium8):
    geo, db, srv = medium8
    idxs = [0, 42, 1000, 100000, 524287] * 4  # 20: a 17..32-query DB pass, two collapse passes
    qs, sts = [], []
    for t, idx in enumerate(idxs):
        q, st = ol.make_query(geo, SEED, idx, 7000 + t)
        qs.append(q)
        sts.append(st)
    singles = [srv.answer_batch([q])[0] for q in qs]
    batch = srv.answer_batch(qs)
    for i in range(len(qs)):
        assert np.array_equal(batch[i], singles[i]), f"batch[{i}] != single"
        assert np.array_equal(ol.extract(geo, sts[i], batch[i].reshape(geo["T"], -1)),
                              ol.expected_slots(geo, db, idxs[i]))
    again = srv.answer_batch(qs)
    assert np.array_equal(again, batch), "rerun not reproducible"
    for k in (3, 2):  # partial batches
        part = srv.answer_batch(qs[:k])
        assert np.array_equal(part, batch[:k])


def test_ragged_batch_over_sub_batches(medium8):
    """A batch of 33 on a server of max_batch 40: a 32-query sub-batch then a
    one-query sub-batch, each with its own pipeline shape and Horner kernel,
    queued on the same streams, over equal chunks or with the lowest chunk
    split; equals the serial answer (both sub-batches' chains in one launch
    after them) and extracts. A serial 32-query sub-batch followed by a
    chunked or streamed one-query one launches its chains before the
    latter's."""
    geo, db, _ = medium8
    n, w, rows = 524288, 120, 2048
    srv = ia.Server(db, n, w, rows, max_batch=40, crs_seed=bytes(SEED))
    try:
        idxs = [(37 * k + 5) % n for k in range(33)]
        qs, sts = [], []
        for t, idx in enumerate(idxs):
            q, st = ol.make_query(geo, SEED, idx, 9000 + t)
            qs.append(q)
            sts.append(st)
        assert srv.lib.ipir_dbg_set_group_chunk(srv.h, geo["G"]) == 0
        assert srv.lib.ipir_dbg_set_pipeline(srv.h, 0, 0) == 0
        serial = srv.answer_batch(qs)
        # one chunk of all groups: the collapse beside the whole pass (3)
        for shapes in ((3, 3), (3, 0), (0, 3)):
            assert srv.lib.ipir_dbg_set_pipeline(srv.h, *shapes) == 0
            assert np.array_equal(srv.answer_batch(qs), serial), f"one chunk, shapes {shapes}"
        # chunks of 2; of 4 with the lowest split 2+1+1 (bit 0), the highest
        # 1+1+2 (bit 1), both, or neither
        for groups, split in ((2, 1), (4, 1), (4, 0), (4, 2), (4, 3)):
            assert srv.lib.ipir_dbg_set_group_chunk(srv.h, groups) == 0
            assert srv.lib.ipir_dbg_set_tail_split(srv.h, split) == 0
            for shapes in ((3, 2), (2, 3), (3, 3), (4, 4), (4, 2), (5, 5), (3, 5), (3, 0)):
                assert srv.lib.ipir_dbg_set_pipeline(srv.h, *shapes) == 0
                got = srv.answer_batch(qs)
                assert np.array_equal(got, serial), f"groups {groups} split {split} shapes {shapes}"
        # streamed shapes: each sub-batch with its own set of per-group
        # completion words, the one-query one behind the 32-query one on the
        # same streams; mixed with a chunked shape
        assert srv.lib.ipir_dbg_take_streaming(srv.h) == 0
        assert srv.lib.ipir_dbg_set_group_chunk(srv.h, 2) == 0
        for shapes in ((6, 6), (7, 7), (6, 3), (3, 7), (8, 8), (8, 2), (6, 0)):
            assert srv.lib.ipir_dbg_set_pipeline(srv.h, *shapes) == 0
            for sl in (1, 4):
                assert srv.lib.ipir_dbg_set_stream_slices(srv.h, 1, sl) == 0
                assert srv.lib.ipir_dbg_set_stream_slices(srv.h, 32, sl) == 0
                n0 = srv.lib.ipir_dbg_streamed_count(srv.h)
                got = srv.answer_batch(qs)
                assert np.array_equal(got, serial), f"streamed shapes {shapes} slices {sl}"
                assert srv.lib.ipir_dbg_streamed_count(srv.h) - n0 == (shapes[0] >= 6) + (shapes[1] >= 6)
        # both sub-batches captured into one graph, replayed with new inputs
        # (a captured answer takes the chunked shapes instead of the streamed ones)
        assert srv.lib.ipir_dbg_set_pipeline(srv.h, 7, 7) == 0
        assert srv.lib.ipir_dbg_set_graph(srv.h, 1) == 0
        try:
            for rep in range(3):
                order = qs if rep % 2 == 0 else qs[::-1]
                want = serial if rep % 2 == 0 else serial[::-1]
                assert np.array_equal(srv.answer_batch(order), want), f"graph replay {rep}"
        finally:
            assert srv.lib.ipir_dbg_set_graph(srv.h, 0) == 0
        for i in (0, 31, 32):
            assert np.array_equal(ol.extract(geo, sts[i], got[i].reshape(geo["T"], -1)),
                                  ol.expected_slots(geo, db, idxs[i]))
    finally:
        srv.close()


def test_collapse_variants_equal(medium8):
    """Every form of the 17..32-query collapse pass (two-step prefetch;
    load rings of depth 3..5 in 4- or 8-wave blocks) gives the same
    responses for a 20-query batch, and those extract."""
    geo, db, srv = medium8
    idxs = [int(i) for i in np.random.default_rng(77).integers(0, 524288, 20)]
    qs, sts = zip(*[ol.make_query(geo, SEED, idx, 9100 + k) for k, idx in enumerate(idxs)])
    tuned = srv.lib.ipir_dbg_get_collapse_variant(srv.h)
    assert 0 <= tuned < 6
    assert srv.lib.ipir_dbg_set_collapse_variant(srv.h, 6) == 1
    try:
        outs = []
        for v in range(6):
            assert srv.lib.ipir_dbg_set_collapse_variant(srv.h, v) == 0
            outs.append(srv.answer_batch(list(qs)))
    finally:
        assert srv.lib.ipir_dbg_set_collapse_variant(srv.h, tuned) == 0
    for v in range(1, 6):
        assert np.array_equal(outs[v], outs[0]), f"collapse variant {v}"
    for k in (0, 19):
        assert np.array_equal(ol.extract(geo, sts[k], outs[0][k].reshape(geo["T"], -1)),
                              ol.expected_slots(geo, db, idxs[k]))


def test_pair_horner_answers_equal_single(medium8):
    """The workgroup-pair Horner gives the same responses as the one-workgroup
    kernel, alone and batched (Horner chain of 7 steps, 10 chains)."""
    geo, db, srv = medium8
    qs = [ol.make_query(geo, SEED, idx, 300 + k)[0] for k, idx in enumerate(range(0, 524288, 52429))]
    assert srv.lib.ipir_dbg_set_horner(srv.h, 0) == 0
    try:
        base = srv.answer_batch(qs)
        one_single = srv.answer_batch(qs[:1])
        assert srv.lib.ipir_dbg_set_horner(srv.h, 1) == 0
        pair = srv.answer_batch(qs)
        one = srv.answer_batch(qs[:1])
    finally:
        assert srv.lib.ipir_dbg_set_horner(srv.h, -1) == 0
    assert np.array_equal(pair, base)
    assert np.array_equal(one[0], base[0]) and np.array_equal(one_single[0], base[0])


@pytest.mark.parametrize("groups,shape,slices", [(1, 2, 0), (2, 2, 0), (4, 2, 0), (1, 3, 0), (2, 3, 0),
                                                   (2, 2, 1), (4, 2, 2), (1, 4, 0), (2, 4, 0), (2, 4, 2),
                                                   (1, 5, 0), (2, 5, 1)])
def test_pipelined_group_chunks_equal_unchunked(medium8, groups, shape, slices):
    """The answer pipelined over group chunks (Horner chains on a second
    stream, resuming between launches; shape 3 also the collapse on a third,
    5 with that stream a chunk ahead; shape 4 the collapse of all groups
    before the chunks)
    equals the serial one-chunk answer, for both Horner kernels and both
    pack kernels, one query and three; chunk database passes with the
    narrowing rule's row slices or a set count (one: straight into the
    chunk's columns)."""
    geo, db, srv = medium8
    qs = [ol.make_query(geo, SEED, idx, 800 + k)[0] for k, idx in enumerate([3, 99999, 524287])]
    base = srv.answer_batch(qs)
    one = srv.answer_batch(qs[:1])
    assert srv.lib.ipir_dbg_set_pipeline(srv.h, 1, shape) == 1
    assert srv.lib.ipir_dbg_set_group_chunk(srv.h, groups) == 0
    assert srv.lib.ipir_dbg_set_pipeline(srv.h, shape, shape) == 0
    assert srv.lib.ipir_dbg_set_chunk_slices(srv.h, 3, slices) == 0
    assert srv.lib.ipir_dbg_set_chunk_slices(srv.h, 1, slices) == 0
    try:
        got = srv.answer_batch(qs)
        got1 = srv.answer_batch(qs[:1])
        # the chunks' one-workgroup pack kernel: for three queries, and for
        # one query alone
        assert srv.lib.ipir_dbg_set_pack(srv.h, 3) == 1
        assert srv.lib.ipir_dbg_set_pack(srv.h, 1) == 0
        got_wg = srv.answer_batch(qs)
        assert srv.lib.ipir_dbg_set_pack(srv.h, -1) == 0
        got1_wg = srv.answer_batch(qs[:1])  # -1: one workgroup for one query
        assert srv.lib.ipir_dbg_set_pack(srv.h, 2) == 0  # fused into the Horner chains
        got_fused = srv.answer_batch(qs)
        got1_fused = srv.answer_batch(qs[:1])
        assert srv.lib.ipir_dbg_set_pack(srv.h, 0) == 0
        assert srv.lib.ipir_dbg_set_horner(srv.h, 1) == 0
        pair = srv.answer_batch(qs)
    finally:
        assert srv.lib.ipir_dbg_set_pack(srv.h, 0) == 0
        assert srv.lib.ipir_dbg_set_horner(srv.h, -1) == 0
        assert srv.lib.ipir_dbg_set_pipeline(srv.h, 3, 2) == 0
        assert srv.lib.ipir_dbg_set_chunk_slices(srv.h, 3, 0) == 0
        assert srv.lib.ipir_dbg_set_chunk_slices(srv.h, 1, 0) == 0
        assert srv.lib.ipir_dbg_set_group_chunk(srv.h, geo["G"]) == 0
    assert np.array_equal(got, base) and np.array_equal(got1[0], one[0]) and np.array_equal(pair, base)
    assert np.array_equal(got_wg, base) and np.array_equal(got1_wg[0], one[0])
    assert np.array_equal(got_fused, base) and np.array_equal(got1_fused[0], one[0])


@pytest.mark.parametrize("shape", [6, 7, 8])
def test_streamed_pass_equals_serial(medium8, shape):
    """One full-width database pass that hands each group over as soon as
    its columns are final, a pack kernel packing each group as it arrives
    and the Horner chains waiting for it inside the kernel; after the collapse of all
    groups (6), with the collapse beside (7) or inside the pass's launch
    group by group (8; 20 queries then take 6); at 1 to 16 row slices
    (one: straight into interm; more: the pack sums the partials), in one
    launch or split in two (the chains released after the first). One
    query, three, sixteen and seventeen (the pack's 2 queries per workgroup
    up to 16, filling the flag sets; 4 from 17) and twenty (the power-row
    pass) equal the serial answers, host-staged and through the device
    entry point."""
    geo, db, srv = medium8
    lib = srv.lib
    idxs = [int(i) for i in np.random.default_rng(606).integers(0, 524288, 20)]
    qs = [ol.make_query(geo, SEED, idx, 6100 + k)[0] for k, idx in enumerate(idxs)]
    assert lib.ipir_dbg_take_streaming(srv.h) == 0
    assert lib.ipir_dbg_stream_ok(srv.h) == 1
    assert lib.ipir_dbg_set_stream_slices(srv.h, 1, 3) == 1  # not a power of two
    assert lib.ipir_dbg_set_pipeline(srv.h, 0, 0) == 0
    try:
        base = {k: srv.answer_batch(qs[:k]) for k in (1, 3, 16, 17, 20)}
        assert lib.ipir_dbg_set_pipeline(srv.h, shape, shape) == 0
        assert lib.ipir_dbg_set_stream_split(srv.h, 1, 8) == 1
        for sl, split in ((1, 0), (2, 0), (4, 0), (8, 0), (16, 0), (2, 3), (8, 4), (1, 7)):
            assert lib.ipir_dbg_set_stream_slices(srv.h, 1, sl) == 0
            assert lib.ipir_dbg_set_stream_slices(srv.h, 20, sl) == 0
            assert lib.ipir_dbg_set_stream_split(srv.h, 1, split) == 0
            assert lib.ipir_dbg_set_stream_split(srv.h, 20, split) == 0
            n0 = lib.ipir_dbg_streamed_count(srv.h)
            for k in (1, 3, 16, 17, 20):
                say(f"[streamed] shape={shape} slices={sl} split={split} queries={k}")
                assert np.array_equal(srv.answer_batch(qs[:k]), base[k]), f"slices {sl} split {split}, {k} queries"
            assert lib.ipir_dbg_streamed_count(srv.h) == n0 + 5, "every answer streamed"
        flat = np.ascontiguousarray(np.stack(qs))
        dq = ia.DeviceBuffer(flat.nbytes, lib)
        dq.upload(flat)
        dout = ia.DeviceBuffer(base[20].nbytes, lib)
        for _ in range(2):
            assert lib.ipir_answer_batch_device(srv.h, C.c_void_p(dq.p), 20, C.c_void_p(dout.p), None) == 0
        assert lib.ipir_dev_sync() == 0
        assert np.array_equal(dout.download(np.zeros_like(base[20])), base[20]), "device entry point"
        dq.free()
        dout.free()
    finally:
        assert lib.ipir_dbg_set_stream_slices(srv.h, 1, 4) == 0
        assert lib.ipir_dbg_set_stream_slices(srv.h, 20, 4) == 0
        assert lib.ipir_dbg_set_stream_split(srv.h, 1, 0) == 0
        assert lib.ipir_dbg_set_stream_split(srv.h, 20, 0) == 0
        assert lib.ipir_dbg_set_pipeline(srv.h, 3, 2) == 0


def test_streamed_answers_back_to_back(medium8):
    """Streamed answers queued without a synchronization in between, so
    each call's flag set is cleared and reused while the previous call's
    pass, pack and chains may still run: six asynchronous device calls
    alternate two query sets (1 and 5 queries, 6 and 8 shapes, 1 and 4 row
    slices) into separate output buffers; every response equals the serial
    answer."""
    geo, db, srv = medium8
    lib = srv.lib
    rng = np.random.default_rng(808)
    sets = [[ol.make_query(geo, SEED, int(i), 7300 + 10 * k + j)[0] for j, i in enumerate(rng.integers(0, 524288, n))]
            for k, n in enumerate((1, 5))]
    assert lib.ipir_dbg_take_streaming(srv.h) == 0
    assert lib.ipir_dbg_set_pipeline(srv.h, 0, 0) == 0
    try:
        want = [srv.answer_batch(qs) for qs in sets]
        dq = [ia.DeviceBuffer(np.stack(qs).nbytes, lib) for qs in sets]
        for b, qs in zip(dq, sets):
            b.upload(np.ascontiguousarray(np.stack(qs)))
        for shapes, sl in (((6, 6), 1), ((8, 8), 4), ((8, 6), 2)):
            assert lib.ipir_dbg_set_pipeline(srv.h, *shapes) == 0
            assert lib.ipir_dbg_set_stream_slices(srv.h, 1, sl) == 0
            assert lib.ipir_dbg_set_stream_slices(srv.h, 5, sl) == 0
            outs = [ia.DeviceBuffer(want[k % 2].nbytes, lib) for k in range(6)]
            n0 = lib.ipir_dbg_streamed_count(srv.h)
            for k in range(6):
                say(f"[back-to-back] shapes={shapes} slices={sl} call={k}")
                assert lib.ipir_answer_batch_device(srv.h, C.c_void_p(dq[k % 2].p), len(sets[k % 2]),
                                                    C.c_void_p(outs[k].p), None) == 0
            assert lib.ipir_dev_sync() == 0
            assert lib.ipir_device_status(srv.h) == 0
            assert lib.ipir_dbg_streamed_count(srv.h) - n0 == 6
            for k in range(6):
                got = outs[k].download(np.zeros_like(want[k % 2]))
                assert np.array_equal(got, want[k % 2]), f"shapes {shapes} slices {sl} call {k}"
                outs[k].free()
        for b in dq:
            b.free()
    finally:
        assert lib.ipir_dbg_set_stream_slices(srv.h, 1, 4) == 0
        assert lib.ipir_dbg_set_stream_slices(srv.h, 5, 4) == 0
        assert lib.ipir_dbg_set_pipeline(srv.h, 3, 2) == 0


def test_streamed_wait_overrun_reported(medium8):
    """The streamed chains' bounded wait, forced to give up at once (spin
    limit 0; the pass and every kernel still end normally): the blocking
    call returns 4; the asynchronous device entry point returns 0 with
    all-ones response words (no residue has that value) and
    ipir_device_status then reports 4 once; the serving loop reports 4 for
    its request; once the bound is restored the next call is not blamed
    for the earlier failures."""
    geo, db, srv = medium8
    lib = srv.lib
    q, st = ol.make_query(geo, SEED, 777, 41)
    assert lib.ipir_dbg_take_streaming(srv.h) == 0
    assert lib.ipir_dbg_set_pipeline(srv.h, 0, 0) == 0
    want = srv.answer_batch([q])[0]
    out = np.zeros((1, geo["T"] * 4 * N), np.uint64)
    views = srv.views([q])
    assert lib.ipir_dbg_set_pipeline(srv.h, 7, 7) == 0
    assert lib.ipir_dbg_set_stream_spin_limit(srv.h, 0) == 0
    try:
        assert lib.ipir_answer_batch(srv.h, views, 1, ia.ptr(out)) == 4
        dq = ia.DeviceBuffer(q.nbytes, lib)
        dq.upload(q)
        dout = ia.DeviceBuffer(want.nbytes, lib)
        assert lib.ipir_device_status(srv.h) == 0, "nothing to report yet"
        assert lib.ipir_answer_batch_device(srv.h, C.c_void_p(dq.p), 1, C.c_void_p(dout.p), None) == 0
        assert lib.ipir_device_status(srv.h) == 4, "the asynchronous call's overrun is reported"
        assert lib.ipir_device_status(srv.h) == 0, "and then cleared"
        got = dout.download(np.zeros_like(want))
        assert np.all(got == np.uint64(2**64 - 1)), "invalid responses are all-ones"
        dq.free()
        dout.free()
        loop = lib.ipir_loop_start(srv.h, 2)
        assert loop
        try:
            o2 = np.zeros(geo["T"] * 4 * N, np.uint64)
            t = lib.ipir_loop_submit(loop, views, ia.ptr(o2))
            assert lib.ipir_loop_wait(loop, t, None) == 4
        finally:
            lib.ipir_loop_stop(loop)
    finally:
        assert lib.ipir_dbg_set_stream_spin_limit(srv.h, 1 << 23) == 0
    try:
        assert lib.ipir_answer_batch(srv.h, views, 1, ia.ptr(out)) == 0, "a later call is not blamed"
        assert np.array_equal(out[0], want)
    finally:
        assert lib.ipir_dbg_set_pipeline(srv.h, 3, 2) == 0


def test_late_chunk_slices_equal_unchunked(medium8):
    """Chunks from a given index on with their own row-slice count (one
    query and three, both pipeline shapes with the collapse beside) give the
    serial answer."""
    geo, db, srv = medium8
    qs = [ol.make_query(geo, SEED, idx, 850 + k)[0] for k, idx in enumerate([11, 262144, 524000])]
    base = srv.answer_batch(qs)
    one = srv.answer_batch(qs[:1])
    assert srv.lib.ipir_dbg_set_late_slices(srv.h, 3, 1, 3) == 1  # not a power of two
    assert srv.lib.ipir_dbg_set_group_chunk(srv.h, 2) == 0
    try:
        for shape in (3, 2):
            assert srv.lib.ipir_dbg_set_pipeline(srv.h, shape, shape) == 0
            for frm, sl in ((1, 2), (2, 4), (0, 1)):
                assert srv.lib.ipir_dbg_set_late_slices(srv.h, 3, frm, sl) == 0
                assert srv.lib.ipir_dbg_set_late_slices(srv.h, 1, frm, sl) == 0
                assert np.array_equal(srv.answer_batch(qs), base), f"shape {shape} late {frm},{sl}"
                assert np.array_equal(srv.answer_batch(qs[:1])[0], one[0]), f"one query, shape {shape} late {frm},{sl}"
    finally:
        assert srv.lib.ipir_dbg_set_late_slices(srv.h, 3, 0, 0) == 0
        assert srv.lib.ipir_dbg_set_late_slices(srv.h, 1, 0, 0) == 0
        assert srv.lib.ipir_dbg_set_pipeline(srv.h, 3, 2) == 0
        assert srv.lib.ipir_dbg_set_group_chunk(srv.h, geo["G"]) == 0


def test_packed_wire_answers_match_host_codecs(medium8):
    """Queries in the 53-bit packed wire form, unpacked on the device, and
    responses compressed on the device give the bytes of the client
    library's host codecs (pack, flat answer, compress), extract to their
    entries, and the device-buffer entry point agrees."""
    geo, db, srv = medium8
    cl = ia.Client(srv.params, srv.lib)
    idxs = [0, 77777, 524287, 4242]
    hs, qs = [], []
    for k, idx in enumerate(idxs):
        h, q = cl.build(idx, 900 + k)
        hs.append(h)
        qs.append(q)
    packed = np.stack([cl.pack_query(q) for q in qs])
    assert packed.shape[1] == srv.lib.ipir_query_packed_bytes(srv.params.ref())
    got = srv.answer_packed(packed)
    flat = srv.answer_batch(qs)
    for k, idx in enumerate(idxs):
        assert np.array_equal(got[k], cl.compress(flat[k])), f"query {k}"
        assert np.array_equal(cl.extract_compressed(hs[k], got[k]), ol.expected_slots(geo, db, idx))
    assert np.array_equal(srv.answer_packed(packed[2:3])[0], got[2])  # alone equals batched
    dq = ia.DeviceBuffer(packed.nbytes, srv.lib)
    dq.upload(packed)
    dout = ia.DeviceBuffer(got.nbytes, srv.lib)
    assert srv.lib.ipir_answer_batch_packed_device(srv.h, C.c_void_p(dq.p), len(idxs), C.c_void_p(dout.p),
                                                   None) == 0
    srv.lib.ipir_dev_sync()
    assert np.array_equal(dout.download(np.zeros_like(got)), got)
    out = np.zeros((MB8 + 1, 12288), np.uint8)
    big = np.zeros((MB8 + 1, packed.shape[1]), np.uint8)
    assert srv.lib.ipir_answer_batch_packed(srv.h, ia.ptr(big), MB8 + 1, ia.ptr(out)) == 2
    assert srv.lib.ipir_answer_batch_packed(srv.h, None, 1, ia.ptr(out)) == 1
    assert srv.lib.ipir_answer_batch_packed(srv.h, ia.ptr(big), 0, ia.ptr(out)) == 0
    for h in hs:
        cl.destroy(h)


def test_generation_rollover_while_serving(medium8):
    """A database generation is one server (reference src/gpu_protocol.h:
    11-15, docs/ARCHITECTURE.md:292-297): the next generation (a new DB) is
    preprocessed on another host thread while the current one keeps
    answering, admitted because device_free_bytes holds another resident
    footprint; both answer their own DB side by side, and freeing one
    generation leaves the other answering."""
    import threading
    geo, db, srv = medium8
    n, w, rows = 524288, 120, 2048
    caps = srv.caps()
    assert caps.device_free_bytes > 2 * caps.resident_bytes
    db2 = np.random.default_rng(43).integers(0, P, size=(geo["R"], geo["C"]), dtype=np.uint16)
    box = {}

    def build():
        box["srv"] = ia.Server(db2, n, w, rows, max_batch=4, crs_seed=bytes(SEED))
    th = threading.Thread(target=build)
    th.start()
    served = 0
    while th.is_alive() or served == 0:
        q, st = ol.make_query(geo, SEED, 1000 + served, 5000 + served)
        r = srv.answer_batch([q])[0]
        assert np.array_equal(ol.extract(geo, st, r.reshape(geo["T"], -1)), ol.expected_slots(geo, db, 1000 + served))
        served += 1
    th.join()
    new = box["srv"]
    try:
        for k, idx in enumerate([7, 300000]):
            q, st = ol.make_query(geo, SEED, idx, 6000 + k)
            r_new = new.answer_batch([q])[0]
            assert np.array_equal(ol.extract(geo, st, r_new.reshape(geo["T"], -1)), ol.expected_slots(geo, db2, idx))
            r_old = srv.answer_batch([q])[0]
            assert np.array_equal(ol.extract(geo, st, r_old.reshape(geo["T"], -1)), ol.expected_slots(geo, db, idx))
        say(f"[rollover] {served} answers served on the old generation while the new one was built")
    finally:
        new.close()
    q, st = ol.make_query(geo, SEED, 424242, 6100)
    r = srv.answer_batch([q])[0]
    assert np.array_equal(ol.extract(geo, st, r.reshape(geo["T"], -1)), ol.expected_slots(geo, db, 424242))


@pytest.fixture(scope="module")
def tall():
    """2^16 DB rows (one group) with a workgroup target of 1 and no
    autotuning: the plans then cut rows into slices only where the exactness
    limits require it (one slice of 2^16 rows for the 16-tile forms, two of
    2^15 for the power-row forms)."""
    import os
    n, w, rows = 2 ** 21, 120, 65536
    rc, geo = ol.setup(n, w, rows)
    assert rc == 0 and geo["R"] == 65536 and geo["C"] == 2048
    saved = {k: os.environ.get(k) for k in ("INSPIRE_AUTOTUNE", "INSPIRE_MFMA_TARGET")}
    os.environ["INSPIRE_AUTOTUNE"] = "0"
    os.environ["INSPIRE_MFMA_TARGET"] = "1"
    try:
        say("[tall] GPU create")
        zero = ia.Server(np.zeros((geo["R"], geo["C"]), np.uint16), n, w, rows, max_batch=32,
                         crs_seed=bytes(SEED))
        db = np.random.default_rng(44).integers(0, P, size=(geo["R"], geo["C"]), dtype=np.uint16)
        rnd = ia.Server(db, n, w, rows, max_batch=32, crs_seed=bytes(SEED))
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    yield geo, zero, rnd
    zero.close()
    rnd.close()


@pytest.mark.parametrize("form", [0, 1, 2, 3, 4, 5])
def test_slice_limits_hold_at_extreme_values(tall, form):
    """Largest dot products the int8 passes can meet: an all-zero DB (every
    plane byte reads as -128) against query rows whose CRT value has six
    balanced digits -128 and top digit 32 (each power-row sum grows by
    2 x 2^14 per row: 2^31 over 2^16 rows, past i32), or whose per-limb
    residues have digits -128, -128, -128, 5 (each per-plane sum grows by
    2^14 per row). The exact product is 0, so a slice past its limit shows
    as a nonzero residue. Then a random DB against the oracle's mat-vec."""
    geo, zero, rnd = tall
    R, Cc = geo["R"], geo["C"]
    B = 2 ** 53 - 0x808080808080  # CRT value, balanced digits: six times -128, then 32
    b4 = 5 * 2 ** 24 - 0x808080   # residue, balanced digits: three times -128, then 5
    exts = [np.array([B % Q0, B % Q1], np.uint64), np.array([b4, b4], np.uint64)]
    for srv in (zero, rnd):
        assert srv.lib.ipir_dbg_set_mfma_depth(srv.h, 3, form) == 0
    for nq in (16, 32):
        name = C.create_string_buffer(128)
        assert zero.lib.ipir_dbg_dbpass_kernel(zero.h, nq, name, 128) == 0
        slices = int(name.value.decode().split("slices=")[1])
        assert slices == (2 if form == 5 else 1), name.value  # the limits, not grid filling, cut the rows
        for ext in exts:
            b = np.broadcast_to(ext[None, :, None], (nq, 2, R)).copy()
            got = np.ones((nq, 2, Cc), np.uint32)
            assert zero.lib.ipir_dbg_db_pass(zero.h, ia.ptr(b), nq, 1, ia.ptr(got)) == 0
            assert not got.any(), f"form {form} nq {nq}: {np.count_nonzero(got)} nonzero residues"
    ext = exts[0]
    enc = np.zeros((R, Cc), np.uint16)
    assert rnd.lib.ipir_dbg_db_rows(rnd.h, 0, R, ia.ptr(enc)) == 0
    rng = np.random.default_rng(500 + form)
    nq = 17
    b = np.zeros((nq, 2, R), np.uint64)
    b[:, 0] = rng.integers(0, Q0, (nq, R), dtype=np.uint64)
    b[:, 1] = rng.integers(0, Q1, (nq, R), dtype=np.uint64)
    b[-1] = ext[:, None]
    got = np.zeros((nq, 2, Cc), np.uint32)
    assert rnd.lib.ipir_dbg_db_pass(rnd.h, ia.ptr(b), nq, 1, ia.ptr(got)) == 0
    for k in (0, nq - 1):
        want = np.zeros(2 * Cc, np.uint64)
        ol.oracle().orc_matvec(enc.reshape(-1), R, Cc, np.ascontiguousarray(b[k, 0]), np.ascontiguousarray(b[k, 1]),
                               0, Cc, want)
        assert np.array_equal(got[k].reshape(-1).astype(np.uint64), want), f"form {form} query {k}"


def test_batch_validation_codes(medium8):
    geo, db, srv = medium8
    lib = srv.lib
    q, _ = ol.make_query(geo, SEED, 1, 1)
    views = srv.views([q] * (MB8 + 1))
    out = np.zeros((MB8 + 1, geo["T"] * 4 * N), np.uint64)
    assert lib.ipir_answer_batch(srv.h, views, MB8 + 1, ia.ptr(out)) == 2   # over max_batch
    assert lib.ipir_answer_batch(srv.h, views, 0, ia.ptr(out)) == 0    # empty batch
    assert lib.ipir_answer_batch(None, views, 1, ia.ptr(out)) == 1
    bad = srv.views([q])
    bad[0].rgsw = None
    assert lib.ipir_answer_batch(srv.h, bad, 1, ia.ptr(out)) == 3
    assert not out.any(), "nothing written on validation failure"


def test_capi_round_trip_with_product_client(medium8):
    geo, db, srv = medium8
    p = srv.params
    assert p.seed == bytes(SEED), "pinned CRS seed honoured"
    assert (p.ring_n, p.d_eff, p.db_rows, p.db_cols, p.n_packed) == (N, 2, geo["R"], geo["C"], geo["G"])
    caps = srv.caps()
    assert caps.max_batch == MB8 and caps.num_cts == geo["T"] and caps.resident_bytes > 0
    assert caps.device_free_bytes > 0
    cl = ia.Client(p, ia.load("server"))
    hs, qs, idxs = [], [], [5, 77777, 524000]
    for idx in idxs:
        h, q = cl.build(idx)  # fresh secret (random_device)
        hs.append(h)
        qs.append(q)
    resp = srv.answer_batch(qs)
    for h, r, idx in zip(hs, resp, idxs):
        assert np.array_equal(cl.extract(h, r), ol.expected_slots(geo, db, idx))
        cl.destroy(h)


def test_device_resident_entry_matches_host_path(medium8):
    geo, db, srv = medium8
    qs = [ol.make_query(geo, SEED, idx, 40 + idx)[0] for idx in (3, 300000, 524287)]
    host = srv.answer_batch(qs)
    flat = np.ascontiguousarray(np.stack(qs))
    dq = ia.DeviceBuffer(flat.nbytes)
    dout = ia.DeviceBuffer(host.nbytes)
    dq.upload(flat)
    assert srv.lib.ipir_answer_batch_device(srv.h, C.c_void_p(dq.p), 3, C.c_void_p(dout.p), None) == 0
    assert srv.lib.ipir_dev_sync() == 0
    got = dout.download(np.zeros_like(host))
    assert np.array_equal(got, host)
    dq.free()
    dout.free()


@pytest.mark.parametrize("groups,shape", [(8, 0), (2, 2), (2, 3), (2, 4), (2, 5), (8, 6), (2, 7), (8, 8)])
def test_graph_replays_equal_eager(medium8, groups, shape):
    """Answers replayed from captured hipGraphs (serial, and pipelined over
    group chunks on two and three streams) equal the launch-by-launch ones,
    on every replay with new queries in the same buffers, with eager work and
    synchronizations between the replays, for one query, three, twenty (a
    17..32-query pass), the device-buffer entry and the packed wire form."""
    geo, db, srv = medium8
    cl = ia.Client(srv.params, srv.lib)
    sets = {k: [[ol.make_query(geo, SEED, (97 * r + 131071 * j) % 524288, 5000 + 100 * r + j)[0]
                 for j in range(k)] for r in range(3)] for k in (1, 3, 20)}
    pset = [np.stack([cl.pack_query(q) for q in sets[3][r]]) for r in range(3)]
    if shape >= 6:  # eager answers stream; captured ones take the chunked (or serial) shape
        assert srv.lib.ipir_dbg_take_streaming(srv.h) == 0
    assert srv.lib.ipir_dbg_set_group_chunk(srv.h, groups) == 0
    assert srv.lib.ipir_dbg_set_pipeline(srv.h, shape, shape) == 0
    try:
        eager = {k: [srv.answer_batch(v) for v in sets[k]] for k in sets}
        peager = [srv.answer_packed(x) for x in pset]
        assert srv.lib.ipir_dbg_set_graph(srv.h, 2) == 1
        assert srv.lib.ipir_dbg_set_graph(srv.h, 1) == 0
        src = np.arange(2 * 2048, dtype=np.uint32)  # one polynomial, two limbs of N = 2048
        ntt_out = np.zeros_like(src)
        flat = np.ascontiguousarray(np.stack(sets[3][0]))
        dq = ia.DeviceBuffer(flat.nbytes, srv.lib)
        dout = ia.DeviceBuffer(eager[3][0].nbytes, srv.lib)
        for rep in range(2):
            for r in range(3):
                say(f"[graph] groups={groups} shape={shape} rep={rep} set={r}")
                for k in sets:
                    assert np.array_equal(srv.answer_batch(sets[k][r]), eager[k][r]), f"count {k} set {r}"
                    # eager device work and a synchronization between replays
                    assert srv.lib.ipir_dbg_ntt(ia.ptr(src), ia.ptr(ntt_out), 1, 0) == 0
                assert np.array_equal(srv.answer_packed(pset[r]), peager[r]), f"packed set {r}"
                dq.upload(np.ascontiguousarray(np.stack(sets[3][r])))
                assert srv.lib.ipir_answer_batch_device(srv.h, C.c_void_p(dq.p), 3, C.c_void_p(dout.p), None) == 0
                assert srv.lib.ipir_dev_sync() == 0
                assert np.array_equal(dout.download(np.zeros_like(eager[3][r])), eager[3][r]), f"device set {r}"
        dq.free()
        dout.free()
    finally:
        assert srv.lib.ipir_dbg_set_graph(srv.h, 0) == 0
        assert srv.lib.ipir_dbg_set_pipeline(srv.h, 3, 2) == 0
        assert srv.lib.ipir_dbg_set_group_chunk(srv.h, geo["G"]) == 0
    assert np.array_equal(srv.answer_batch(sets[3][1]), eager[3][1])


def _loop_stats(lib, loop):
    st = ia.ipir_loop_stats()
    assert lib.ipir_loop_get_stats(loop, C.byref(st)) == 0
    return st


def _wait_until(pred, what, seconds=30.0):
    import time
    t0 = time.time()
    while not pred():
        assert time.time() - t0 < seconds, f"timed out waiting for {what}"
        time.sleep(0.002)


def test_serving_loop_admits_while_batches_in_flight(medium8):
    """Continuous batching, deterministically: with the loop's partitions
    held busy after their device work, a first request occupies partition
    0; a second wave submitted while it is in flight is admitted into
    partition 1 and launched; a third waits for a free partition. Once
    released, every result equals the same query answered alone and
    extracts to its entry. While the loop runs, the server's other entry
    points answer 5 (busy) and a second loop cannot start; after it stops
    they work again."""
    geo, db, srv = medium8
    lib = srv.lib
    n = 524288
    idxs = [(104729 * k + 17) % n for k in range(9)]
    built = [ol.make_query(geo, SEED, idx, 900 + k) for k, idx in enumerate(idxs)]
    qs = [q for q, _ in built]
    alone = [srv.answer_batch([q])[0] for q in qs]
    loop = lib.ipir_loop_start(srv.h, 2)  # two partitions of 16 slots
    assert loop
    views = [srv.views([q]) for q in qs]
    outs = [np.zeros(geo["T"] * 4 * N, np.uint64) for _ in qs]
    tickets = []
    try:
        assert lib.ipir_loop_start(srv.h, 2) is None, "one loop per server"
        out = np.zeros((1, geo["T"] * 4 * N), np.uint64)
        assert lib.ipir_answer_batch(srv.h, views[0], 1, ia.ptr(out)) == 5
        dq = ia.DeviceBuffer(qs[0].nbytes, lib)
        assert lib.ipir_answer_batch_device(srv.h, C.c_void_p(dq.p), 1, C.c_void_p(dq.p), None) == 5
        assert lib.ipir_dbg_set_pipeline(srv.h, 3, 2) == 5
        assert lib.ipir_dbg_loop_hold(loop, 1) == 0
        tickets.append(lib.ipir_loop_submit(loop, views[0], ia.ptr(outs[0])))
        _wait_until(lambda: _loop_stats(lib, loop).batches == 1, "the first batch")
        for k in range(1, 6):  # second wave while partition 0 is in flight
            tickets.append(lib.ipir_loop_submit(loop, views[k], ia.ptr(outs[k])))
        _wait_until(lambda: _loop_stats(lib, loop).busy_partitions == 2, "the second partition")
        for k in range(6, 9):  # third wave: both partitions busy
            tickets.append(lib.ipir_loop_submit(loop, views[k], ia.ptr(outs[k])))
        _wait_until(lambda: _loop_stats(lib, loop).submitted == 9, "the third wave")
        import time
        time.sleep(0.05)
        st = _loop_stats(lib, loop)
        say(f"[loop] held: batches {st.batches}, pending {st.pending}, busy {st.busy_partitions}, "
            f"completed {st.completed}")
        assert st.max_busy_partitions == 2 and st.admitted_while_busy >= 1 and st.batches == 2
        assert st.completed == 0, "held partitions complete nothing"
        assert 3 <= st.pending <= 7, st.pending  # the third wave (and what partition 1 did not take) waits
        assert lib.ipir_dbg_loop_hold(loop, 0) == 0
        lat = C.c_double()
        for t in tickets:
            assert t > 0
            assert lib.ipir_loop_wait(loop, t, C.byref(lat)) == 0
            assert lat.value > 0
        st = _loop_stats(lib, loop)
        assert st.completed == 9 and st.failed == 0 and st.batches >= 3
        say(f"[loop] batches {st.batches}, largest {st.max_batch_size}, admitted while busy "
            f"{st.admitted_while_busy}, device overlaps {st.device_overlaps}")
    finally:
        lib.ipir_loop_stop(loop)
    for i in range(len(qs)):
        assert np.array_equal(outs[i], alone[i]), f"loop result {i} differs from alone"
        assert np.array_equal(ol.extract(geo, built[i][1], outs[i].reshape(geo["T"], -1)),
                              ol.expected_slots(geo, db, idxs[i]))
    assert np.array_equal(srv.answer_batch(qs[:2]), np.stack(alone[:2])), "entry points work after stop"
    assert lib.ipir_dbg_set_pipeline(srv.h, 3, 2) == 0


def test_serving_loop_sizes_partitions_from_the_slots(medium8):
    """ipir_loop_start(srv, 0): partitions of 32 slots (the whole batch if
    fewer), as many as max_batch holds; a burst larger than one partition
    comes back bit-identical to the alone answers."""
    geo, db, srv = medium8
    lib = srv.lib
    qs = [ol.make_query(geo, SEED, (7919 * k) % 524288, 950 + k)[0] for k in range(40)]
    alone = srv.answer_batch(qs[:32])
    alone = np.concatenate([alone, srv.answer_batch(qs[32:])])
    loop = lib.ipir_loop_start(srv.h, 0)
    assert loop
    outs = np.zeros((40, geo["T"] * 4 * N), np.uint64)
    views = [srv.views([q]) for q in qs]
    try:
        st = _loop_stats(lib, loop)
        assert (st.partitions, st.slots_per_partition) == (1, MB8)
        tickets = [lib.ipir_loop_submit(loop, views[k], ia.ptr(outs[k])) for k in range(40)]
        for t in tickets:
            assert lib.ipir_loop_wait(loop, t, None) == 0
        st = _loop_stats(lib, loop)
        assert st.completed == 40 and st.max_batch_size <= MB8
    finally:
        lib.ipir_loop_stop(loop)
    assert np.array_equal(outs, alone)


def test_pair_overrun_reported_on_the_right_call(medium8):
    """The pair Horner kernel's bounded wait, forced to give up at once
    (spin limit 0; every kernel still ends normally): the blocking call
    that ran it returns 4, the serving loop reports 4 for its request, the
    asynchronous device entry point never uses the pair and answers
    correctly, and once the bound is restored the next call is not blamed
    for the earlier failure."""
    geo, db, srv = medium8
    lib = srv.lib
    q, st = ol.make_query(geo, SEED, 4242, 31)
    want = srv.answer_batch([q])[0]
    out = np.zeros((1, geo["T"] * 4 * N), np.uint64)
    views = srv.views([q])
    assert lib.ipir_dbg_set_horner(srv.h, 1) == 0
    assert lib.ipir_dbg_set_horner_spin_limit(srv.h, 0) == 0
    try:
        assert lib.ipir_answer_batch(srv.h, views, 1, ia.ptr(out)) == 4
        dq = ia.DeviceBuffer(q.nbytes, lib)
        dq.upload(q)
        dout = ia.DeviceBuffer(want.nbytes, lib)
        assert lib.ipir_answer_batch_device(srv.h, C.c_void_p(dq.p), 1, C.c_void_p(dout.p), None) == 0
        assert lib.ipir_dev_sync() == 0
        assert np.array_equal(dout.download(np.zeros_like(want)), want), "device entry point (no pair)"
        loop = lib.ipir_loop_start(srv.h, 2)
        assert loop
        try:
            o2 = np.zeros(geo["T"] * 4 * N, np.uint64)
            t = lib.ipir_loop_submit(loop, views, ia.ptr(o2))
            assert lib.ipir_loop_wait(loop, t, None) == 4
        finally:
            lib.ipir_loop_stop(loop)
    finally:
        assert lib.ipir_dbg_set_horner_spin_limit(srv.h, 1 << 26) == 0
    try:
        assert lib.ipir_answer_batch(srv.h, views, 1, ia.ptr(out)) == 0, "a later call is not blamed"
        assert np.array_equal(out[0], want)
    finally:
        assert lib.ipir_dbg_set_horner(srv.h, -1) == 0


def test_device_calls_on_two_streams_are_ordered(medium8):
    """Asynchronous calls share the server's slots: one on stream A, the
    next on stream B, then a blocking host-staged call, repeated without
    synchronising in between; each call's responses equal its queries
    answered alone (the library orders each call after the previous one)."""
    geo, db, srv = medium8
    lib = srv.lib
    sets = [[ol.make_query(geo, SEED, (99991 * (3 * r + j)) % 524288, 1300 + 10 * r + j)[0] for j in range(3)]
            for r in range(3)]
    alone = [srv.answer_batch(x) for x in sets]
    sa, sb = lib.ipir_dev_stream_create(), lib.ipir_dev_stream_create()
    assert sa and sb
    bufs = []
    try:
        for r in range(2):
            dqa = ia.DeviceBuffer(np.stack(sets[0]).nbytes, lib)
            dqb = ia.DeviceBuffer(np.stack(sets[1]).nbytes, lib)
            da = ia.DeviceBuffer(alone[0].nbytes, lib)
            db_ = ia.DeviceBuffer(alone[1].nbytes, lib)
            bufs += [dqa, dqb, da, db_]
            dqa.upload(np.stack(sets[0]))
            dqb.upload(np.stack(sets[1]))
            assert lib.ipir_answer_batch_device(srv.h, C.c_void_p(dqa.p), 3, C.c_void_p(da.p), sa) == 0
            assert lib.ipir_answer_batch_device(srv.h, C.c_void_p(dqb.p), 3, C.c_void_p(db_.p), sb) == 0
            host = srv.answer_batch(sets[2])
            assert lib.ipir_dev_stream_sync(sa) == 0 and lib.ipir_dev_stream_sync(sb) == 0
            assert np.array_equal(da.download(np.zeros_like(alone[0])), alone[0]), f"stream A, round {r}"
            assert np.array_equal(db_.download(np.zeros_like(alone[1])), alone[1]), f"stream B, round {r}"
            assert np.array_equal(host, alone[2]), f"host-staged, round {r}"
    finally:
        for b in bufs:
            b.free()
        lib.ipir_dev_stream_destroy(sa)
        lib.ipir_dev_stream_destroy(sb)


# ---------------------------------------------------------------- several response ciphertexts
def test_two_ciphertext_responses_extract():
    """n_packed = 512 > D = 256: two response ciphertexts per query (the
    16 GB-and-up geometry's Horner split, reference src/gpu_protocol.cu:
    1300-1339), 4 GiB DB at db_rows 2048. Every response extracts to its
    entry, in a batch and alone, and the two Horner kernels agree; the
    streamed shapes (two chains per query waiting on one pack workgroup)
    give the serial answers."""
    n, w, rows = 2 ** 25, 120, 2048
    rc, geo = ol.setup(n, w, rows)
    assert geo["G"] == 512 and geo["T"] == 2
    db = np.random.default_rng(9).integers(0, P, size=(geo["R"], geo["C"]), dtype=np.uint16)
    say("[wide] GPU create (G=512, T=2)")
    srv = ia.Server(db, n, w, rows, max_batch=4, crs_seed=bytes(SEED))
    try:
        idxs = [5, 2 ** 24 + 77, n - 1]  # first half, second half (other ciphertext), last entry
        qs, sts = [], []
        for k, idx in enumerate(idxs):
            q, st = ol.make_query(geo, SEED, idx, 60 + k)
            qs.append(q)
            sts.append(st)
        batch = srv.answer_batch(qs)
        for k, idx in enumerate(idxs):
            got = ol.extract(geo, sts[k], batch[k].reshape(geo["T"], -1))
            assert np.array_equal(got, ol.expected_slots(geo, db, idx)), f"entry {idx}"
        single = srv.answer_batch(qs[:1])
        assert np.array_equal(single[0], batch[0])
        assert srv.lib.ipir_dbg_set_horner(srv.h, 1) == 0
        pair = srv.answer_batch(qs)
        assert srv.lib.ipir_dbg_set_horner(srv.h, 0) == 0
        one = srv.answer_batch(qs)
        assert np.array_equal(pair, batch) and np.array_equal(one, batch)
        # streamed: each query's two chains wait for the same pack workgroup's groups
        lib = srv.lib
        assert lib.ipir_dbg_take_streaming(srv.h) == 0
        assert lib.ipir_dbg_set_pipeline(srv.h, 0, 0) == 0
        serial, serial1 = srv.answer_batch(qs), srv.answer_batch(qs[:1])
        assert np.array_equal(serial, batch)
        for shape in (6, 8):
            for sl in (2, 4):
                assert lib.ipir_dbg_set_pipeline(srv.h, shape, shape) == 0
                assert lib.ipir_dbg_set_stream_slices(srv.h, 1, sl) == 0
                assert lib.ipir_dbg_set_stream_slices(srv.h, 3, sl) == 0
                n0 = lib.ipir_dbg_streamed_count(srv.h)
                say(f"[wide] streamed shape={shape} slices={sl}")
                assert np.array_equal(srv.answer_batch(qs), serial), f"shape {shape} slices {sl}, 3 queries"
                assert np.array_equal(srv.answer_batch(qs[:1]), serial1), f"shape {shape} slices {sl}, 1 query"
                assert lib.ipir_dbg_streamed_count(srv.h) == n0 + 2, "both answers streamed"
        # the packed wire path: both cts compressed on the device, query-major
        cl = ia.Client(srv.params, srv.lib)
        packed = np.stack([cl.pack_query(q) for q in qs])
        comp = srv.answer_packed(packed)
        assert comp.shape == (3, 2 * 12288)
        for k in range(3):
            assert np.array_equal(comp[k], cl.compress(batch[k])), f"compressed response {k}"
    finally:
        srv.close()
