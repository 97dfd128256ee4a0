"""GPU parity at the 16 GB geometry the benchmark runs: 2^27 entries x
120 B, db_rows 32768, 128 groups (encode over D_a = 128, 127 Horner
steps), uniform u16 DB from seed 0, served with max_batch 32 (the
reference runs this tier behind a flag, tests/test_e2e_gpu.cu:124-147).

The whole DB is too large for the oracle, so it is compared where the
oracle finishes in seconds and through properties at full size:
* the encode of sampled rows (oracle direct sum over the row's D_a
  blocks) against the resident DB;
* the database pass of 1, 2 and 32 queries (the one-query pass, a
  2..8-query pass and the 32-query power-row pass) over sampled
  2048-column windows against the oracle's mat-vec of the device's
  encoded block;
* one sampled group's precomp recomputed by the oracle from the
  device-encoded block of that group;
* full-size answers: idx 0, 2^27 - 1 and 64 random indices extract to
  their entries, and a batch of 32 equals the 32 single answers.
"""
import numpy as np
import pytest

import inspire_amd as ia
import oracle_lib as ol
import scale_lib as sl

pytestmark = pytest.mark.gpu
N_ENTRIES, W, ROWS = 2 ** 27, 120, 32768


@pytest.fixture(scope="module")
def gb16():
    rc, geo = ol.setup(N_ENTRIES, W, ROWS)
    assert rc == 0 and (geo["R"], geo["C"], geo["G"], geo["Da"], geo["T"]) == (32768, 262144, 128, 128, 1)
    sl.need_host(40 * 2 ** 30, "16 GB geometry")
    sl.say("[2^27 x 120 B] generating DB (17.2 GB)")
    db = sl.gen_db(geo["R"], geo["C"], 0)
    sl.say("[2^27 x 120 B] GPU create")
    srv = ia.Server(db, N_ENTRIES, W, ROWS, max_batch=32, crs_seed=bytes(sl.SEED))
    sl.say("[2^27 x 120 B] server ready")
    yield geo, db, srv
    srv.close()


def test_sampled_rows_encode_like_the_oracle(gb16):
    geo, db, srv = gb16
    R, Cc = geo["R"], geo["C"]
    rows = [0, 1, 4097, R - 1] + [int(r) for r in np.random.default_rng(21).integers(0, R, 4)]

    def one(r):
        # the oracle encodes row r of a one-row DB copy in place
        row = np.ascontiguousarray(db[r:r + 1]).copy()
        ol.oracle().orc_encode_rows(row.reshape(-1), 1, Cc, geo["D"], np.zeros(1, np.uint64), 1)
        return row[0]
    want = sl.pmap(one, rows)
    for r, w in zip(rows, want):
        got = np.zeros((1, Cc), np.uint16)
        assert srv.lib.ipir_dbg_db_rows(srv.h, r, 1, ia.ptr(got)) == 0
        assert np.array_equal(got[0], w), f"row {r}"


def test_db_pass_on_a_column_window_matches_oracle(gb16):
    geo, db, srv = gb16
    R, Cc = geo["R"], geo["C"]
    rng = np.random.default_rng(22)
    for nq in (1, 2, 32):  # the one-query pass, a 2..8-query pass, the 32-query pass
        b = np.zeros((nq, 2, R), np.uint64)
        b[:, 0] = rng.integers(0, ol.Q0, (nq, R), dtype=np.uint64)
        b[:, 1] = rng.integers(0, ol.Q1, (nq, R), dtype=np.uint64)
        b[-1, :, :R // 16] = np.array([ol.Q0 - 1, ol.Q1 - 1], np.uint64)[:, None]
        got = np.zeros((nq, 2, Cc), np.uint32)
        assert srv.lib.ipir_dbg_db_pass(srv.h, ia.ptr(b), nq, -1, ia.ptr(got)) == 0
        for k in sorted({0, nq - 1}):
            col0 = int(rng.integers(0, Cc // 2048)) * 2048
            blk = srv.block(0, R, col0, 2048)
            want = np.zeros(2 * 2048, np.uint64)
            ol.oracle().orc_matvec(blk.reshape(-1), R, 2048, np.ascontiguousarray(b[k, 0]),
                                   np.ascontiguousarray(b[k, 1]), 0, 2048, want)
            assert np.array_equal(got[k, :, col0:col0 + 2048].reshape(-1).astype(np.uint64), want), \
                f"{nq} queries, query {k}, columns {col0}.."


def test_sampled_group_precomp_from_the_encoded_db(gb16):
    geo, db, srv = gb16
    R = geo["R"]
    g = int(np.random.default_rng(23).integers(0, geo["G"]))
    blk = np.ascontiguousarray(srv.block(0, R, g * ol.N, ol.N))
    sl.say(f"[2^27 x 120 B] oracle precomp of group {g} from its encoded columns")
    sub = dict(geo, C=ol.N, G=1, Da=1)
    want = sl.oracle_precomp_groups(blk, sub, [0])[0]
    got = srv.precomp_group(g)
    for k in ("a", "Dp", "Dm", "Df"):
        assert np.array_equal(got[k], want[k]), f"{k} of group {g}"


def test_full_size_answers_extract_and_batch_equals_singles(gb16):
    geo, db, srv = gb16
    idxs = [0, N_ENTRIES - 1] + sl.random_indices(N_ENTRIES, 64)
    built = [ol.make_query(geo, sl.SEED, idx, 3 * 1000003 + k) for k, idx in enumerate(idxs)]
    resp = []
    for c0 in range(0, len(built), 32):
        resp.extend(srv.answer_batch([q for q, _ in built[c0:c0 + 32]]))
    bad = [idx for (q, st), r, idx in zip(built, resp, idxs)
           if not np.array_equal(ol.extract(geo, st, r.reshape(geo["T"], -1)), ol.expected_slots(geo, db, idx))]
    assert not bad, f"{len(bad)} of {len(idxs)} entries not extracted: {bad[:5]}"
    singles = [srv.answer_batch([q])[0] for q, _ in built[:32]]  # the first batch of 32
    for k in range(32):
        assert np.array_equal(resp[k], singles[k]), f"batch of 32 vs single, query {k}"
    sl.say(f"[2^27 x 120 B] {len(idxs)} entries extracted, batch of 32 equals singles")
