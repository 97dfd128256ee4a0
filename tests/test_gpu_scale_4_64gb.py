"""GPU answers at the 64 GB geometry: 2^29 entries x 120 B, db_rows 32768,
512 groups in two response ciphertexts of 256 (68.7 GB DB, ~103 GB
resident), uniform u16 DB from seed 0.

At this size the oracle checks what it can finish in seconds: the
database pass over a sampled column window of the device's encoded DB.
The answers are checked through extract() at full size: idx 0, 2^29 - 1
and random indices landing in both ciphertexts; a batch equals its single
answers, and both Horner kernels give the same responses.
"""
import numpy as np
import pytest

import inspire_amd as ia
import oracle_lib as ol
import scale_lib as sl

pytestmark = pytest.mark.gpu
N_ENTRIES, W, ROWS = 2 ** 29, 120, 32768


@pytest.fixture(scope="module")
def gb64():
    rc, geo = ol.setup(N_ENTRIES, W, ROWS)
    assert rc == 0 and (geo["G"], geo["Da"], geo["T"]) == (512, 256, 2)
    sl.need_host(80 * 2 ** 30, "64 GB geometry")
    sl.say("[2^29 x 120 B] generating DB (68.7 GB)")
    db = sl.gen_db(geo["R"], geo["C"], 0)
    sl.say("[2^29 x 120 B] GPU create")
    srv = ia.Server(db, N_ENTRIES, W, ROWS, max_batch=8, crs_seed=bytes(sl.SEED))
    sl.say("[2^29 x 120 B] server ready")
    yield geo, db, srv
    srv.close()


def test_db_pass_on_a_column_window_matches_oracle(gb64):
    geo, db, srv = gb64
    R, Cc = geo["R"], geo["C"]
    rng = np.random.default_rng(41)
    b = np.zeros((2, 2, R), np.uint64)
    b[:, 0] = rng.integers(0, ol.Q0, (2, R), dtype=np.uint64)
    b[:, 1] = rng.integers(0, ol.Q1, (2, R), dtype=np.uint64)
    got = np.zeros((2, 2, Cc), np.uint32)
    assert srv.lib.ipir_dbg_db_pass(srv.h, ia.ptr(b), 2, -1, ia.ptr(got)) == 0
    for k, col0 in ((0, 0), (1, Cc - 2048)):
        blk = srv.block(0, R, col0, 2048)
        want = np.zeros(2 * 2048, np.uint64)
        ol.oracle().orc_matvec(blk.reshape(-1), R, 2048, np.ascontiguousarray(b[k, 0]), np.ascontiguousarray(b[k, 1]),
                               0, 2048, want)
        assert np.array_equal(got[k, :, col0:col0 + 2048].reshape(-1).astype(np.uint64), want), f"query {k}"


def test_answers_in_both_ciphertexts_extract(gb64):
    geo, db, srv = gb64
    idxs = [0, N_ENTRIES - 1] + sl.random_indices(N_ENTRIES, 6, seed=42)
    assert {sl.ct_of(geo, i) for i in idxs} == {0, 1}
    built = [ol.make_query(geo, sl.SEED, idx, 4 * 1000003 + k) for k, idx in enumerate(idxs)]
    batch = srv.answer_batch([q for q, _ in built])
    for (q, st), r, idx in zip(built, batch, idxs):
        assert np.array_equal(ol.extract(geo, st, r.reshape(geo["T"], -1)), ol.expected_slots(geo, db, idx)), idx
    for k in (0, 1, 7):
        assert np.array_equal(srv.answer_batch([built[k][0]])[0], batch[k]), f"single {k}"
    assert srv.lib.ipir_dbg_set_horner(srv.h, 1) == 0  # the workgroup-pair kernel for every chain
    try:
        pair = srv.answer_batch([q for q, _ in built[:3]])
    finally:
        assert srv.lib.ipir_dbg_set_horner(srv.h, -1) == 0
    assert np.array_equal(pair, batch[:3])
    sl.say(f"[2^29 x 120 B] {len(idxs)} entries extracted over both ciphertexts")
