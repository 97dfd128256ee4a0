"""GPU parity at the smallest benchmark geometry: 2^20 entries x 60 B,
db_rows 32768 (one group, so no encode and no Horner step), uniform u16
DB from seed 0, 16 single queries with indices from seed 1 (the
reference's e2e test shape, tests/test_e2e_gpu.cu:95-104, at this size).

Everything the oracle computes is compared bit for bit: the resident
(encoded) DB, every precomp tensor, and per query the database-pass
output, the packed b after the collapse, the response, and extract().
"""
import numpy as np
import pytest

import inspire_amd as ia
import oracle_lib as ol
import scale_lib as sl

pytestmark = pytest.mark.gpu
N_ENTRIES, W, ROWS = 2 ** 20, 60, 32768


@pytest.fixture(scope="module")
def small():
    rc, geo = ol.setup(N_ENTRIES, W, ROWS)
    assert rc == 0 and (geo["R"], geo["C"], geo["G"], geo["T"]) == (32768, 2048, 1, 1)
    sl.say("[2^20 x 60 B] generating DB, oracle preprocess")
    db = sl.gen_db(geo["R"], geo["C"], 0)
    pc = ol.OraclePrecomp(db, geo, sl.SEED)
    sl.say("[2^20 x 60 B] GPU create")
    srv = ia.Server(db, N_ENTRIES, W, ROWS, max_batch=16, crs_seed=bytes(sl.SEED))
    yield geo, db, pc, srv
    srv.close()


def test_encoded_db_and_precomp_match_oracle(small):
    geo, db, pc, srv = small
    got = np.zeros((geo["R"], geo["C"]), np.uint16)
    assert srv.lib.ipir_dbg_db_rows(srv.h, 0, geo["R"], ia.ptr(got)) == 0
    assert np.array_equal(got, pc.db)
    t = srv.precomp_group(0)
    assert np.array_equal(t["a"], pc.a[0].reshape(-1)), "a"
    assert np.array_equal(t["Dp"], pc.Dp[0].reshape(-1)), "D+"
    assert np.array_equal(t["Dm"], pc.Dm[0].reshape(-1)), "D-"
    assert np.array_equal(t["Df"], pc.Df[0].reshape(-1)), "D_final"


def test_sixteen_queries_every_stage_and_extract(small):
    geo, db, pc, srv = small
    idxs = sl.random_indices(N_ENTRIES, 16)
    idxs[0], idxs[-1] = 0, N_ENTRIES - 1
    qs = [ol.make_query(geo, sl.SEED, idx, 2 * 1000003 + k) for k, idx in enumerate(idxs)]
    want = sl.pmap(lambda q: pc.answer(q[0], taps=True), qs)
    for k, ((q, st), idx) in enumerate(zip(qs, idxs)):
        interm = np.zeros(2 * geo["C"], np.uint32)
        packed = np.zeros(geo["G"] * 2 * ol.N, np.uint32)
        resp = np.zeros(geo["T"] * 4 * ol.N, np.uint64)
        assert srv.lib.ipir_dbg_answer_taps(srv.h, ia.ptr(q), ia.ptr(interm), ia.ptr(packed), ia.ptr(resp)) == 0
        w_resp, w_interm, w_packed = want[k]
        assert np.array_equal(interm.astype(np.uint64), w_interm), f"DB pass, query {k}"
        assert np.array_equal(packed.astype(np.uint64), w_packed), f"packed b, query {k}"
        assert np.array_equal(resp, w_resp.reshape(-1)), f"response, query {k}"
        assert np.array_equal(ol.extract(geo, st, resp.reshape(geo["T"], -1)), ol.expected_slots(geo, db, idx))
    batch = srv.answer_batch([q for q, _ in qs])
    for k in range(16):
        assert np.array_equal(batch[k], want[k][0].reshape(-1)), f"batch of 16, query {k}"
    sl.say("[2^20 x 60 B] 16 queries bit-exact at every stage, batch of 16 equal")
