"""GPU parity at the 1 GB geometry: 2^23 entries x 120 B, db_rows 32768,
8 groups (encode over D_a = 8, a 7-step Horner chain), uniform u16 DB
from seed 0 (the reference's default e2e geometry,
tests/test_e2e_gpu.cu:164-166).

The oracle encodes the whole DB and computes all 8 groups' precomp
(spread over host threads); the GPU's resident DB and every precomp
tensor must equal it, and for idx 0, 2^23 - 1 and six random indices the
database-pass output, the packed b, the response and extract() too. A
batch of those 8 equals the 8 single answers.
"""
import numpy as np
import pytest

import inspire_amd as ia
import oracle_lib as ol
import scale_lib as sl

pytestmark = pytest.mark.gpu
N_ENTRIES, W, ROWS = 2 ** 23, 120, 32768


@pytest.fixture(scope="module")
def gb1():
    rc, geo = ol.setup(N_ENTRIES, W, ROWS)
    assert rc == 0 and (geo["R"], geo["C"], geo["G"], geo["Da"], geo["T"]) == (32768, 16384, 8, 8, 1)
    sl.need_host(8 * 2 ** 30, "1 GB geometry")
    sl.say("[2^23 x 120 B] generating DB")
    db = sl.gen_db(geo["R"], geo["C"], 0)
    sl.say("[2^23 x 120 B] GPU create")
    srv = ia.Server(db, N_ENTRIES, W, ROWS, max_batch=8, crs_seed=bytes(sl.SEED))
    sl.say("[2^23 x 120 B] oracle encode + precomp of 8 groups")
    enc = sl.oracle_encode(db, geo)
    pre = sl.oracle_precomp_groups(enc, geo, range(geo["G"]))
    yield geo, db, enc, pre, srv
    srv.close()


def test_encoded_db_matches_oracle(gb1):
    geo, db, enc, pre, srv = gb1
    got = np.zeros((geo["R"], geo["C"]), np.uint16)
    assert srv.lib.ipir_dbg_db_rows(srv.h, 0, geo["R"], ia.ptr(got)) == 0
    bad = np.argwhere(got != enc)
    assert bad.size == 0, f"{len(bad)} encoded values differ, first at {bad[:3].tolist()}"


def test_precomp_of_every_group_matches_oracle(gb1):
    geo, db, enc, pre, srv = gb1
    for g in range(geo["G"]):
        t = srv.precomp_group(g)
        for k in ("a", "Dp", "Dm", "Df"):
            assert np.array_equal(t[k], pre[g][k]), f"{k} of group {g}"


def test_queries_every_stage_batch_and_extract(gb1):
    geo, db, enc, pre, srv = gb1
    idxs = [0, N_ENTRIES - 1] + sl.random_indices(N_ENTRIES, 6)
    qs = [ol.make_query(geo, sl.SEED, idx, 2 * 1000003 + k) for k, idx in enumerate(idxs)]
    sl.say("[2^23 x 120 B] oracle answers of 8 queries")
    want = sl.pmap(lambda q: sl.oracle_answer(enc, geo, pre, q[0], taps=True), qs)
    for k, ((q, st), idx) in enumerate(zip(qs, idxs)):
        interm = np.zeros(2 * geo["C"], np.uint32)
        packed = np.zeros(geo["G"] * 2 * ol.N, np.uint32)
        resp = np.zeros(geo["T"] * 4 * ol.N, np.uint64)
        assert srv.lib.ipir_dbg_answer_taps(srv.h, ia.ptr(q), ia.ptr(interm), ia.ptr(packed), ia.ptr(resp)) == 0
        w_resp, w_interm, w_packed = want[k]
        assert np.array_equal(interm.astype(np.uint64), w_interm), f"DB pass, idx {idx}"
        assert np.array_equal(packed.astype(np.uint64), w_packed), f"packed b, idx {idx}"
        assert np.array_equal(resp, w_resp.reshape(-1)), f"response, idx {idx}"
        assert np.array_equal(ol.extract(geo, st, resp.reshape(geo["T"], -1)), ol.expected_slots(geo, db, idx))
    batch = srv.answer_batch([q for q, _ in qs])
    for k in range(len(qs)):
        assert np.array_equal(batch[k], want[k][0].reshape(-1)), f"batch of 8, idx {idxs[k]}"
    sl.say("[2^23 x 120 B] 8 queries bit-exact at every stage, batch of 8 equal")
