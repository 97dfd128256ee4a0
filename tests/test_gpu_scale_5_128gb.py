"""GPU answers at the 128 GB wide-record geometry: 2^25 entries x 3840 B
(one entry fills a group's 2048 slots), db_rows 32768, 1024 groups in
four response ciphertexts (137.4 GB DB, ~206 GB resident), uniform u16
DB from seed 0, with max_batch sized from the device's free memory.

* The server's footprint: ipir_fit_max_batch sizes max_batch against the
  free device memory; the server then reports resident_bytes = fixed +
  max_batch x per-slot bytes, and the device's free memory drops by that
  much (within allocation granularity).
* The database pass over a sampled column window against the oracle's
  mat-vec of the device's encoded block.
* Answers at full size: entries in each of the four ciphertexts, idx 0
  and 2^25 - 1, extract to their entries; a batch equals its singles.
"""
import numpy as np
import pytest

import inspire_amd as ia
import oracle_lib as ol
import scale_lib as sl

pytestmark = pytest.mark.gpu
N_ENTRIES, W, ROWS = 2 ** 25, 3840, 32768


@pytest.fixture(scope="module")
def gb128():
    rc, geo = ol.setup(N_ENTRIES, W, ROWS)
    assert rc == 0 and (geo["G"], geo["Da"], geo["T"]) == (1024, 256, 4)
    sl.need_host(150 * 2 ** 30, "128 GB geometry")
    lib = ia.load("server")
    fp = ia.footprint(N_ENTRIES, W, ROWS)
    free0, total = ia.mem_info()
    if total < fp["preprocess_peak"] + 8 * 2 ** 30:
        pytest.skip(f"device memory {total / 2**30:.0f} GiB below this geometry's footprint")
    mb = lib.ipir_fit_max_batch(N_ENTRIES, W, ROWS, 4 * 2 ** 30, 8)
    assert 1 <= mb <= 8
    sl.say(f"[2^25 x 3840 B] generating DB (137.4 GB); max_batch {mb} fits the free device memory")
    db = sl.gen_db(geo["R"], geo["C"], 0)
    free1, _ = ia.mem_info()
    sl.say("[2^25 x 3840 B] GPU create")
    srv = ia.Server(db, N_ENTRIES, W, ROWS, max_batch=mb, crs_seed=bytes(sl.SEED))
    free2, _ = ia.mem_info()
    sl.say(f"[2^25 x 3840 B] server ready; device memory taken {(free1 - free2) / 2**30:.2f} GiB")
    yield geo, db, srv, fp, mb, free1 - free2
    srv.close()


def test_footprint_sizes_the_server(gb128):
    geo, db, srv, fp, mb, taken = gb128
    caps = srv.caps()
    assert caps.max_batch == mb and caps.num_cts == 4
    assert caps.resident_bytes == fp["resident_fixed"] + mb * fp["per_slot"]
    # hipMalloc rounds each allocation up; a few hundred buffers at most
    assert abs(taken - caps.resident_bytes) <= 0.005 * caps.resident_bytes + 2 ** 28, (taken, caps.resident_bytes)


def test_db_pass_on_a_column_window_matches_oracle(gb128):
    geo, db, srv, *_ = gb128
    R, Cc = geo["R"], geo["C"]
    rng = np.random.default_rng(51)
    b = np.zeros((1, 2, R), np.uint64)
    b[:, 0] = rng.integers(0, ol.Q0, (1, R), dtype=np.uint64)
    b[:, 1] = rng.integers(0, ol.Q1, (1, R), dtype=np.uint64)
    got = np.zeros((1, 2, Cc), np.uint32)
    assert srv.lib.ipir_dbg_db_pass(srv.h, ia.ptr(b), 1, -1, ia.ptr(got)) == 0
    col0 = int(rng.integers(0, Cc // 2048)) * 2048
    blk = srv.block(0, R, col0, 2048)
    want = np.zeros(2 * 2048, np.uint64)
    ol.oracle().orc_matvec(blk.reshape(-1), R, 2048, np.ascontiguousarray(b[0, 0]), np.ascontiguousarray(b[0, 1]),
                           0, 2048, want)
    assert np.array_equal(got[0, :, col0:col0 + 2048].reshape(-1).astype(np.uint64), want)


def test_answers_in_all_four_ciphertexts_extract(gb128):
    geo, db, srv, fp, mb, _ = gb128
    R = geo["R"]
    rng = np.random.default_rng(52)
    # one random entry per ciphertext (entry idx sits in group idx // R), plus both ends
    idxs = [int(ct * 256 * R + rng.integers(0, 256 * R)) for ct in range(4)] + [0, N_ENTRIES - 1]
    idxs = idxs[:mb] if mb < len(idxs) else idxs
    assert {sl.ct_of(geo, i) for i in idxs} >= ({0, 1, 2, 3} if mb >= 4 else {0})
    built = [ol.make_query(geo, sl.SEED, idx, 5 * 1000003 + k) for k, idx in enumerate(idxs)]
    batch = srv.answer_batch([q for q, _ in built])
    for (q, st), r, idx in zip(built, batch, idxs):
        assert np.array_equal(ol.extract(geo, st, r.reshape(geo["T"], -1)), ol.expected_slots(geo, db, idx)), idx
    for k in range(len(idxs)):
        assert np.array_equal(srv.answer_batch([built[k][0]])[0], batch[k]), f"single {k}"
    sl.say(f"[2^25 x 3840 B] {len(idxs)} entries extracted over all four ciphertexts")
