"""CPU tests of the product libraries: every C-ABI symbol the headers declare
is exported; the client half (query build, extract, wire packing,
compression) agrees bit for bit with the oracle and the reference CPU
library; hostile parameters and bad arguments are rejected. No GPU calls.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import inspire_amd as ia
import oracle_lib as ol
from oracle_lib import N, P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
rng = np.random.default_rng(77)


def declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ipir_\w+)\s*\(", src)))


@pytest.mark.parametrize("header", ["capi.h", "inspire_debug.h"])
def test_server_library_exports_every_declared_symbol(header):
    lib = C.CDLL(ia.SERVER_SO)
    names = declared(header)
    assert len(names) > 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_client_library_exports_client_half():
    lib = C.CDLL(ia.CLIENT_SO)
    for n, _, _ in ia.CLIENT_FUNCS:
        assert hasattr(lib, n), n


def params_for(n, w, rows, seed):
    rc, geo = ol.setup(n, w, rows)
    assert rc == 0
    return geo, ia.make_params(n, w, rows, bytes(seed), geo["R"], geo["C"], geo["G"], geo["D"], geo["T"])


@pytest.mark.parametrize("n,w,rows,idx", [(256, 120, 2048, 7), (262144, 120, 2048, 262143),
                                         (2**23, 120, 32768, 1000000)])
def test_client_query_equals_oracle(n, w, rows, idx):
    seed = rng.integers(0, 256, 64, dtype=np.uint8)
    geo, p = params_for(n, w, rows, seed)
    cl = ia.Client(p)
    h, q = cl.build(idx, 31)
    try:
        q2, st = ol.make_query(geo, seed, idx, 31)
        assert np.array_equal(q, q2)
        # extract a synthetic "response" that encrypts known slots: use the
        # oracle answer on a tiny DB when the geometry allows it
    finally:
        cl.destroy(h)


@pytest.fixture(scope="module")
def small():
    seed = np.zeros(64, np.uint8)
    seed[0] = 0x42
    geo, p = params_for(262144, 120, 2048, seed)
    db = np.random.default_rng(0).integers(0, P, size=(geo["R"], geo["C"]), dtype=np.uint16)
    return geo, p, db, seed, ol.OraclePrecomp(db, geo, seed)


def test_client_extract_and_compressed_extract(small):
    geo, p, db, seed, pc = small
    cl = ia.Client(p)
    for idx in (0, 4242, 262143):
        h, q = cl.build(idx, 9 + idx)
        resp = pc.answer(q).reshape(-1)
        assert np.array_equal(cl.extract(h, resp), ol.expected_slots(geo, db, idx))
        nb = cl.lib.ipir_response_compressed_bytes(p.ref())
        assert nb == geo["T"] * 12288
        buf = np.zeros(nb, np.uint8)
        assert cl.lib.ipir_response_compress(p.ref(), ia.ptr(resp), ia.ptr(buf)) == 0
        ob = np.zeros(12288, np.uint8)
        ol.oracle().orc_compress_ct(np.ascontiguousarray(resp[:4 * N]), ob)
        assert np.array_equal(buf[:12288], ob)
        slots = np.zeros(cl.entry_slots(), np.uint16)
        assert cl.lib.ipir_extract_compressed(p.ref(), h, ia.ptr(buf), ia.ptr(slots)) == 0
        assert np.array_equal(slots, ol.expected_slots(geo, db, idx))
        cl.destroy(h)


def unpack_bits(buf, nvals, bits):
    b = np.unpackbits(buf)[:nvals * bits].reshape(nvals, bits).astype(np.uint64)
    return b.dot(np.uint64(1) << np.arange(bits - 1, -1, -1, dtype=np.uint64))


def test_wire_pack_roundtrip_and_format(small):
    geo, p, db, seed, pc = small
    cl = ia.Client(p)
    h, q = cl.build(123, 5)
    L = cl.lib
    nb = L.ipir_query_packed_bytes(p.ref())
    nvals = geo["R"] + 6 * 2 * N
    assert nb == (nvals * 53 + 7) // 8
    buf = np.zeros(nb, np.uint8)
    assert L.ipir_query_pack(p.ref(), ia.ptr(q), ia.ptr(buf)) == 0
    # format: MSB-first 53-bit Garner values of the (limb0, limb1) pairs
    vals = unpack_bits(buf, nvals, 53)
    R = geo["R"]
    pairs = [(q[i], q[R + i]) for i in range(0, R, 97)]
    for k, i in enumerate(range(0, R, 97)):
        v = int(vals[i])
        assert v < ol.Q0 * ol.Q1 and v % ol.Q0 == pairs[k][0] and v % ol.Q1 == pairs[k][1]
        assert v == ol.oracle().orc_garner(int(pairs[k][0]), int(pairs[k][1]))
    q2 = np.zeros_like(q)
    assert L.ipir_query_unpack(p.ref(), ia.ptr(buf), ia.ptr(q2)) == 0
    assert np.array_equal(q, q2)
    resp = pc.answer(q).reshape(-1)
    rb = np.zeros(L.ipir_response_packed_bytes(p.ref()), np.uint8)
    assert L.ipir_response_pack(p.ref(), ia.ptr(resp), ia.ptr(rb)) == 0
    r2 = np.zeros_like(resp)
    assert L.ipir_response_unpack(p.ref(), ia.ptr(rb), ia.ptr(r2)) == 0
    assert np.array_equal(resp, r2)
    cl.destroy(h)


def test_query_view_from_flat_and_sizes(small):
    geo, p, *_ = small
    L = ia.load("client")
    R = geo["R"]
    assert L.ipir_query_u64s(p.ref()) == 2 * R + 6 * ia.BLK
    assert L.ipir_response_u64s(p.ref()) == geo["T"] * 4 * N
    assert L.ipir_entry_slots(p.ref()) == 64
    flat = np.zeros(2 * R + 6 * ia.BLK, np.uint64)
    v = ia.ipir_query_view()
    assert L.ipir_query_view_from_flat(p.ref(), ia.ptr(flat), C.byref(v)) == 0
    base = flat.ctypes.data
    assert (v.lwe_b_limb1 - base, v.ksk5_b - base, v.kskneg1_b - base, v.rgsw - base) == \
        (8 * R, 16 * R, 8 * (2 * R + ia.BLK), 8 * (2 * R + 2 * ia.BLK))


def test_hostile_params_and_bad_args_rejected(small):
    geo, p, *_ = small
    L = ia.load("client")
    out = np.zeros(2 * geo["R"] + 6 * ia.BLK, np.uint64)
    # index out of range -> NULL (reference src/capi_client.cpp:98-99)
    assert not L.ipir_query_build(p.ref(), geo["n_entries"], ia.ptr(out))
    for field, val in (("ring_n", 1024), ("d_eff", 3), ("db_rows", 0), ("n_packed", 3), ("num_cts", 7),
                       ("entry_bytes", 32), ("db_cols", 4096), ("interp_d", 0)):
        bad = ia.ipir_params()
        C.pointer(bad)[0] = p.raw
        setattr(bad, field, val)
        assert L.ipir_query_u64s(C.byref(bad)) == 0, field
        assert not L.ipir_query_build(C.byref(bad), 0, ia.ptr(out)), field
    assert L.ipir_query_pack(p.ref(), None, None) == 1
    assert L.ipir_extract(p.ref(), None, ia.ptr(out), ia.ptr(out)) == 1


# ---------------------------------------------------------------- product setup()
GEOS = [((2**20, 60, 32768), 0), ((2**23, 120, 32768), 0), ((2**27, 120, 32768), 0),
        ((2**29, 120, 32768), 0), ((2**25, 3840, 32768), 0), ((256, 120, 2048), 0), ((256, 120, 0), 0),
        # the literal benchmark shapes: ceil(8w/15) does not divide N = 2048
        ((2**20, 32, 32768), 1), ((2**24, 64, 32768), 1), ((2**28, 64, 32768), 1), ((2**25, 4096, 32768), 1),
        # n_packed not a power of two (D_a must divide 2N); db_rows not a multiple of N
        ((3 * 2**20, 120, 32768), 1), ((2**20, 60, 1000), 1)]


@pytest.mark.parametrize("args,rc", GEOS)
def test_product_geometry_follows_setup_rules(args, rc):
    """The product's setup() (through ipir_geometry) accepts and rejects the
    shapes the reference's setup() does (src/params.cpp:46-60), with the
    same geometry as the oracle and, where built, the reference library."""
    raw = ia.ipir_params()
    got = ia.load("client").ipir_geometry(*args, C.byref(raw))
    assert got == rc, args
    orc, geo = ol.setup(args[0], args[1], args[2] or 32768)
    assert (orc != 0) == (rc != 0)  # the oracle names the rule that failed
    if ol.have_ref():
        out = np.zeros(5, np.uint64)
        assert (ol.ref().ref_setup(args[0], args[1], args[2] or 32768, out) != 0) == (rc != 0)
        if rc == 0:
            assert tuple(int(v) for v in out) == (geo["R"], geo["C"], geo["G"], geo["D"], geo["T"])
    if rc == 0:
        assert (raw.db_rows, raw.db_cols, raw.n_packed, raw.interp_d, raw.num_cts) == \
            (geo["R"], geo["C"], geo["G"], geo["D"], geo["T"])


@pytest.mark.parametrize("n,w,rows", [(2**20, 60, 32768), (2**27, 120, 32768), (2**25, 3840, 32768)])
def test_server_footprint_arithmetic(n, w, rows):
    """ipir_server_footprint: the resident DB (2 bytes per value) and
    precomp (7174 N words per group: a, Dd at 3.5 bytes per entry, Df) are
    in the fixed part; a slot holds at
    least its flat query and response staging; preprocessing peaks above
    the resident DB and precomp. Rejected geometries return 1."""
    f = ia.footprint(n, w, rows)
    rc, geo = ol.setup(n, w, rows)
    db_bytes = geo["R"] * geo["C"] * 2
    pc_bytes = geo["G"] * 7174 * N * 4
    assert db_bytes + pc_bytes < f["resident_fixed"] < db_bytes + pc_bytes + 2**30
    assert f["per_slot"] >= (2 * geo["R"] + 6 * ol.BLK) * 8 + geo["T"] * 4 * N * 8
    assert f["preprocess_peak"] > db_bytes + pc_bytes
    with pytest.raises(ValueError):
        ia.footprint(2**24, 64, 32768)
