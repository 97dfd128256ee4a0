"""The allocation-free Horner entry point (ipir_horner_noalloc: device
buffers in, device buffer out, its scratch a static per-GPU array)
against the allocating one (ipir_dbg_horner) and the oracle, on calls in
a row with a shorter chain after a longer one; sizes past its maximum
are rejected with nothing launched."""
import ctypes as C

import numpy as np
import pytest

import inspire_amd as ia
import oracle_lib as ol
from oracle_lib import BLK, N

pytestmark = pytest.mark.gpu
SEED = np.array([0x42] + [0] * 63, np.uint8)


def _rgsw():
    rc, geo = ol.setup(262144, 120, 2048)
    q, _ = ol.make_query(geo, SEED, 1000, 3)
    return np.ascontiguousarray(q[2 * geo["R"] + 2 * BLK:])


def test_noalloc_horner_equals_allocating_entry_point():
    lib = ia.load("server")
    f = lib.ipir_horner_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
    rg = _rgsw()
    rng = np.random.default_rng(2024)
    d_rg = ia.DeviceBuffer(rg.nbytes, lib)
    d_rg.upload(rg)
    d_out = ia.DeviceBuffer(4 * N * 8, lib)
    d_cts = ia.DeviceBuffer(256 * 4 * N * 8, lib)
    stream = lib.ipir_dev_stream_create()
    try:
        for L in (9, 3, 1, 256):  # a shorter chain after a longer one, the one-element chain, the maximum
            cts = ol.rand_poly(rng, (L, 2)).reshape(-1)
            cts[: 2 * N] += np.uint64(ol.Q0)  # non-canonical residues are reduced like the other entry points
            d_cts.upload(cts)
            assert f(C.c_void_p(d_cts.p), L, C.c_void_p(d_rg.p), C.c_void_p(d_out.p), stream) == 0
            assert lib.ipir_dev_stream_sync(stream) == 0
            got = d_out.download(np.zeros(4 * N, np.uint64))
            want = np.zeros(4 * N, np.uint64)
            assert lib.ipir_dbg_horner(ia.ptr(cts), L, ia.ptr(rg), ia.ptr(want)) == 0
            assert np.array_equal(got, want), f"L={L}"
            if L <= 9:
                orc = np.zeros(4 * N, np.uint64)
                ol.oracle().orc_horner(ol.reduce_limbs(cts.reshape(L, 2, 2, N)).reshape(-1), L, rg, orc)
                assert np.array_equal(got, orc), f"L={L} vs oracle"
        before = d_out.download(np.zeros(4 * N, np.uint64))
        assert f(C.c_void_p(d_cts.p), 257, C.c_void_p(d_rg.p), C.c_void_p(d_out.p), stream) == 2
        assert f(C.c_void_p(d_cts.p), 0, C.c_void_p(d_rg.p), C.c_void_p(d_out.p), stream) == 1
        assert f(None, 3, C.c_void_p(d_rg.p), C.c_void_p(d_out.p), stream) == 1
        assert lib.ipir_dev_stream_sync(stream) == 0
        assert np.array_equal(d_out.download(np.zeros(4 * N, np.uint64)), before), "rejected calls write nothing"
        # the null stream works too
        assert f(C.c_void_p(d_cts.p), 2, C.c_void_p(d_rg.p), C.c_void_p(d_out.p), None) == 0
        assert lib.ipir_dev_sync() == 0
        want = np.zeros(4 * N, np.uint64)
        cts = d_cts.download(np.zeros(256 * 4 * N, np.uint64))[: 2 * 4 * N].copy()
        assert lib.ipir_dbg_horner(ia.ptr(cts), 2, ia.ptr(rg), ia.ptr(want)) == 0
        assert np.array_equal(d_out.download(np.zeros(4 * N, np.uint64)), want), "null stream"
    finally:
        lib.ipir_dev_stream_destroy(stream)
        d_rg.free()
        d_out.free()
        d_cts.free()
