"""Write tests/golden/ref_vectors.npz from the reference's own CPU library.

The library is oracle/_ref/libinspire_ref.so, compiled from the reference's
source files by oracle/Makefile (only where the reference tree exists). Every
input is regenerated from a fixed numpy seed, so the fixture holds the seeds
and, per output, its SHA-256 digest plus the first 8 words (a readable
check). tests/test_golden.py recomputes the outputs with the oracle and
compares. Run: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as ol  # noqa: E402
from oracle_lib import BLK, N, QS  # noqa: E402


def digest(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), np.uint8)


def cases(lib, prefix):
    """Yield (name, output array) using lib's functions (prefix 'ref_' or 'orc_')."""
    f = lambda name: getattr(lib, prefix + name)  # noqa: E731
    rng = np.random.default_rng(20240601)
    for limb in (0, 1):
        x = rng.integers(0, QS[limb], N, dtype=np.uint64)
        y = x.copy()
        f("ntt_fwd")(y, limb)
        yield f"ntt_fwd_l{limb}", y.copy()
        f("ntt_inv")(y, limb)
        yield f"ntt_roundtrip_l{limb}", y
        z = x.copy()
        f("ntt_inv")(z, limb)
        yield f"ntt_inv_l{limb}", z
    p = ol.rand_poly(rng).reshape(-1)
    d = np.zeros(4 * N, np.uint64)
    f("decomp")(p, d)
    yield "decomp", d
    for k in (5, 3277, 2 * N - 1, 25):
        a = np.zeros(2 * N, np.uint64)
        f("automorphism")(p, k, a)
        yield f"automorphism_{k}", a
    seed = np.array([0x42] + [0] * 63, np.uint8)
    for dom, idx in ((b"rlwe", 0), (b"rlwe", 7), (b"ksk", 501), (b"ksk", 204700 + 2)):
        e = np.zeros(2 * N, np.uint64)
        f("expand_seed")(seed, dom, idx, e)
        yield f"expand_{dom.decode()}_{idx}", e
    # seeded client query at the reference's Small and Medium test geometries
    for n, w, rows, idx in ((256, 120, 2048, 42), (262144, 120, 2048, 100000)):
        out = np.zeros(5, np.uint64)
        assert f("setup")(n, w, rows, out) == 0
        R, Cc, G, D, T = (int(v) for v in out)
        q = np.zeros(2 * R + 6 * BLK, np.uint64)
        st = np.zeros(4 * N + 2, np.uint64)
        f("query_seeded")(R, Cc, D, w, seed, idx, 99, q, st)
        yield f"query_{n}_{idx}", q
        yield f"query_state_{n}_{idx}", st
        if n == 262144:
            rg = np.ascontiguousarray(q[2 * R + 2 * BLK:])
            ct = ol.rand_poly(rng, (2,)).reshape(-1)
            r = np.zeros(4 * N, np.uint64)
            f("ext_prod")(rg, ct, r)
            yield "ext_prod", r
            for L in (2, 5):
                cts = ol.rand_poly(rng, (L, 2)).reshape(-1)
                h = np.zeros(4 * N, np.uint64)
                f("horner")(cts, L, rg, h)
                yield f"horner_L{L}", h
            m = np.zeros(N, np.uint64)
            f("rlwe_dec")(np.ascontiguousarray(st[2 * N:4 * N]), ct, m)
            yield "rlwe_dec", m


def main():
    lib = ol.ref()
    if lib is None:
        sys.exit("oracle/_ref/libinspire_ref.so is not built (make -C oracle ref, with the reference tree present)")
    out = {}
    for name, arr in cases(lib, "ref_"):
        out[name + "__sha256"] = digest(arr)
        out[name + "__head"] = np.ascontiguousarray(arr[:8], np.uint64)
    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez(path, **out)
    print(f"wrote {len(out) // 2} vectors to {path}")


if __name__ == "__main__":
    main()
