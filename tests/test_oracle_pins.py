"""Pins of the oracle (oracle/inspire_oracle.cpp) against independent facts:
brute-force closed forms, hashlib's SHAKE-256, known-answer tests from the
reference's own test suite, and -- where it was built -- the reference's own
CPU library (oracle/_ref). CPU only.
"""
import hashlib

import numpy as np
import pytest

import oracle_lib as ol
from oracle_lib import BLK, N, P, Q0, Q1, QS

rng = np.random.default_rng(1234)
needs_ref = pytest.mark.skipif(not ol.have_ref(), reason="reference CPU library not built here")


def bitrev(x, bits=11):
    return int(f"{x:0{bits}b}"[::-1], 2)


def tables(limb):
    arrs = [np.zeros(N, np.uint64) for _ in range(4)]
    sc = np.zeros(4, np.uint64)
    ol.oracle().orc_ntt_tables(limb, *arrs, sc)
    return arrs, sc


# ---------------------------------------------------------------- NTT
@pytest.mark.parametrize("limb", [0, 1])
def test_psi_rule_and_twiddles(limb):
    q = QS[limb]
    (fwd, inv, fwdp, invp), sc = tables(limb)
    # smallest g >= 2 whose (q-1)/2N-th power has order exactly 2N
    e = (q - 1) // (2 * N)
    g = 2
    while pow(pow(g, e, q), N, q) != q - 1:
        g += 1
    psi = pow(g, e, q)
    assert int(sc[0]) == psi and int(sc[3]) == q
    assert pow(psi, 2 * N, q) == 1 and pow(psi, N, q) == q - 1
    for i in (0, 1, 2, 3, 1000, N // 2, N - 1):
        assert int(fwd[i]) == pow(psi, bitrev(i), q)
        assert int(fwd[i]) * int(inv[i]) % q == 1
        assert int(fwdp[i]) == (int(fwd[i]) << 32) // q
    assert int(sc[1]) * N % q == 1


def brute_ntt(f, limb):
    """NTT(f)[i] = sum_k f_k psi^((2BR(i)+1) k): the evaluation closed form."""
    q = QS[limb]
    psi = int(tables(limb)[1][0])
    pw = np.array([pow(psi, e, q) for e in range(2 * N)], dtype=np.uint64)
    br = np.array([bitrev(i) for i in range(N)], dtype=np.uint64)
    E = ((2 * br[:, None] + 1) * np.arange(N, dtype=np.uint64)[None, :]) % (2 * N)
    prod = (f[None, :].astype(np.uint64) * pw[E]) % np.uint64(q)
    return (prod.sum(axis=1) % np.uint64(q)).astype(np.uint64)


@pytest.mark.parametrize("limb", [0, 1])
def test_ntt_closed_form_and_roundtrip(limb):
    q = QS[limb]
    f = rng.integers(0, q, N, dtype=np.uint64)
    y = f.copy()
    ol.oracle().orc_ntt_fwd(y, limb)
    assert np.array_equal(y, brute_ntt(f, limb))
    ol.oracle().orc_ntt_inv(y, limb)
    assert np.array_equal(y, f)


def test_poly_mul_kat():
    # (1+X)^2 = 1 + 2X + X^2 (reference tests/test_ring.cpp:23-34)
    for limb in (0, 1):
        a = np.zeros(N, np.uint64)
        a[0] = a[1] = 1
        ol.oracle().orc_ntt_fwd(a, limb)
        a = (a * a) % np.uint64(QS[limb])
        ol.oracle().orc_ntt_inv(a, limb)
        assert list(a[:4]) == [1, 2, 1, 0] and not a[3:].any()


# ---------------------------------------------------------------- SHAKE / seeds
@pytest.mark.parametrize("inlen,outlen", [(0, 32), (1, 200), (135, 136), (136, 137), (300, 1000)])
def test_shake256_vs_hashlib(inlen, outlen):
    msg = rng.integers(0, 256, inlen, dtype=np.uint8)
    out = np.zeros(max(outlen, 1), np.uint8)
    ol.oracle().orc_shake256(msg if inlen else np.zeros(1, np.uint8), inlen, out, outlen)
    assert bytes(out[:outlen]) == hashlib.shake_256(bytes(msg)).digest(outlen)


def test_expand_seed_vs_hashlib():
    seed = rng.integers(0, 256, 64, dtype=np.uint8)
    for dom, idx in ((b"rlwe", 0), (b"rlwe", 15), (b"ksk", 502), (b"ksk", 409501), (b"rgsw1", 2)):
        out = np.zeros(2 * N, np.uint64)
        ol.oracle().orc_expand_seed(seed, dom, idx, out)
        raw = np.frombuffer(hashlib.shake_256(bytes(seed) + dom + idx.to_bytes(8, "little")).digest(2 * N * 8),
                            np.uint64)
        assert np.array_equal(out[:N], raw[:N] % np.uint64(Q0))
        assert np.array_equal(out[N:], raw[N:] % np.uint64(Q1))


# ---------------------------------------------------------------- decomposition
def test_decomp_invariants():
    f = ol.rand_poly(rng).reshape(-1)
    d = np.zeros(2 * 2 * N, np.uint64)
    ol.oracle().orc_decomp(f, d)
    d = d.reshape(2, 2, N)
    Q = Q0 * Q1
    B = 1 << 18
    for i in range(0, N, 7):
        v = int(ol.oracle().orc_crt(int(f[i]), int(f[N + i])))
        assert v % Q0 == f[i] and v % Q1 == f[N + i]  # CRT lift
        c = v if v <= Q // 2 else v - Q               # centring
        digs = []
        for j in range(2):
            r0, r1 = int(d[j, 0, i]), int(d[j, 1, i])
            s = r0 if r0 < Q0 // 2 else r0 - Q0
            assert (s - r1) % Q1 == 0                  # same signed digit in both limbs
            assert -B // 2 <= s < B // 2 + 1
            digs.append(s)
        # digit 0 dropped: c - (d1*B + d2*B^2) is the low digit, |.| <= B/2
        low = c - (digs[0] * B + digs[1] * B * B)
        assert -B // 2 <= low <= B // 2


# ---------------------------------------------------------------- automorphism
def test_automorphism_kat_and_definition():
    f = np.zeros(2 * N, np.uint64)
    f[1] = f[N + 1] = 1
    out = np.zeros(2 * N, np.uint64)
    ol.oracle().orc_automorphism(f, 5, out)  # tau_5(X) = X^5
    assert out[5] == 1 and out[N + 5] == 1 and out.sum() == 2
    g = ol.rand_poly(rng).reshape(-1)
    k = 2 * N - 1
    ol.oracle().orc_automorphism(g, k, out)
    for limb, q in enumerate(QS):
        exp = np.zeros(N, dtype=object)
        for j in range(N):
            e = (j * k) % (2 * N)
            if e < N:
                exp[e] += int(g[limb * N + j])
            else:
                exp[e - N] -= int(g[limb * N + j])
        assert all(int(out[limb * N + i]) == exp[i] % q for i in range(N))


def test_perm_closed_form_matches_automorphism():
    """NTT(tau_t f)[i] = NTT(f)[perm_t(i)] with the closed-form permutation."""
    f = ol.rand_poly(rng).reshape(-1)
    o = ol.oracle()
    a, b = f[:N].copy(), f[N:].copy()
    o.orc_ntt_fwd(a, 0)
    o.orc_ntt_fwd(b, 1)
    for t in (5, 25, 2 * N - 5, 2 * N - 1, pow(5, 1021, 2 * N)):
        g = np.zeros(2 * N, np.uint64)
        o.orc_automorphism(f, t, g)
        ga, gb = g[:N].copy(), g[N:].copy()
        o.orc_ntt_fwd(ga, 0)
        o.orc_ntt_fwd(gb, 1)
        perm = np.array([o.orc_perm_idx(t, i) for i in range(N)])
        assert np.array_equal(ga, a[perm]) and np.array_equal(gb, b[perm])
        if t > N:
            pf = np.array([o.orc_perm_idx(2 * N - t, i) for i in range(N)])
            assert np.array_equal(perm, N - 1 - pf)  # conjugate = reversal


# ---------------------------------------------------------------- setup
@pytest.mark.parametrize("args,rc", [
    ((2**20, 60, 32768), 0), ((2**23, 120, 32768), 0), ((2**27, 120, 32768), 0),
    ((2**29, 120, 32768), 0), ((2**25, 3840, 32768), 0), ((256, 120, 2048), 0),
    ((2**20, 32, 32768), 1), ((2**24, 64, 32768), 1), ((2**28, 64, 32768), 1), ((2**25, 4096, 32768), 1),
])
def test_setup_geometry(args, rc):
    got, geo = ol.setup(*args)
    assert got == rc
    if rc == 0:
        cpe = (args[1] * 8 + 14) // 15
        assert geo["C"] % N == 0 and geo["G"] * N == geo["C"]
        assert geo["C"] * geo["R"] >= args[0] * cpe


# ---------------------------------------------------------------- client / crypto
def test_ext_prod_kat_rgsw_one():
    """dec(ext_prod(RGSW(1), RLWE(42))) = 42 (reference tests/test_crypto.cpp:54-89),
    with RGSW(1) taken from a real query whose j* = 0 (mu = X^0 = 1)."""
    rc, geo = ol.setup(256, 120, 2048)
    seed = np.zeros(64, np.uint8)
    q, st = ol.make_query(geo, seed, 0, 77)
    rg = q[2 * geo["R"] + 2 * BLK:]
    s_ntt = st[2 * N:4 * N]
    # RLWE(42): b = a*s + e + 42*Delta, NTT form
    a = ol.rand_poly(rng).reshape(-1)
    delta = (Q0 * Q1) // P
    b = np.zeros(2 * N, np.uint64)
    for limb, qq in enumerate(QS):
        prod = (a[limb * N:(limb + 1) * N] * s_ntt[limb * N:(limb + 1) * N]) % np.uint64(qq)
        m = np.zeros(N, np.uint64)
        m[0] = 42 * delta % qq
        m[1] = 3  # small noise
        mn = m.copy()
        ol.oracle().orc_ntt_fwd(mn, limb)
        b[limb * N:(limb + 1) * N] = (prod + mn) % np.uint64(qq)
    ct = np.concatenate([a, b])
    res = np.zeros(4 * N, np.uint64)
    ol.oracle().orc_ext_prod(rg, ct, res)
    m = np.zeros(N, np.uint64)
    ol.oracle().orc_rlwe_dec(s_ntt, res, m)
    assert m[0] == 42 and not m[1:].any()


@needs_ref
def test_oracle_primitives_equal_reference_library():
    r, o = ol.ref(), ol.oracle()
    for limb in (0, 1):
        a1 = [np.zeros(N, np.uint64) for _ in range(4)] + [np.zeros(4, np.uint64)]
        a2 = [np.zeros(N, np.uint64) for _ in range(4)] + [np.zeros(4, np.uint64)]
        o.orc_ntt_tables(limb, *a1)
        r.ref_ntt_tables(limb, *a2)
        assert all(np.array_equal(x, y) for x, y in zip(a1, a2))
        x = rng.integers(0, QS[limb], N, dtype=np.uint64)
        y1, y2 = x.copy(), x.copy()
        o.orc_ntt_fwd(y1, limb)
        r.ref_ntt_fwd(y2, limb)
        assert np.array_equal(y1, y2)
        o.orc_ntt_inv(y1, limb)
        r.ref_ntt_inv(y2, limb)
        assert np.array_equal(y1, y2)
    f = ol.rand_poly(rng).reshape(-1)
    d1, d2 = np.zeros(4 * N, np.uint64), np.zeros(4 * N, np.uint64)
    o.orc_decomp(f, d1)
    r.ref_decomp(f, d2)
    assert np.array_equal(d1, d2)
    for k in (5, 25, 2 * N - 1, 3):
        a1, a2 = np.zeros(2 * N, np.uint64), np.zeros(2 * N, np.uint64)
        o.orc_automorphism(f, k, a1)
        r.ref_automorphism(f, k, a2)
        assert np.array_equal(a1, a2)
    seed = rng.integers(0, 256, 64, dtype=np.uint8)
    e1, e2 = np.zeros(2 * N, np.uint64), np.zeros(2 * N, np.uint64)
    o.orc_expand_seed(seed, b"ksk", 502, e1)
    r.ref_expand_seed(seed, b"ksk", 502, e2)
    assert np.array_equal(e1, e2)


@needs_ref
@pytest.mark.parametrize("n,w,rows,idx", [(256, 120, 2048, 42), (262144, 120, 2048, 100000),
                                         (524288, 120, 2048, 524287)])
def test_seeded_query_equals_reference(n, w, rows, idx):
    rc, geo = ol.setup(n, w, rows)
    seed = rng.integers(0, 256, 64, dtype=np.uint8)
    q1, s1 = ol.make_query(geo, seed, idx, 99)
    q2, s2 = np.zeros_like(q1), np.zeros_like(s1)
    ol.ref().ref_query_seeded(geo["R"], geo["C"], geo["D"], w, seed, idx, 99, q2, s2)
    assert np.array_equal(q1, q2) and np.array_equal(s1, s2)


@needs_ref
def test_ext_prod_horner_dec_equal_reference():
    rc, geo = ol.setup(262144, 120, 2048)
    seed = rng.integers(0, 256, 64, dtype=np.uint8)
    q, st = ol.make_query(geo, seed, 1000, 5)
    rg = q[2 * geo["R"] + 2 * BLK:]
    ct = ol.rand_poly(rng, (2,)).reshape(-1)
    a, b = np.zeros(4 * N, np.uint64), np.zeros(4 * N, np.uint64)
    ol.oracle().orc_ext_prod(rg, ct, a)
    ol.ref().ref_ext_prod(rg, ct, b)
    assert np.array_equal(a, b)
    for L in (1, 2, 7):
        cts = ol.rand_poly(rng, (L, 2)).reshape(-1)
        h1, h2 = np.zeros(4 * N, np.uint64), np.zeros(4 * N, np.uint64)
        ol.oracle().orc_horner(cts, L, rg, h1)
        ol.ref().ref_horner(cts, L, rg, h2)
        assert np.array_equal(h1, h2)
    m1, m2 = np.zeros(N, np.uint64), np.zeros(N, np.uint64)
    ol.oracle().orc_rlwe_dec(st[2 * N:4 * N], ct, m1)
    ol.ref().ref_rlwe_dec(st[2 * N:4 * N], ct, m2)
    assert np.array_equal(m1, m2)


@needs_ref
def test_compress_equals_reference():
    ct = ol.rand_poly(rng, (2,)).reshape(-1)
    ob = np.zeros(12288, np.uint8)
    ol.oracle().orc_compress_ct(ct, ob)
    a, b = np.zeros(N, np.uint32), np.zeros(N, np.uint32)
    ol.ref().ref_compress_ct(ct, a, b)
    bits = np.unpackbits(ob)
    va = bits[:28 * N].reshape(N, 28).dot(1 << np.arange(27, -1, -1, dtype=np.uint64))
    vb = bits[28 * N:].reshape(N, 20).dot(1 << np.arange(19, -1, -1, dtype=np.uint64))
    assert np.array_equal(va, a) and np.array_equal(vb, b)


# ---------------------------------------------------------------- server pieces
def test_encode_recovers_slots_under_forward_dft():
    """Horner's decode evaluates sum_m out[m] * X^(2N j m/D): it must give x[j]."""
    rc, geo = ol.setup(262144, 120, 2048)  # G = D_actual = 4
    R, Cc, Da = geo["R"], geo["C"], geo["Da"]
    db = rng.integers(0, P, size=(R, Cc), dtype=np.uint16)
    rows = np.array([0, 7, R - 1], np.uint64)
    enc = db.copy()
    ol.oracle().orc_encode_rows(enc.reshape(-1), R, Cc, geo["D"], rows, len(rows))
    for r in rows:
        r = int(r)
        out = enc[r].reshape(Da, N).astype(np.int64)
        for j in range(Da):
            acc = np.zeros(N, np.int64)
            for m in range(Da):
                e = (2 * N * ((j * m) % Da)) // Da  # multiply out[m] by X^e (negacyclic)
                shifted = np.zeros(N, np.int64)
                idx = np.arange(N) + e
                wrap = (idx // N) % 2 == 1
                shifted[idx % N] = np.where(wrap, -out[m], out[m])
                acc += shifted
            assert np.array_equal(acc % P, db[r, j * N:(j + 1) * N].astype(np.int64) % P)
    # untouched rows stay untouched
    assert np.array_equal(enc[1], db[1])


def negacyclic_mul(a, b, q):
    """Schoolbook product in Z_q[X]/(X^N+1) (object ints)."""
    res = [0] * N
    for i in np.nonzero(a)[0]:
        ai = int(a[i])
        for j in range(N):
            k = i + j
            if k < N:
                res[k] += ai * int(b[j])
            else:
                res[k - N] -= ai * int(b[j])
    return [x % q for x in res]


def test_a32_equals_negacyclic_product():
    """a32 column j = N^-1 * reversed NTT(sum_r col_r * tau_{2N-1}(a_r)):
    checked with a schoolbook product for a sparse column (one row block)."""
    rc, geo = ol.setup(256, 120, 2048)
    R, Cc = geo["R"], geo["C"]
    db = np.zeros((R, Cc), np.uint16)
    j = 5
    nz = [0, 3, 100, 2047]
    for i in nz:
        db[i, j] = 1000 + i
    seed = rng.integers(0, 256, 64, dtype=np.uint8)
    crs = np.zeros(2 * N, np.uint64)
    ol.oracle().orc_crs(seed, 1, crs)
    a32 = np.zeros(N * 2 * N, np.uint64)
    ol.oracle().orc_a32(db.reshape(-1), R, Cc, 0, crs, a32)
    a = np.zeros(2 * N, np.uint64)
    ol.oracle().orc_expand_seed(seed, b"rlwe", 0, a)
    for limb, q in enumerate(QS):
        ar = a[limb * N:(limb + 1) * N]
        rev = np.array([ar[0]] + [(q - ar[N - i]) % q for i in range(1, N)], dtype=np.uint64)
        col = db[:, j].astype(np.uint64)
        prod = np.array(negacyclic_mul(col, rev, q), dtype=np.uint64)
        ol.oracle().orc_ntt_fwd(prod, limb)
        inv_n = pow(N, q - 2, q)
        expect = np.array([int(prod[N - 1 - k]) * inv_n % q for k in range(N)], dtype=np.uint64)
        got = a32.reshape(N, 2, N)[j, limb]
        assert np.array_equal(got, expect)


def test_ring_embed_reformulation_matches_direct_formula():
    """The NTT-along-j ring embed equals the reference's direct 2N-point
    monomial sum (src/gpu_preprocess.cu:39-92) followed by the permutation."""
    a32 = ol.rand_poly(rng, (N,)).reshape(-1)
    af = np.zeros(ol.HALF * 2 * N, np.uint64)
    ac = np.zeros_like(af)
    ol.oracle().orc_ring_embed(a32, af, ac)
    af, ac = af.reshape(ol.HALF, 2, N), ac.reshape(ol.HALF, 2, N)
    picks = [(k, c, i) for k in (0, 1, 2, 511, 1022, 1023) for c in (0, 1) for i in rng.integers(0, N, 5)]
    kci = np.array(picks, dtype=np.uint64).reshape(-1)
    df, dc = np.zeros(len(picks), np.uint64), np.zeros(len(picks), np.uint64)
    ol.oracle().orc_ring_embed_direct(a32, kci, len(picks), df, dc)
    for t, (k, c, i) in enumerate(picks):
        assert af[k, c, i] == df[t] and ac[k, c, i] == dc[t]


def test_matvec_brute_force():
    R, Cc = 2048, 4096
    db = rng.integers(0, P, size=(R, Cc), dtype=np.uint16)
    b0 = rng.integers(0, 2**40, R, dtype=np.uint64)  # non-canonical: reduced first
    b1 = rng.integers(0, Q1, R, dtype=np.uint64)
    out = np.zeros(2 * 64, np.uint64)
    ol.oracle().orc_matvec(db.reshape(-1), R, Cc, b0, b1, 100, 164, out)
    for k, col in enumerate(range(100, 164)):
        assert out[k] == sum(int(db[i, col]) * (int(b0[i]) % Q0) for i in range(R)) % Q0
        assert out[64 + k] == sum(int(db[i, col]) * int(b1[i]) for i in range(R)) % Q1


@pytest.fixture(scope="module")
def small_medium():
    """Oracle preprocess of Small (G=1) and Medium (G=4) databases."""
    seed = np.zeros(64, np.uint8)
    seed[0] = 0x42
    out = {}
    for name, (n, w, rows) in {"small": (256, 120, 2048), "medium": (262144, 120, 2048)}.items():
        rc, geo = ol.setup(n, w, rows)
        db = np.random.default_rng(0).integers(0, P, size=(geo["R"], geo["C"]), dtype=np.uint16)
        out[name] = (geo, db, seed, ol.OraclePrecomp(db, geo, seed))
    return out


@pytest.mark.parametrize("name,idxs", [("small", [0, 42, 255]), ("medium", [0, 1000, 100000, 262143])])
def test_oracle_end_to_end_extract(small_medium, name, idxs):
    """The composed oracle server decrypts to the queried entry (the reference
    e2e contract, tests/test_e2e_gpu.cu:95-104)."""
    geo, db, seed, pc = small_medium[name]
    for t, idx in enumerate(idxs):
        q, st = ol.make_query(geo, seed, idx, 1000 + t)
        resp = pc.answer(q)
        assert np.array_equal(ol.extract(geo, st, resp), ol.expected_slots(geo, db, idx))
