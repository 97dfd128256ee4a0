"""Exactness of the batched database pass's CRT-digit arithmetic, on the CPU.

The 17..32-query (and 9..16-query power-row) passes replace each query
row's two residues b_0 (mod Q0) and b_1 (mod Q1) by the CRT value
B in [0, Q0 Q1), expand B into six balanced base-256 digits and a top
digit, and accumulate int8 x int8 products in 32-bit sums over at most
2^16 (per (digit, plane)) or 2^15 (per power of 256) DB rows, then
recombine in 64-bit integers with offsets that keep every partial sum
non-negative. This checks that arithmetic against the direct product
sum_r v_r b_c[r] mod q_c, at the largest magnitudes the kernels can meet
(every DB byte -128 or +127, every digit -128, top digit 32, the full
row count) and on random data, including every bound the recombination
relies on. A wrong digit rule, offset, weight or slice limit fails here.
"""
import numpy as np

Q0, Q1 = 94519297, 95293441
Q1INV = pow(Q1, Q0 - 2, Q0)


def digits7(r0, r1):
    """CRT value of (r0, r1) and its 7 balanced base-256 digits."""
    t = ((r0 - r1 % Q0) % Q0) * Q1INV % Q0
    v = r1 + Q1 * t
    assert v % Q0 == r0 and v % Q1 == r1 and v < Q0 * Q1 < 2 ** 53
    d = []
    for k in range(7):
        dk = ((v & 255) ^ 128) - 128 if k < 6 else v
        v = (v - dk) >> 8 if k < 6 else 0
        d.append(dk)
    assert all(-128 <= x <= 127 for x in d[:6]) and 0 <= d[6] <= 32
    return d


def residue_per_digit(t, Sc, c):
    """sum_k 256^k t_k + 32896 S mod q_c, t_k = 256 H_k + L_k (7-digit form)."""
    qc = Q1 if c else Q0
    lo = t[0] + t[1] * 256 + t[2] * 65536
    assert abs(lo) < 2 ** 54.02
    s = lo + qc * ((1 << 55) // qc) + 32896 * Sc
    assert 0 <= s
    off40 = qc * ((1 << 40) // qc)
    for k in range(3, 7):
        x = t[k] + off40
        assert 0 <= x < 2 ** 64
        s += (x % qc) * pow(256, k, qc)
    assert s < 2 ** 57.1
    return s % qc


def residue_per_power(p, Sc, c):
    """sum_k 256^k p_k + 32896 S mod q_c (power-row form)."""
    qc = Q1 if c else Q0
    lo = p[0] + p[1] * 256 + p[2] * 65536 + p[3] * 16777216
    assert abs(lo) < 2 ** 54.01
    s = lo + qc * ((1 << 55) // qc) + 32896 * Sc
    assert 0 <= s
    off32 = qc * ((1 << 32) // qc)
    for k in range(4, 8):
        y = p[k] + off32
        assert 0 <= y
        s += (y % qc) * pow(256, k, qc)
    assert s < 2 ** 57
    return s % qc


def check(db, b0, b1):
    """Both recombinations against the direct product, for every column."""
    R = db.shape[0]
    D = np.array([digits7(int(x), int(y)) for x, y in zip(b0, b1)], dtype=np.int64)  # R x 7
    hp = (db >> 8) - 128   # the high byte read as int8 after XOR 0x80
    lp = (db & 255) - 128
    H, L = hp.T @ D, lp.T @ D                                       # per (digit, plane)
    Dk = np.concatenate([D, np.zeros((R, 1), np.int64)], 1)          # d_k, d_7 = 0
    Dkm1 = np.concatenate([np.zeros((R, 1), np.int64), D], 1)        # d_{k-1}, d_{-1} = 0
    Pw = hp.T @ Dkm1 + lp.T @ Dk                                     # per power of 256
    assert np.abs(H).max() <= 2 ** 30 and np.abs(L).max() <= 2 ** 30  # i32, 2^16 rows
    S = [int(b0.sum()) % Q0, int(b1.sum()) % Q1]
    for col in range(db.shape[1]):
        for c, (b, q) in enumerate(((b0, Q0), (b1, Q1))):
            want = int(sum(int(x) * int(y) for x, y in zip(db[:, col], b)) % q)
            t = [int(256 * H[col, k] + L[col, k]) for k in range(7)]
            assert residue_per_digit(t, S[c], c) == want
            if R <= 2 ** 15:
                assert np.abs(Pw).max() <= 2 ** 30                  # i32, 2^15 rows
                assert residue_per_power([int(v) for v in Pw[col]], S[c], c) == want
    return Pw


def extreme_b(R):
    B = 2 ** 53 - 0x808080808080  # six balanced digits -128, top digit 32
    assert digits7(B % Q0, B % Q1)[:6] == [-128] * 6
    return np.full(R, B % Q0, np.int64), np.full(R, B % Q1, np.int64)


def test_extreme_values_at_the_slice_limits():
    # 2^15 rows: both forms; DB values 0 (bytes -128, -128), 65535 (127, 127), 255 (-128, 127)
    R = 2 ** 15
    db = np.stack([np.zeros(R, np.int64), np.full(R, 65535, np.int64), np.full(R, 255, np.int64)], 1)
    b0, b1 = extreme_b(R)
    Pw = check(db, b0, b1)
    assert np.abs(Pw).max() == 2 ** 30  # the power-row sums reach the bound exactly
    # 2^16 rows: the per-(digit, plane) form alone (the power rows would pass 2^31)
    R = 2 ** 16
    db = np.zeros((R, 1), np.int64)
    b0, b1 = extreme_b(R)
    check(db, b0, b1)


def test_random_rows_and_max_residues():
    rng = np.random.default_rng(7)
    R = 2048
    db = rng.integers(0, 65536, (R, 4))
    b0 = rng.integers(0, Q0, R)
    b1 = rng.integers(0, Q1, R)
    b0[:64], b1[:64] = Q0 - 1, Q1 - 1
    b0[64:128], b1[64:128] = 0, 0
    check(db, b0, b1)
