"""The server's collapse pass computes the reference's per-step sum in
dot-product form over int8 byte planes (kernels_collapse.hip). These CPU
checks pin both steps of that reformulation against the oracle's direct
step sum (orc_pack_group, which follows src/gpu_online.cu:547-632):

1. scattering D+/D- to the key slots their steps gather reaches every slot
   once (two per coefficient stay empty) and gives the same sum;
2. the int8 arithmetic (signed byte planes of Dd with the top byte kept,
   balanced base-256 key digits, 7 weight levels, the 128*65793*sum(K)
   correction) reproduces that dot product mod q exactly, including the
   extreme residues q-1.
"""
import numpy as np

import oracle_lib as ol
from oracle_lib import N, NSTEPS, Q0, Q1

QS = (Q0, Q1)
HALF = N // 2
BR = np.array([int(format(x, "011b")[::-1], 2) for x in range(N)], np.int64)


def scatter_dot_form(Dp, Dm):
    """Dd[c][j][p][i] from D+/-[c][s][j][i] (reference order)."""
    Dd = np.zeros((2, 2, N, N), np.uint64)
    hit = np.zeros((N, N), np.int32)  # [p][i] hits per (limb, digit)
    v = 2 * BR + 1
    cols = np.arange(N)
    for s in range(NSTEPS):
        t = pow(5, HALF - 2 - s, 2 * N)
        u = (t * v) % (2 * N)
        pf = BR[(u - 1) // 2]
        pc = N - 1 - pf
        Dd[:, :, pf, cols] = Dp[:, s, :, :]
        Dd[:, :, pc, cols] = Dm[:, s, :, :]
        hit[pf, cols] += 1
        hit[pc, cols] += 1
    return Dd, hit


def dot_mod(Dd_c, K_c, q):
    """sum_{j,p} Dd[j][p][i] K[j][p] mod q, exactly (split K into 14-bit halves)."""
    A = Dd_c.reshape(2 * N, N)
    k = K_c.reshape(2 * N)
    lo, hi = k & np.uint64(0x3FFF), k >> np.uint64(14)
    s_lo = (lo @ A) % np.uint64(q)  # < 2^53 before the reduction
    s_hi = (hi @ A) % np.uint64(q)
    return (s_lo + (s_hi * np.uint64(1 << 14)) % np.uint64(q)) % np.uint64(q)


def int8_form(Dd_c, K_c, q):
    """The kernel's arithmetic: planes x_a, digits d_b, A_ab in int32 range."""
    x = Dd_c.reshape(2 * N, N).astype(np.int64)
    planes = [((x >> (8 * a)) & 255) - 128 for a in range(3)] + [x >> 24]
    v = K_c.reshape(2 * N).astype(np.int64)
    digits = []
    for b in range(3):
        d = ((v & 255) ^ 128) - 128
        digits.append(d)
        v = (v - d) >> 8
    digits.append(v)
    assert all(np.abs(p).max() <= 128 for p in planes) and planes[3].max() < 8
    assert digits[3].min() >= 0 and digits[3].max() <= 8
    lv = [np.zeros(N, np.int64) for _ in range(7)]
    for a in range(4):
        for b in range(4):
            Aab = digits[b] @ planes[a]
            assert np.abs(Aab).max() < 2 ** 31
            lv[a + b] += Aab
    s = np.full(N, 8421504 * (int(K_c.sum()) % q), dtype=object)
    for k in range(7):
        s = s + (pow(256, k, q) * lv[k].astype(object))
    return np.array([int(z) % q for z in s], np.uint64)


def test_collapse_dot_form_matches_oracle_step_sum():
    rng = np.random.default_rng(7)
    Dp = ol.rand_poly(rng, (NSTEPS, 2))          # [s][j][limb][N]
    Dm = ol.rand_poly(rng, (NSTEPS, 2))
    Dp = np.ascontiguousarray(np.transpose(Dp, (2, 0, 1, 3)))  # [limb][s][j][N]
    Dm = np.ascontiguousarray(np.transpose(Dm, (2, 0, 1, 3)))
    Dp[0, 0, 0, :8] = Q0 - 1                     # extreme residues
    Dm[1, 5, 1, :8] = Q1 - 1
    k5 = ol.rand_poly(rng, (2,))                 # client key b-parts, coeff form [j][limb][N]
    k5[0, 0, :16] = Q0 - 1
    zeros2 = np.zeros(2 * N, np.uint64)
    zeros4 = np.zeros(4 * N, np.uint64)
    out = np.zeros(2 * N, np.uint64)
    # interm = 0 and D_final = 0 leave packed_b = -(collapse sum) mod q
    ol.oracle().orc_pack_group(zeros2, Dp.reshape(-1), Dm.reshape(-1), zeros4, k5.reshape(-1), zeros4, out)
    want = out.reshape(2, N)
    K5 = k5.copy()
    for j in range(2):
        for c in range(2):
            vv = K5[j, c].copy()
            ol.oracle().orc_ntt_fwd(vv, c)
            K5[j, c] = vv
    Dd, hit = scatter_dot_form(Dp, Dm)
    assert hit.max() == 1 and (hit.sum(axis=0) == 2 * NSTEPS).all(), "steps reach distinct slots"
    for c in range(2):
        q = QS[c]
        dot = dot_mod(Dd[c], K5[:, c], q)
        assert np.array_equal((np.uint64(q) - dot) % np.uint64(q), want[c]), f"limb {c}: dot form"
        assert np.array_equal(int8_form(Dd[c], K5[:, c], q), dot), f"limb {c}: int8 planes"
