"""ctypes access to the test references: the oracle (oracle/liboracle.so) and, where
it was built, the reference's own CPU library (oracle/_ref/libinspire_ref.so).

Test infrastructure only.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libinspire_ref.so")

N = 2048
Q0, Q1 = 94519297, 95293441
QS = (Q0, Q1)
P = 65535
D_EFF = 2
HALF = N // 2
NSTEPS = HALF - 1
BLK = D_EFF * 2 * N

u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
U64 = C.c_uint64


def _decl(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so missing: run __graft_entry__.build()")
        L = C.CDLL(ORACLE_SO)
        _decl(L, "orc_ntt_tables", None, C.c_int, u64p, u64p, u64p, u64p, u64p)
        _decl(L, "orc_ntt_fwd", None, u64p, C.c_int)
        _decl(L, "orc_ntt_inv", None, u64p, C.c_int)
        _decl(L, "orc_automorphism", None, u64p, U64, u64p)
        _decl(L, "orc_decomp", None, u64p, u64p)
        _decl(L, "orc_crt", U64, U64, U64)
        _decl(L, "orc_shake256", None, u8p, C.c_size_t, u8p, C.c_size_t)
        _decl(L, "orc_expand_seed", None, u8p, C.c_char_p, U64, u64p)
        _decl(L, "orc_ext_prod", None, u64p, u64p, u64p)
        _decl(L, "orc_rlwe_dec", None, u64p, u64p, u64p)
        _decl(L, "orc_perm_idx", C.c_uint32, U64, C.c_uint32)
        _decl(L, "orc_horner", None, u64p, U64, u64p, u64p)
        _decl(L, "orc_setup", C.c_int, U64, U64, U64, u64p)
        _decl(L, "orc_query", None, U64, U64, U64, U64, u8p, U64, U64, u64p, u64p)
        _decl(L, "orc_extract", C.c_int, U64, U64, u64p, u64p, u16p)
        _decl(L, "orc_encode_rows", None, u16p, U64, U64, U64, u64p, U64)
        _decl(L, "orc_crs", None, u8p, U64, u64p)
        _decl(L, "orc_a32", None, u16p, U64, U64, U64, u64p, u64p)
        _decl(L, "orc_ring_embed_direct", None, u64p, u64p, U64, u64p, u64p)
        _decl(L, "orc_ring_embed", None, u64p, u64p, u64p)
        _decl(L, "orc_ksk_a", None, u8p, u64p, u64p)
        _decl(L, "orc_collapse", None, u64p, u64p, u64p, u64p, u64p, u64p, u64p, u64p)
        _decl(L, "orc_preprocess_groups", C.c_int, u16p, U64, U64, U64, u8p, U64, U64, C.c_int,
              u64p, u64p, u64p, u64p)
        _decl(L, "orc_matvec", None, u16p, U64, U64, u64p, u64p, U64, U64, u64p)
        _decl(L, "orc_pack_group", None, u64p, u64p, u64p, u64p, u64p, u64p, u64p)
        _decl(L, "orc_answer", None, u16p, U64, U64, U64, u64p, u64p, u64p, u64p, u64p, u64p,
              C.c_void_p, C.c_void_p)
        _decl(L, "orc_garner", U64, U64, U64)
        _decl(L, "orc_compress_ct", None, u64p, u8p)
        _oracle = L
    return _oracle


def have_ref():
    return os.path.exists(REF_SO)


def ref():
    """The reference's CPU library built from its sources, or None."""
    global _ref
    if _ref is None and have_ref():
        L = C.CDLL(REF_SO)
        _decl(L, "ref_ntt_tables", None, C.c_int, u64p, u64p, u64p, u64p, u64p)
        _decl(L, "ref_ntt_fwd", None, u64p, C.c_int)
        _decl(L, "ref_ntt_inv", None, u64p, C.c_int)
        _decl(L, "ref_automorphism", None, u64p, U64, u64p)
        _decl(L, "ref_automorphism_ntt", None, u64p, U64, u64p)
        _decl(L, "ref_decomp", None, u64p, u64p)
        _decl(L, "ref_expand_seed", None, u8p, C.c_char_p, U64, u64p)
        _decl(L, "ref_ext_prod", None, u64p, u64p, u64p)
        _decl(L, "ref_horner", None, u64p, U64, u64p, u64p)
        _decl(L, "ref_rlwe_dec", None, u64p, u64p, u64p)
        _decl(L, "ref_setup", C.c_int, U64, U64, U64, u64p)
        _decl(L, "ref_query_seeded", None, U64, U64, U64, U64, u8p, U64, U64, u64p, u64p)
        _decl(L, "ref_compress_ct", None, u64p, u32p, u32p)
        _ref = L
    return _ref


# ---------------------------------------------------------------- helpers
def rand_poly(rng, shape=()):
    """Uniform canonical residues, [..., 2, N] (limb-major)."""
    a = np.empty(shape + (2, N), np.uint64)
    a[..., 0, :] = rng.integers(0, Q0, size=shape + (N,), dtype=np.uint64)
    a[..., 1, :] = rng.integers(0, Q1, size=shape + (N,), dtype=np.uint64)
    return a


def reduce_limbs(x):
    """Reduce a [..., 2, N] array limb-wise."""
    x = np.array(x, dtype=np.uint64, copy=True)
    x[..., 0, :] %= Q0
    x[..., 1, :] %= Q1
    return x


def setup(n_entries, w, db_rows):
    out = np.zeros(5, np.uint64)
    rc = oracle().orc_setup(n_entries, w, db_rows, out)
    R, Cc, G, D, T = (int(v) for v in out)
    return rc, dict(R=R, C=Cc, G=G, D=D, T=T, Da=min(D, G), w=w, n_entries=n_entries)


def query_u64s(R):
    return 2 * R + 6 * BLK


def make_query(geo, seed, idx, rng_seed):
    """Seeded oracle query: (flat query u64[...], state u64[4N+2])."""
    q = np.zeros(query_u64s(geo["R"]), np.uint64)
    st = np.zeros(4 * N + 2, np.uint64)
    oracle().orc_query(geo["R"], geo["C"], geo["D"], geo["w"], seed, idx, rng_seed, q, st)
    return q, st


def extract(geo, state, resp):
    cpe = (geo["w"] * 8 + 14) // 15
    slots = np.zeros(cpe, np.uint16)
    rc = oracle().orc_extract(geo["T"], geo["w"], state, np.ascontiguousarray(resp, np.uint64), slots)
    assert rc == 0
    return slots


def expected_slots(geo, db_plain, idx):
    cpe = (geo["w"] * 8 + 14) // 15
    R = geo["R"]
    i_star = idx % R
    col = (idx // R) * cpe
    return db_plain[i_star, col:col + cpe]


class OraclePrecomp:
    """Oracle preprocess of a whole DB: encoded DB + all precomp tensors."""

    def __init__(self, db_plain, geo, seed):
        o = oracle()
        R, Cc, G, D = geo["R"], geo["C"], geo["G"], geo["D"]
        self.db = np.ascontiguousarray(db_plain, np.uint16).copy()
        self.a = np.zeros((G, 2, N), np.uint64)
        self.Dp = np.zeros((G, 2, NSTEPS, D_EFF, N), np.uint64)
        self.Dm = np.zeros((G, 2, NSTEPS, D_EFF, N), np.uint64)
        self.Df = np.zeros((G, 2, D_EFF, N), np.uint64)
        o.orc_preprocess_groups(self.db.reshape(-1), R, Cc, D, seed, 0, G, 1,
                                self.a.reshape(-1), self.Dp.reshape(-1), self.Dm.reshape(-1),
                                self.Df.reshape(-1))
        self.geo = geo

    def answer(self, q, taps=False):
        g = self.geo
        out = np.zeros((g["T"], 4 * N), np.uint64)
        ti = np.zeros(2 * g["C"], np.uint64) if taps else None
        tp = np.zeros(g["G"] * 2 * N, np.uint64) if taps else None
        oracle().orc_answer(self.db.reshape(-1), g["R"], g["C"], g["D"], self.a.reshape(-1),
                            self.Dp.reshape(-1), self.Dm.reshape(-1), self.Df.reshape(-1),
                            np.ascontiguousarray(q, np.uint64), out.reshape(-1),
                            None if ti is None else ti.ctypes.data,
                            None if tp is None else tp.ctypes.data)
        if taps:
            return out, ti, tp
        return out
