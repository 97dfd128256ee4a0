"""ctypes binding of the InsPIRe C ABI (include/capi.h, include/inspire_debug.h).

This is the Python-side binding a caller of the C ABI would write (the same
role a Rust/Go FFI wrapper plays for the reference's capi.h). It loads the
in-tree libraries built by the package Makefile:
  lib/libinspire_amd.so    - server (HIP, gfx950) + client half
  lib/libinspire_client.so - client half only (CPU)
and fails loudly if they are missing: there is no fallback implementation.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
SERVER_SO = os.path.join(LIB_DIR, "libinspire_amd.so")
CLIENT_SO = os.path.join(LIB_DIR, "libinspire_client.so")

N = 2048
D_EFF = 2
BLK = D_EFF * 2 * N


class ipir_params(C.Structure):
    _fields_ = [("n_entries", C.c_size_t), ("entry_bytes", C.c_size_t), ("db_rows", C.c_size_t),
                ("db_cols", C.c_size_t), ("n_packed", C.c_size_t), ("interp_d", C.c_size_t),
                ("num_cts", C.c_size_t), ("ring_n", C.c_size_t), ("d_eff", C.c_size_t),
                ("seed", C.c_uint8 * 64)]


class ipir_caps(C.Structure):
    _fields_ = [("max_batch", C.c_size_t), ("resident_bytes", C.c_size_t),
                ("device_free_bytes", C.c_size_t), ("num_cts", C.c_size_t)]


class ipir_footprint(C.Structure):
    _fields_ = [("resident_fixed", C.c_size_t), ("per_slot", C.c_size_t), ("preprocess_peak", C.c_size_t)]


class ipir_loop_stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("partitions", "slots_per_partition", "submitted", "completed", "failed",
                                          "batches", "max_batch_size", "max_busy_partitions",
                                          "admitted_while_busy", "device_overlaps", "pending",
                                          "busy_partitions")]


class ipir_query_view(C.Structure):
    _fields_ = [("lwe_b_limb0", C.c_void_p), ("lwe_b_limb1", C.c_void_p), ("ksk5_b", C.c_void_p),
                ("kskneg1_b", C.c_void_p), ("rgsw", C.c_void_p)]


P = C.POINTER
VP = C.c_void_p
SZ = C.c_size_t
U64 = C.c_uint64

# (name, restype, argtypes) of every function the headers declare
CLIENT_FUNCS = [
    ("ipir_geometry", C.c_int, [SZ, SZ, SZ, P(ipir_params)]),
    ("ipir_query_u64s", SZ, [P(ipir_params)]),
    ("ipir_response_u64s", SZ, [P(ipir_params)]),
    ("ipir_entry_slots", SZ, [P(ipir_params)]),
    ("ipir_query_view_from_flat", C.c_int, [P(ipir_params), VP, P(ipir_query_view)]),
    ("ipir_query_build", VP, [P(ipir_params), U64, VP]),
    ("ipir_query_build_seeded", VP, [P(ipir_params), U64, U64, VP]),
    ("ipir_extract", C.c_int, [P(ipir_params), VP, VP, VP]),
    ("ipir_client_query_destroy", None, [VP]),
    ("ipir_query_packed_bytes", SZ, [P(ipir_params)]),
    ("ipir_response_packed_bytes", SZ, [P(ipir_params)]),
    ("ipir_query_pack", C.c_int, [P(ipir_params), VP, VP]),
    ("ipir_response_pack", C.c_int, [P(ipir_params), VP, VP]),
    ("ipir_query_unpack", C.c_int, [P(ipir_params), VP, VP]),
    ("ipir_response_unpack", C.c_int, [P(ipir_params), VP, VP]),
    ("ipir_response_compressed_bytes", SZ, [P(ipir_params)]),
    ("ipir_response_compress", C.c_int, [P(ipir_params), VP, VP]),
    ("ipir_extract_compressed", C.c_int, [P(ipir_params), VP, VP, VP]),
]
SERVER_FUNCS = [
    ("ipir_server_create", VP, [SZ, SZ, SZ, SZ, VP, VP, P(ipir_params)]),
    ("ipir_server_caps", None, [VP, P(ipir_caps)]),
    ("ipir_answer_batch", C.c_int, [VP, P(ipir_query_view), SZ, VP]),
    ("ipir_server_destroy", None, [VP]),
    ("ipir_answer_batch_device", C.c_int, [VP, VP, SZ, VP, VP]),
    ("ipir_answer_batch_packed", C.c_int, [VP, VP, SZ, VP]),
    ("ipir_answer_batch_packed_device", C.c_int, [VP, VP, SZ, VP, VP]),
    ("ipir_device_status", C.c_int, [VP]),
    ("ipir_loop_start", VP, [VP, SZ]),
    ("ipir_loop_submit", U64, [VP, P(ipir_query_view), VP]),
    ("ipir_loop_wait", C.c_int, [VP, U64, P(C.c_double)]),
    ("ipir_loop_stop", None, [VP]),
    ("ipir_loop_get_stats", C.c_int, [VP, P(ipir_loop_stats)]),
    ("ipir_server_footprint", C.c_int, [SZ, SZ, SZ, P(ipir_footprint)]),
    ("ipir_fit_max_batch", SZ, [SZ, SZ, SZ, SZ, SZ]),
    ("ipir_dbg_db_block", C.c_int, [VP, SZ, SZ, SZ, SZ, VP]),
    ("ipir_dbg_set_horner_spin_limit", C.c_int, [VP, C.c_uint32]),
    ("ipir_dbg_set_stream_spin_limit", C.c_int, [VP, C.c_uint32]),
    ("ipir_dbg_loop_hold", C.c_int, [VP, C.c_int]),
    ("ipir_dev_stream_create", VP, []),
    ("ipir_dev_stream_sync", C.c_int, [VP]),
    ("ipir_dev_stream_destroy", C.c_int, [VP]),
    ("ipir_dev_mem_info", C.c_int, [P(C.c_size_t), P(C.c_size_t)]),
    ("ipir_dbg_precomp_group_words", SZ, []),
    ("ipir_dbg_precomp_group", C.c_int, [VP, SZ, VP]),
    ("ipir_dbg_db_rows", C.c_int, [VP, SZ, SZ, VP]),
    ("ipir_dbg_db_pass", C.c_int, [VP, VP, SZ, C.c_int, VP]),
    ("ipir_dbg_dbpass_kernel", C.c_int, [VP, SZ, C.c_char_p, SZ]),
    ("ipir_dbg_tuning", C.c_int, [VP, C.c_char_p, SZ]),
    ("ipir_dbg_set_tail_split", C.c_int, [VP, C.c_int]),
    ("ipir_dbg_time_dbpass", C.c_int, [VP, SZ, C.c_int, P(C.c_double)]),
    ("ipir_dbg_answer_taps", C.c_int, [VP, VP, VP, VP, VP]),
    ("ipir_dbg_ntt", C.c_int, [VP, VP, SZ, C.c_int]),
    ("ipir_dbg_ntt_bench", C.c_int, [SZ, C.c_int, C.c_int, P(C.c_double)]),
    ("ipir_dbg_copy_bench", C.c_int, [SZ, C.c_int, P(C.c_double)]),
    ("ipir_dbg_horner", C.c_int, [VP, SZ, VP, VP]),
    ("ipir_dbg_horner_variant", C.c_int, [VP, SZ, VP, VP, C.c_int]),
    ("ipir_horner_noalloc", C.c_int, [VP, SZ, VP, VP, VP]),
    ("ipir_dbg_set_horner", C.c_int, [VP, C.c_int]),
    ("ipir_dbg_set_graph", C.c_int, [VP, C.c_int]),
    ("ipir_dbg_set_pack", C.c_int, [VP, C.c_int]),
    ("ipir_dbg_set_group_chunk", C.c_int, [VP, SZ]),
    ("ipir_dbg_set_mfma_depth", C.c_int, [VP, C.c_uint32, C.c_int]),
    ("ipir_dbg_set_chunk_slices", C.c_int, [VP, SZ, C.c_uint32]),
    ("ipir_dbg_set_late_slices", C.c_int, [VP, SZ, SZ, C.c_uint32]),
    ("ipir_dbg_set_collapse_variant", C.c_int, [VP, C.c_int]),
    ("ipir_dbg_get_collapse_variant", C.c_int, [VP]),
    ("ipir_dbg_set_pipeline", C.c_int, [VP, C.c_int, C.c_int]),
    ("ipir_dbg_set_stream_slices", C.c_int, [VP, SZ, C.c_uint32]),
    ("ipir_dbg_stream_ok", C.c_int, [VP]),
    ("ipir_dbg_set_stream_split", C.c_int, [VP, SZ, C.c_uint32]),
    ("ipir_dbg_take_streaming", C.c_int, [VP]),
    ("ipir_dbg_streamed_count", SZ, [VP]),
    ("ipir_dbg_pass_timing", C.c_int, [VP, C.c_int, P(C.c_double), P(SZ)]),
    ("ipir_dbg_stream_parts", C.c_int, [VP, SZ, C.c_uint32, P(C.c_double)]),
    ("ipir_dbg_profile_batch", C.c_int, [VP, VP, SZ, VP, VP]),
    ("ipir_dev_alloc", VP, [SZ]),
    ("ipir_dev_free", C.c_int, [VP]),
    ("ipir_dev_copy", C.c_int, [VP, VP, SZ, C.c_int]),
    ("ipir_dev_sync", C.c_int, []),
]

_libs = {}


def load(which="server"):
    """Load and declare a library: 'server' (full) or 'client' (CPU half)."""
    if which in _libs:
        return _libs[which]
    path = SERVER_SO if which == "server" else CLIENT_SO
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() (make -C inspire-gpu_amd)")
    lib = C.CDLL(path)
    funcs = CLIENT_FUNCS + (SERVER_FUNCS if which == "server" else [])
    for name, res, args in funcs:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _libs[which] = lib
    return lib


def ptr(a):
    return C.c_void_p(a.ctypes.data)


class DeviceBuffer:
    """Device allocation through the library's own HIP runtime."""

    def __init__(self, nbytes, lib=None):
        self.lib = lib or load("server")
        self.nbytes = nbytes
        self.p = self.lib.ipir_dev_alloc(nbytes)
        if not self.p:
            raise MemoryError(f"ipir_dev_alloc({nbytes}) failed")

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        if self.lib.ipir_dev_copy(self.p, ptr(arr), arr.nbytes, 1):
            raise RuntimeError("upload failed")

    def download(self, arr):
        assert arr.nbytes <= self.nbytes and arr.flags["C_CONTIGUOUS"]
        if self.lib.ipir_dev_copy(ptr(arr), self.p, arr.nbytes, 2):
            raise RuntimeError("download failed")
        return arr

    def free(self):
        if self.p:
            self.lib.ipir_dev_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Params:
    """Published parameters of a server (wraps ipir_params)."""

    def __init__(self, raw):
        self.raw = raw

    def __getattr__(self, k):
        return getattr(self.raw, k)

    @property
    def seed(self):
        return bytes(self.raw.seed)

    def ref(self):
        return C.byref(self.raw)


class Client:
    """Client half over the C ABI (query build / extract / codecs)."""

    def __init__(self, params, lib=None):
        self.lib = lib or load("client")
        self.p = params

    def query_u64s(self):
        return self.lib.ipir_query_u64s(self.p.ref())

    def response_u64s(self):
        return self.lib.ipir_response_u64s(self.p.ref())

    def entry_slots(self):
        return self.lib.ipir_entry_slots(self.p.ref())

    def build(self, idx, rng_seed=None):
        q = np.zeros(self.query_u64s(), np.uint64)
        if rng_seed is None:
            h = self.lib.ipir_query_build(self.p.ref(), idx, ptr(q))
        else:
            h = self.lib.ipir_query_build_seeded(self.p.ref(), idx, rng_seed, ptr(q))
        if not h:
            raise ValueError("ipir_query_build failed")
        return h, q

    def pack_query(self, q):
        out = np.zeros(self.lib.ipir_query_packed_bytes(self.p.ref()), np.uint8)
        if self.lib.ipir_query_pack(self.p.ref(), ptr(np.ascontiguousarray(q, np.uint64)), ptr(out)):
            raise ValueError("ipir_query_pack failed")
        return out

    def compress(self, resp):
        out = np.zeros(self.lib.ipir_response_compressed_bytes(self.p.ref()), np.uint8)
        if self.lib.ipir_response_compress(self.p.ref(), ptr(np.ascontiguousarray(resp, np.uint64)), ptr(out)):
            raise ValueError("ipir_response_compress failed")
        return out

    def extract_compressed(self, h, buf):
        out = np.zeros(self.entry_slots(), np.uint16)
        rc = self.lib.ipir_extract_compressed(self.p.ref(), h, ptr(np.ascontiguousarray(buf, np.uint8)), ptr(out))
        if rc:
            raise ValueError(f"ipir_extract_compressed rc={rc}")
        return out

    def extract(self, h, resp):
        out = np.zeros(self.entry_slots(), np.uint16)
        rc = self.lib.ipir_extract(self.p.ref(), h, ptr(np.ascontiguousarray(resp, np.uint64)), ptr(out))
        if rc:
            raise ValueError(f"ipir_extract rc={rc}")
        return out

    def destroy(self, h):
        self.lib.ipir_client_query_destroy(h)


def geometry(n_entries, entry_bytes, db_rows=0, lib=None):
    """setup()'s geometry as a dict (R, C, G, D, T, cpe), or ValueError."""
    raw = ipir_params()
    if (lib or load("client")).ipir_geometry(n_entries, entry_bytes, db_rows, C.byref(raw)):
        raise ValueError("setup rejects this geometry")
    return dict(R=raw.db_rows, C=raw.db_cols, G=raw.n_packed, D=raw.interp_d, T=raw.num_cts,
                cpe=(entry_bytes * 8 + 14) // 15)


def footprint(n_entries, entry_bytes, db_rows=0, lib=None):
    """Device bytes of a server of this geometry: dict(resident_fixed,
    per_slot, preprocess_peak), or ValueError if setup() rejects it."""
    f = ipir_footprint()
    if (lib or load("server")).ipir_server_footprint(n_entries, entry_bytes, db_rows, C.byref(f)):
        raise ValueError("setup rejects this geometry")
    return dict(resident_fixed=f.resident_fixed, per_slot=f.per_slot, preprocess_peak=f.preprocess_peak)


def mem_info(lib=None):
    """(free, total) device bytes now."""
    fr, tot = C.c_size_t(), C.c_size_t()
    if (lib or load("server")).ipir_dev_mem_info(C.byref(fr), C.byref(tot)):
        raise RuntimeError("ipir_dev_mem_info failed")
    return fr.value, tot.value


def make_params(n_entries, entry_bytes, db_rows, seed, R, Cc, G, D, T):
    raw = ipir_params()
    raw.n_entries, raw.entry_bytes, raw.db_rows, raw.db_cols = n_entries, entry_bytes, R, Cc
    raw.n_packed, raw.interp_d, raw.num_cts, raw.ring_n, raw.d_eff = G, D, T, N, D_EFF
    for i, b in enumerate(bytes(seed)):
        raw.seed[i] = b
    return Params(raw)


class Server:
    """One server (database generation) over the C ABI."""

    def __init__(self, db, n_entries, entry_bytes, db_rows=0, max_batch=16, crs_seed=None):
        self.lib = load("server")
        db = np.ascontiguousarray(db, np.uint16)
        raw = ipir_params()
        seed = None
        if crs_seed is not None:
            seed = (C.c_uint8 * 64)(*bytes(crs_seed))
        h = self.lib.ipir_server_create(n_entries, entry_bytes, db_rows, max_batch, ptr(db),
                                        C.cast(seed, C.c_void_p) if seed is not None else None,
                                        C.byref(raw))
        if not h:
            raise RuntimeError("ipir_server_create failed")
        self.h = h
        self.params = Params(raw)

    def caps(self):
        c = ipir_caps()
        self.lib.ipir_server_caps(self.h, C.byref(c))
        return c

    def views(self, queries):
        arr = (ipir_query_view * len(queries))()
        R = self.params.db_rows
        for i, q in enumerate(queries):
            base = q.ctypes.data
            arr[i].lwe_b_limb0 = base
            arr[i].lwe_b_limb1 = base + 8 * R
            arr[i].ksk5_b = base + 8 * 2 * R
            arr[i].kskneg1_b = base + 8 * (2 * R + BLK)
            arr[i].rgsw = base + 8 * (2 * R + 2 * BLK)
        return arr

    def answer_batch(self, queries):
        queries = [np.ascontiguousarray(q, np.uint64) for q in queries]
        T = self.params.num_cts
        out = np.zeros((len(queries), T * 4 * N), np.uint64)
        rc = self.lib.ipir_answer_batch(self.h, self.views(queries), len(queries), ptr(out))
        if rc:
            raise RuntimeError(f"ipir_answer_batch rc={rc}")
        return out

    def block(self, row0, nrows, col0, ncols):
        """Encoded DB rows [row0, row0+nrows) x columns [col0, col0+ncols), row-major u16."""
        out = np.zeros((nrows, ncols), np.uint16)
        if self.lib.ipir_dbg_db_block(self.h, row0, nrows, col0, ncols, ptr(out)):
            raise RuntimeError("ipir_dbg_db_block failed")
        return out

    def precomp_group(self, g):
        """Group g's precomp in the reference layout: dict(a, Dp, Dm, Df) as u64
        arrays shaped like the oracle's ([2][N], [2][N/2-1][2][N], ..., [2][2][N])."""
        words = self.lib.ipir_dbg_precomp_group_words()
        blk = np.zeros(words, np.uint32)
        if self.lib.ipir_dbg_precomp_group(self.h, g, ptr(blk)):
            raise RuntimeError("ipir_dbg_precomp_group failed")
        dsz = 2 * (N // 2 - 1) * D_EFF * N
        return dict(a=blk[:2 * N].astype(np.uint64), Dp=blk[2 * N:2 * N + dsz].astype(np.uint64),
                    Dm=blk[2 * N + dsz:2 * N + 2 * dsz].astype(np.uint64),
                    Df=blk[2 * N + 2 * dsz:].astype(np.uint64))

    def answer_packed(self, packed):
        """Packed wire queries (ipir_query_pack bytes, one row per query) ->
        compressed responses (num_cts * 12288 bytes per query)."""
        packed = np.ascontiguousarray(packed, np.uint8)
        T = self.params.num_cts
        out = np.zeros((packed.shape[0], T * 12288), np.uint8)
        rc = self.lib.ipir_answer_batch_packed(self.h, ptr(packed), packed.shape[0], ptr(out))
        if rc:
            raise RuntimeError(f"ipir_answer_batch_packed rc={rc}")
        return out

    def close(self):
        if self.h:
            self.lib.ipir_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
