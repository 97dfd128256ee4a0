"""Shared pieces of the GPU tests at the benchmark geometries (db_rows
32768): seeded parallel database generation, host-memory guards, oracle
work spread over host threads, and progress lines that name each step.

Test infrastructure only.
"""
import concurrent.futures as cf
import os
import sys
import time

import numpy as np
import pytest

import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SEED = np.array([0x42] + [0] * 63, np.uint8)
THREADS = min(16, os.cpu_count() or 1)
_T0 = time.time()


def say(*a):
    """Progress line on stdout and, when the repository has a run-output
    folder, appended to a progress log in it (so a long step is seen
    working)."""
    msg = f"[{time.time() - _T0:7.1f} s] " + " ".join(str(x) for x in a)
    print(msg, flush=True)
    sys.stdout.flush()
    d = os.path.join(ROOT, "gpurun_out")
    if os.path.isdir(d):
        with open(os.path.join(d, "gpu_tests_progress.log"), "a") as f:
            f.write(msg + "\n")


def host_available_bytes():
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemAvailable:"):
                return int(line.split()[1]) * 1024
    except OSError:
        pass
    return 0


def need_host(nbytes, what):
    avail = host_available_bytes()
    if avail and avail < nbytes:
        pytest.skip(f"{what}: needs {nbytes / 2**30:.0f} GiB of host memory, {avail / 2**30:.0f} GiB available")


def gen_db(R, Cc, seed=0):
    """Uniform u16 in [0, 65535) (the reference's value range), row chunks
    from SeedSequence(seed).spawn, generated on THREADS host threads."""
    db = np.empty((R, Cc), np.uint16)
    rows = max(1, (1 << 26) // Cc)  # ~128 MiB per chunk
    starts = list(range(0, R, rows))
    kids = np.random.SeedSequence(seed).spawn(len(starts))

    def fill(k):
        r0 = starts[k]
        g = np.random.default_rng(kids[k])
        db[r0:r0 + rows] = g.integers(0, 65535, size=(min(rows, R - r0), Cc), dtype=np.uint16)
    with cf.ThreadPoolExecutor(THREADS) as ex:
        list(ex.map(fill, range(len(starts))))
    return db


def pmap(fn, items):
    """fn over items on host threads (oracle calls release the GIL)."""
    with cf.ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(fn, items))


def oracle_precomp_groups(db_enc, geo, groups, seed=SEED):
    """Oracle precomp (a, D+, D-, D_final) of the listed groups of an
    ENCODED row-major DB, one group per host thread: dict g -> tensors."""
    R, Cc, D = geo["R"], geo["C"], geo["D"]
    dsz = 2 * ol.NSTEPS * ol.D_EFF * ol.N

    def one(g):
        a = np.zeros(2 * ol.N, np.uint64)
        Dp = np.zeros(dsz, np.uint64)
        Dm = np.zeros(dsz, np.uint64)
        Df = np.zeros(2 * ol.D_EFF * ol.N, np.uint64)
        ol.oracle().orc_preprocess_groups(db_enc.reshape(-1), R, Cc, D, seed, g, g + 1, 0, a, Dp, Dm, Df)
        return g, dict(a=a, Dp=Dp, Dm=Dm, Df=Df)
    return dict(pmap(one, list(groups)))


def oracle_encode(db_plain, geo):
    """The oracle's encode of a whole row-major DB (a copy), rows spread over
    host threads."""
    enc = np.ascontiguousarray(db_plain).copy()
    if geo["Da"] <= 1:
        return enc
    R, Cc, D = geo["R"], geo["C"], geo["D"]
    chunks = np.array_split(np.arange(R, dtype=np.uint64), THREADS * 4)
    pmap(lambda rows: ol.oracle().orc_encode_rows(enc.reshape(-1), R, Cc, D, np.ascontiguousarray(rows), len(rows)),
         chunks)
    return enc


def oracle_answer(db_enc, geo, pre, q, taps=False):
    """orc_answer against an encoded DB and per-group tensors {g: dict}."""
    G = geo["G"]
    a = np.stack([pre[g]["a"] for g in range(G)])
    Dp = np.stack([pre[g]["Dp"] for g in range(G)])
    Dm = np.stack([pre[g]["Dm"] for g in range(G)])
    Df = np.stack([pre[g]["Df"] for g in range(G)])
    out = np.zeros((geo["T"], 4 * ol.N), np.uint64)
    ti = np.zeros(2 * geo["C"], np.uint64) if taps else None
    tp = np.zeros(G * 2 * ol.N, np.uint64) if taps else None
    ol.oracle().orc_answer(db_enc.reshape(-1), geo["R"], geo["C"], geo["D"], a.reshape(-1), Dp.reshape(-1),
                           Dm.reshape(-1), Df.reshape(-1), np.ascontiguousarray(q, np.uint64), out.reshape(-1),
                           None if ti is None else ti.ctypes.data, None if tp is None else tp.ctypes.data)
    return (out, ti, tp) if taps else out


def random_indices(n_entries, k, seed=1):
    return [int(i) for i in np.random.default_rng(seed).integers(0, n_entries, k)]


def ct_of(geo, idx):
    """Response ciphertext that carries entry idx (its column's group / D_a)."""
    cpe = (geo["w"] * 8 + 14) // 15
    col = (idx // geo["R"]) * cpe
    return (col // ol.N) // geo["Da"]
