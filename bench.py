"""Benchmark of the InsPIRe PIR server on one MI355X.

Metric: queries/s at the 16 GB database (2^27 entries x 120 B, db_rows
32768, 128 groups), batch 32 versus batch 1, responses bit-exact; plus
p50/p99 batch latency. One "step" is one answer of a batch of queries whose
flat wire queries are already resident in HBM (the timed region covers
on-device unmarshal, key NTTs, the database pass, the precomp-tensor pass,
pack and Horner, writing flat u64 responses to HBM).

Synthetic data: DB values uniform in [0, 65535) from seed 0, query indices
uniform from seed 1, client secrets from seed 2 (seeded client query).

Prints one JSON line (see the repository's benchmark contract).
"""
import argparse
import ctypes as C
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "inspire-gpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

CONFIGS = {
    # name: (n_entries, entry_bytes, db_rows, label)
    "c1": (2**20, 60, 32768, "plain small, 2^20 x 60 B (0.13 GB DB)"),
    "c2": (2**23, 120, 32768, "1 GB DB, 2^23 x 120 B"),
    "c3": (2**27, 120, 32768, "16 GB DB, 2^27 x 120 B"),
    "c4": (2**29, 120, 32768, "64 GB DB, 2^29 x 120 B"),
    "c5": (2**25, 3840, 32768, "128 GB DB, 2^25 x 3840 B (wide records)"),
}
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gen_db(R, Cc, seed=0, threads=16):
    """Uniform u16 in [0, 65535), row chunks from SeedSequence(seed).spawn,
    generated on host threads (numpy's generators release the GIL)."""
    import concurrent.futures as cf
    db = np.empty((R, Cc), np.uint16)
    rows = max(1, (1 << 26) // Cc)  # ~128 MiB per chunk
    starts = list(range(0, R, rows))
    kids = np.random.SeedSequence(seed).spawn(len(starts))

    def fill(k):
        g = np.random.default_rng(kids[k])
        db[starts[k]:starts[k] + rows] = g.integers(0, 65535, size=(min(rows, R - starts[k]), Cc), dtype=np.uint16)
    with cf.ThreadPoolExecutor(min(threads, os.cpu_count() or 1)) as ex:
        list(ex.map(fill, range(len(starts))))
    return db


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cpus():
    """(logical CPUs of the machine, CPUs this process may run on)."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return os.cpu_count() or 1, allowed


def ref_primitives_us():
    """The reference's own CPU primitives (its ntt.cpp / ring.cpp /
    crypto.cpp compiled from source into oracle/_ref when the build saw the
    reference tree), one thread, microseconds per call; None if absent."""
    import oracle_lib as ol
    if not ol.have_ref():
        return None
    r = ol.ref()
    N = ol.N
    rng = np.random.default_rng(6)
    poly = ol.rand_poly(rng).reshape(-1)

    def us(fn, reps):
        fn()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        return round((time.perf_counter() - t) / reps * 1e6, 1)
    a = poly[:N].copy()
    out = {"ntt_fwd_q0": us(lambda: r.ref_ntt_fwd(a, 0), 50), "ntt_fwd_q1": us(lambda: r.ref_ntt_fwd(a, 1), 50),
           "ntt_inv_q0": us(lambda: r.ref_ntt_inv(a, 0), 50), "ntt_inv_q1": us(lambda: r.ref_ntt_inv(a, 1), 50)}
    dec = np.zeros(2 * 2 * N, np.uint64)
    out["decomp"] = us(lambda: r.ref_decomp(poly, dec), 50)
    rg = ol.rand_poly(rng, (4, 2)).reshape(-1)
    ct = ol.rand_poly(rng, (2,)).reshape(-1)
    res = np.zeros(4 * N, np.uint64)
    out["ext_prod"] = us(lambda: r.ref_ext_prod(rg, ct, res), 10)
    return out


def cpu_baseline(geo, n_groups_sample=8, sample_cols=65536, horner_steps=16, threads=None):
    """Time the oracle (scalar C++ restatement of the reference's CPU
    primitives) on the host cores over a bounded sample of one query and
    scale to a whole query: the mat-vec over sample_cols columns (all rows)
    split across `threads` workers, the collapse+pack of groups spread over
    the workers, the Horner chains (one per response ct, in parallel).
    Also one-thread figures and per-primitive times."""
    import concurrent.futures as cf
    import oracle_lib as ol
    o = ol.oracle()
    N = ol.N
    threads = threads or min(16, os.cpu_count() or 1)
    R, Cc, G, Da, T = geo["R"], geo["C"], geo["G"], geo["Da"], geo["T"]
    rng = np.random.default_rng(5)
    cw = min(Cc, sample_cols)
    dbs = rng.integers(0, 65535, size=(R, cw), dtype=np.uint16)
    flat = dbs.reshape(-1)
    b0 = rng.integers(0, ol.Q0, R, dtype=np.uint64)
    b1 = rng.integers(0, ol.Q1, R, dtype=np.uint64)
    pool = cf.ThreadPoolExecutor(threads)  # ctypes calls release the GIL

    def mv(c0, c1):
        out = np.zeros(2 * (c1 - c0), np.uint64)
        o.orc_matvec(flat, R, cw, b0, b1, c0, c1, out)
    part = max(1, cw // threads)
    mv(0, 8)  # first call: the oracle's tables
    t = time.perf_counter()
    mv(0, part)
    t_mv1 = (time.perf_counter() - t) * Cc / part
    t = time.perf_counter()
    list(pool.map(lambda k: mv(k * part, min(cw, (k + 1) * part)), range(threads)))
    t_mv = (time.perf_counter() - t) * Cc / (part * threads)
    Dp = ol.rand_poly(rng, (ol.NSTEPS, 2))  # [step][digit][limb] shapes; values are irrelevant to timing
    Dp = np.ascontiguousarray(np.transpose(Dp, (2, 0, 1, 3))).reshape(-1)
    Df = ol.rand_poly(rng, (2,)).reshape(-1)
    k5 = ol.rand_poly(rng, (2,)).reshape(-1)
    interm = ol.rand_poly(rng).reshape(-1)

    def pack(_):
        ob = np.zeros(2 * N, np.uint64)
        o.orc_pack_group(interm, Dp, Dp, Df, k5, k5, ob)
    pack(0)
    t = time.perf_counter()
    pack(0)
    t_pack1 = (time.perf_counter() - t) * G
    ng = max(n_groups_sample, threads)
    t = time.perf_counter()
    list(pool.map(pack, range(ng)))
    t_pack = (time.perf_counter() - t) / ng * G
    rg = ol.rand_poly(rng, (4, 2)).reshape(-1)
    cts = ol.rand_poly(rng, (horner_steps + 1, 2)).reshape(-1)
    ho = np.zeros(4 * N, np.uint64)
    t = time.perf_counter()
    o.orc_horner(cts, horner_steps + 1, rg, ho)
    step = (time.perf_counter() - t) / horner_steps
    t_h1 = step * max(0, Da - 1) * T
    t_h = step * max(0, Da - 1) * -(-T // threads)
    pool.shutdown()
    # per-primitive (one thread), microseconds
    prim = {}
    poly = ol.rand_poly(rng).reshape(-1)

    def us(fn, reps):
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        return round((time.perf_counter() - t) / reps * 1e6, 1)
    a = poly[:N].copy()
    prim["ntt_fwd_q0"] = us(lambda: o.orc_ntt_fwd(a, 0), 50)
    prim["ntt_fwd_q1"] = us(lambda: o.orc_ntt_fwd(a, 1), 50)
    prim["ntt_inv_q0"] = us(lambda: o.orc_ntt_inv(a, 0), 50)
    prim["ntt_inv_q1"] = us(lambda: o.orc_ntt_inv(a, 1), 50)
    dec = np.zeros(2 * 2 * N, np.uint64)
    prim["decomp"] = us(lambda: o.orc_decomp(poly, dec), 50)
    ct = ol.rand_poly(rng, (2,)).reshape(-1)
    res = np.zeros(4 * N, np.uint64)
    prim["ext_prod"] = us(lambda: o.orc_ext_prod(rg, ct, res), 10)
    prim[f"horner_eval_Da{Da}"] = round(step * max(0, Da - 1) * 1e6, 1)
    total = t_mv + t_pack + t_h
    total1 = t_mv1 + t_pack1 + t_h1
    ncpu, allowed = host_cpus()
    refp = ref_primitives_us()
    return {
        "value": 1.0 / total, "unit": "queries/s", "cores": threads, "kind": "port",
        "host": {"cpu_model": cpu_model(), "logical_cpus": ncpu, "cpus_allowed": allowed, "threads_used": threads},
        "sample": (f"EXTRAPOLATED from a bounded sample: the oracle (scalar C++ restatement of the reference CPU "
                   f"primitives) on {threads} worker threads of {ncpu} logical CPUs ({cpu_model()}) times the "
                   f"mat-vec over {cw} of {Cc} columns, collapse+pack of {ng} of {G} groups and {horner_steps} of "
                   f"{Da - 1} Horner steps, each scaled to one whole query (mat-vec {t_mv:.2f} s, collapse "
                   f"{t_pack:.2f} s, Horner {t_h:.2f} s)"),
        "single_thread_qps": 1.0 / total1,
        "primitives_us": prim,
        "reference_primitives_us": refp,
        "reference_primitives_note": ("the reference's own ntt.cpp/ring.cpp/crypto.cpp compiled from source, one "
                                      "thread on this host" if refp else "reference CPU library not built"),
    }


def int8_macs(nq, R, Cc, kernel=""):
    """int8 MACs the MFMA database pass executes for nq queries: 16-row digit
    tiles of 2 queries (4 digits x 2 limbs each) padded to the kernel's tile
    count, or 14 tiles of 7 CRT digits for 32 queries (the kernels built with
    14 tiles), times the DB's two byte planes, per row and column."""
    tiles = 1 if nq <= 2 else 2 if nq <= 4 else 4 if nq <= 8 else 8 if nq <= 16 else 16
    if kernel.startswith("k_dbpass_mfma_sa<14,"):
        tiles = 14
    return 16.0 * tiles * 2 * R * Cc


def serving(lib, srv, qs, capacity_qps, resp_u64s, rates=(0.25, 0.5, 0.9), seconds=3.0, max_inflight=2):
    """Poisson arrivals through the continuous-batching loop (ipir_loop_*):
    at each rate (a fraction of the measured batched capacity) submit
    requests at exponential inter-arrival times for `seconds`, wait for
    each, and report throughput and latency percentiles (submit to
    response written)."""
    import threading
    import inspire_amd as ia
    views = [srv.views([q]) for q in qs]
    ring = 1024
    outs = np.zeros((ring, resp_u64s), np.uint64)
    results = []
    rng = np.random.default_rng(3)
    for rate in rates:
        lam = rate * capacity_qps
        n = max(50, int(lam * seconds))
        gaps = rng.exponential(1.0 / lam, n)
        loop = lib.ipir_loop_start(srv.h, max_inflight)
        assert loop
        st0 = ia.ipir_loop_stats()
        lib.ipir_loop_get_stats(loop, C.byref(st0))
        tickets = [None] * n
        lat = np.zeros(n)
        done_t = [0.0]
        ready = threading.Semaphore(0)

        def waiter():
            for k in range(n):
                ready.acquire()
                l = C.c_double()
                rc = lib.ipir_loop_wait(loop, tickets[k], C.byref(l))
                assert rc == 0
                lat[k] = l.value * 1e-3
            done_t[0] = time.perf_counter()

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        t_next = t0
        for k in range(n):
            t_next += gaps[k]
            while True:
                d = t_next - time.perf_counter()
                if d <= 0:
                    break
                time.sleep(min(d, 0.0005) if d > 0.0002 else 0)
            tickets[k] = lib.ipir_loop_submit(loop, views[k % len(views)], ia.ptr(outs[k % ring]))
            assert tickets[k]
            ready.release()
        th.join()
        st = ia.ipir_loop_stats()
        lib.ipir_loop_get_stats(loop, C.byref(st))
        lib.ipir_loop_stop(loop)
        wall = done_t[0] - t0
        results.append({"offered_qps": round(lam, 1), "achieved_qps": round(n / wall, 1), "requests": n,
                        "p50_ms": round(float(np.percentile(lat, 50)), 3),
                        "p99_ms": round(float(np.percentile(lat, 99)), 3),
                        "p999_ms": round(float(np.percentile(lat, 99.9)), 3),
                        "partitions": int(st0.partitions), "slots_per_partition": int(st0.slots_per_partition),
                        "batches": int(st.batches), "mean_batch": round(n / max(1, st.batches), 2),
                        "max_busy_partitions": int(st.max_busy_partitions),
                        "batches_admitted_while_another_in_flight": int(st.admitted_while_busy),
                        "batches_launched_beside_device_work": int(st.device_overlaps)})
        log(f"[bench] serving at {rate:.2f}x capacity: {results[-1]}")
    return results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--batch", type=int, default=32,
                    help="queries per batch (max_batch of the server); 0: the most slots that fit the device's free "
                         "memory (ipir_fit_max_batch, 4 GiB kept free), at most 128")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verify", type=int, default=0,
                    help="also answer and extract this many further random queries (batches of --batch)")
    ap.add_argument("--serving-inflight", type=int, default=0,
                    help="serving loop partitions (batches in flight), each of max_batch / N slots; "
                         "0: as many 32-slot partitions as the batch holds, at most 4")
    ap.add_argument("--serving", action="store_true",
                    help="also run Poisson arrivals through the serving loop at 0.25/0.5/0.9x capacity")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run a plain device copy of this many GiB (3 launches) and print its rate: the "
                         "known-size reference for the FETCH_SIZE/WRITE_SIZE counter scale")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        import inspire_amd as ia
        nbytes = int(args.copy_calibration * 2 ** 30) // 16 * 16
        ms = C.c_double()
        assert ia.load("server").ipir_dbg_copy_bench(nbytes, 3, C.byref(ms)) == 0
        print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": round(ms.value, 4),
                          "GBs_read_plus_write": round(2 * nbytes / (ms.value * 1e-3) / 1e9, 1)}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported "
                                   "(queries shard by replica or by group, see the design notes)"}))
        sys.exit(2)

    import inspire_amd as ia

    n, w, rows, label = CONFIGS[args.config]
    geo = ia.geometry(n, w, rows)
    R, Cc, G, T = geo["R"], geo["C"], geo["G"], geo["T"]
    geo["Da"] = min(geo["D"], G)
    B = args.batch
    fitted = None
    if B <= 0:
        B = int(ia.load("server").ipir_fit_max_batch(n, w, rows, 4 * 2 ** 30, 128))
        fp = ia.footprint(n, w, rows)
        fitted = {"max_batch": B, "free_bytes_at_start": ia.mem_info()[0], "reserve_bytes": 4 * 2 ** 30,
                  "resident_fixed": fp["resident_fixed"], "per_slot": fp["per_slot"],
                  "preprocess_peak": fp["preprocess_peak"]}
        log(f"[bench] max_batch {B} fits the free device memory ({fitted})")
        if B < 1:
            print(json.dumps({"error": "the server does not fit the device's free memory"}))
            sys.exit(3)
    log(f"[bench] {label}: R={R} C={Cc} G={G} T={T}; generating DB ({R * Cc * 2 / 1e9:.2f} GB)")
    t0 = time.time()
    db = gen_db(R, Cc, 0)
    log(f"[bench] DB generated in {time.time() - t0:.1f} s; preprocessing")
    seed = bytes(np.random.default_rng(2).integers(0, 256, 64, dtype=np.uint8))
    t0 = time.time()
    srv = ia.Server(db, n, w, rows, max_batch=B, crs_seed=seed)
    t_pre = time.time() - t0
    log(f"[bench] preprocess + setup {t_pre:.1f} s")
    lib = srv.lib
    idx = np.random.default_rng(1).integers(0, n, B)
    cl = ia.Client(srv.params, lib)
    qs, hs, build_ms = [], [], []
    for k, i in enumerate(idx):
        t = time.perf_counter()
        h, q = cl.build(int(i), 2 * 1000003 + k)
        build_ms.append((time.perf_counter() - t) * 1e3)
        qs.append(q)
        hs.append(h)
    client = {"query_build_ms_first": round(build_ms[0], 2),
              "query_build_ms_cached_p50": round(float(np.median(build_ms[1:])) if B > 1 else build_ms[0], 2),
              "note": "client CPU, one thread: the first query of a process expands the seed's CRS polynomials, "
                      "later ones reuse them (not part of the server metric)"}
    log(f"[bench] client query build: {client}")
    flat = np.ascontiguousarray(np.stack(qs))
    qlen = flat.shape[1]
    resp_u64s = T * 4 * ia.N
    dq = ia.DeviceBuffer(flat.nbytes, lib)
    dq.upload(flat)
    dout = ia.DeviceBuffer(B * resp_u64s * 8, lib)

    def step(count):
        rc = lib.ipir_answer_batch_device(srv.h, C.c_void_p(dq.p), count, C.c_void_p(dout.p), None)
        assert rc == 0

    # correctness of this run: responses extract to the DB entries
    step(B)
    lib.ipir_dev_sync()
    resp = dout.download(np.zeros((B, resp_u64s), np.uint64))
    def entry(i):  # the u16 slots of entry i in the row-major DB (src/protocol.cpp index math)
        col = (i // R) * geo["cpe"]
        return db[i % R, col:col + geo["cpe"]]

    ok = all(np.array_equal(cl.extract(hs[k], resp[k]), entry(int(idx[k]))) for k in range(B))
    ref1 = srv.answer_batch([qs[0]])[0]
    ok = ok and np.array_equal(ref1, resp[0])
    log(f"[bench] extract check ({B} queries) and batch==single: {'ok' if ok else 'FAILED'}")
    verified = None
    if args.verify > 0:
        vr = np.random.default_rng(11).integers(0, n, args.verify)
        bad = 0
        t0 = time.time()
        for c0 in range(0, args.verify, B):
            chunk = vr[c0:c0 + B]
            built = [cl.build(int(i), 3 * 1000003 + c0 + k) for k, i in enumerate(chunk)]
            dq.upload(np.ascontiguousarray(np.stack([q for _, q in built])))
            step(len(chunk))
            lib.ipir_dev_sync()
            r = dout.download(np.zeros((len(chunk), resp_u64s), np.uint64))
            for k, (h, _) in enumerate(built):
                bad += not np.array_equal(cl.extract(h, r[k]), entry(int(chunk[k])))
                cl.destroy(h)
        verified = {"queries": int(args.verify), "mismatches": int(bad)}
        ok = ok and bad == 0
        log(f"[bench] verified {args.verify} further random queries: {bad} mismatches ({time.time() - t0:.1f} s)")
        dq.upload(flat)
    del db

    def timed(count, steps, warmup):
        for _ in range(warmup):
            step(count)
        lib.ipir_dev_sync()
        lat = []
        t_all = time.perf_counter()
        for _ in range(steps):
            t = time.perf_counter()
            step(count)
            lib.ipir_dev_sync()
            lat.append(time.perf_counter() - t)
        total = time.perf_counter() - t_all
        return total, np.array(lat)

    tot32, lat32 = timed(B, args.steps, args.warmup)
    tot1, lat1 = timed(1, max(args.steps, 10), args.warmup)
    qps32 = B * args.steps / tot32
    qps1 = max(args.steps, 10) / tot1
    log(f"[bench] B={B}: {qps32:.1f} q/s ({tot32 / args.steps * 1e3:.2f} ms/batch); "
        f"B=1: {qps1:.1f} q/s ({tot1 / max(args.steps, 10) * 1e3:.2f} ms/query)")
    # the same answers replayed from captured hipGraphs (a warmup call captures)
    assert lib.ipir_dbg_set_graph(srv.h, 1) == 0
    gtot32, glat32 = timed(B, args.steps, max(args.warmup, 1))
    gtot1, glat1 = timed(1, max(args.steps, 10), max(args.warmup, 1))
    step(B)
    lib.ipir_dev_sync()
    g_equal = bool(np.array_equal(dout.download(np.zeros((B, resp_u64s), np.uint64)), resp))
    assert lib.ipir_dbg_set_graph(srv.h, 0) == 0
    graph = {f"batch{B}_qps": round(B * args.steps / gtot32, 2),
             "batch1_qps": round(max(args.steps, 10) / gtot1, 2),
             f"batch{B}_p50_ms": round(float(np.percentile(glat32, 50)) * 1e3, 3),
             "batch1_p50_ms": round(float(np.percentile(glat1, 50)) * 1e3, 3),
             "responses_equal_launch_by_launch": g_equal}
    log(f"[bench] graph replay: B={B}: {graph[f'batch{B}_qps']} q/s; B=1: {graph['batch1_qps']} q/s; "
        f"responses equal: {g_equal}")
    ok = ok and g_equal

    # per-stage device times (HIP events on the server stream)
    st1 = (C.c_double * 5)()
    lib.ipir_dbg_profile_batch(srv.h, C.c_void_p(dq.p), 1, C.c_void_p(dout.p), st1)
    st32 = (C.c_double * 5)()
    lib.ipir_dbg_profile_batch(srv.h, C.c_void_p(dq.p), B, C.c_void_p(dout.p), st32)
    names = ["unmarshal_keyntt", "db_pass", "collapse_pass", "pack_finalize", "horner"]
    stages1 = {k: round(v, 4) for k, v in zip(names, st1)}
    stages32 = {k: round(v, 4) for k, v in zip(names, st32)}
    log(f"[bench] stages B=1 (ms): {stages1}")
    log(f"[bench] stages B={B} (ms): {stages32}")
    # the dominant kernel of the measured workload: the database pass of a
    # batch of min(B, 32) queries, timed alone over 20 launches with HIP
    # events on the server stream (algorithmic bytes 2RC: the DB read once)
    def dbpass(nq):
        ms = C.c_double()
        assert lib.ipir_dbg_time_dbpass(srv.h, nq, 20, C.byref(ms)) == 0
        name = C.create_string_buffer(128)
        lib.ipir_dbg_dbpass_kernel(srv.h, nq, name, 128)
        return ms.value, name.value.decode()
    db_bytes = 2.0 * R * Cc
    dbpass_ms, kname = dbpass(1)
    achieved = db_bytes / (dbpass_ms * 1e-3) / 1e9
    dbB_ms, knameB = dbpass(min(B, 32))
    log(f"[bench] DB pass: B=1 {kname} {dbpass_ms:.3f} ms ({achieved:.0f} GB/s); "
        f"B={min(B, 32)} {knameB} {dbB_ms:.3f} ms")
    tb = C.create_string_buffer(1024)
    lib.ipir_dbg_tuning(srv.h, tb, 1024)
    tuning = tb.value.decode()
    log(f"[bench] tuning: {tuning}")
    achievedB = db_bytes / (dbB_ms * 1e-3) / 1e9

    # the same pass as the answers run it (streamed or chunked, beside the
    # collapse, pack and chains): HIP events around its launches on the
    # stream they run on, over 10 answers after the timed runs
    def in_answer(count, n=10):
        assert lib.ipir_dbg_pass_timing(srv.h, 1, None, None) == 0
        for _ in range(n):
            step(count)
        lib.ipir_dev_sync()
        tot, npass = C.c_double(), C.c_size_t()
        rc = lib.ipir_dbg_pass_timing(srv.h, 0, C.byref(tot), C.byref(npass))
        return tot.value / npass.value if rc == 0 and npass.value else None
    ia_ms1, ia_msB = in_answer(1), in_answer(B)
    log(f"[bench] DB pass inside the answers: B=1 {ia_ms1:.3f} ms, per {min(B, 32)}-query sub-batch "
        f"{ia_msB:.3f} ms")
    # HBM bytes per launch from the committed copy-calibrated counter runs
    # (profiles/pmc_traffic_*.md), for the same kernel on the same workload
    traffics = {}
    tfile = os.path.join(ROOT, "profiles", "dbpass_traffic.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile))
        if t.get("workload") == label:
            traffics = {k: v["traffic_bytes"] for k, v in t.get("kernels", {}).items()}
    def family(k):  # the kernel without its ring depth (which does not change the bytes it moves)
        m = re.match(r"(k_dbpass_mfma_(?:sa|bs))<(.*)>$", k)
        if not m:
            return k
        a = [x.strip() for x in m.group(2).split(",")]
        a[3 if m.group(1).endswith("sa") else 2] = "*"
        return f"{m.group(1)}<{', '.join(a)}>"

    def traffic_of(k):
        k = k.split(" slices=")[0]
        if k in traffics:
            return traffics[k], k
        same = sorted(x for x in traffics if family(x) == family(k))
        return (traffics[same[0]], same[0]) if same else (None, None)
    traffic, _ = traffic_of(kname)
    trafficB, traffic_kernel = traffic_of(knameB)
    # NTT throughput at N = 2048, both limbs per polynomial, 1024 polynomials per launch
    ntt = {}
    for inv in (0, 1):
        ms = C.c_double()
        assert lib.ipir_dbg_ntt_bench(1024, inv, 20, C.byref(ms)) == 0
        ntt["intt" if inv else "ntt"] = {"polys_per_s": round(1024 * 2 / (ms.value * 1e-3)),
                                         "us_per_launch_1024x2": round(ms.value * 1e3, 2)}
    log(f"[bench] NTT throughput (2048-point, per limb): {ntt}")

    # host-staged wire paths (PCIe-inclusive; the headline value keeps the
    # queries resident): flat u64 residues vs the packed wire form with the
    # codecs on the device
    def host_rate(fn, steps=5):
        fn()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        return B * steps / (time.perf_counter() - t)
    views = srv.views(qs)
    flat_out = np.zeros((B, resp_u64s), np.uint64)
    packed = np.stack([cl.pack_query(q) for q in qs])
    comp = np.zeros((B, T * 12288), np.uint8)
    qps_flat = host_rate(lambda: lib.ipir_answer_batch(srv.h, views, B, ia.ptr(flat_out)))
    qps_packed = host_rate(lambda: lib.ipir_answer_batch_packed(srv.h, ia.ptr(packed), B, ia.ptr(comp)))
    wire_ok = all(np.array_equal(comp[k], cl.compress(flat_out[k])) for k in range(B))
    wire = {"flat_qps": round(qps_flat, 1), "packed_qps": round(qps_packed, 1),
            "h2d_bytes_per_query": {"flat": int(qlen * 8), "packed": int(packed.shape[1])},
            "d2h_bytes_per_query": {"flat": int(resp_u64s * 8), "packed": int(T * 12288)},
            "packed_equals_host_codecs": bool(wire_ok)}
    log(f"[bench] host-staged B={B}: flat {qps_flat:.1f} q/s, packed wire {qps_packed:.1f} q/s "
        f"(codecs on device; bytes equal host codecs: {wire_ok})")

    serve = None
    if args.serving:
        # 0: partitions sized by the loop itself (32 slots each, as many as max_batch holds, up to 4)
        serve = serving(lib, srv, qs, qps32, resp_u64s, max_inflight=max(0, args.serving_inflight))

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle on the host worker threads, bounded sample)")
        cpu = cpu_baseline(geo)

    out = {
        "metric": ("queries/s at 16 GB DB, batch 32 vs batch 1 (bit-exact); p50/p99 latency" if args.config == "c3"
                   else f"queries/s at {label}, batch {B} vs batch 1 (bit-exact); p50/p99 latency"),
        "value": round(qps32, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tot32 / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (uniform u16 DB from seed 0, indices seed 1, client secrets seed 2)",
        "config": {"workload": label, "n_entries": n, "entry_bytes": w, "db_rows": R, "db_cols": Cc,
                   "n_packed": G, "num_cts": T, "batch": B},
        "batch1_qps": round(qps1, 2),
        "batch_speedup": round(qps32 / qps1, 3),
        "latency_ms": {
            f"batch{B}_p50": round(float(np.percentile(lat32, 50)) * 1e3, 3),
            f"batch{B}_p99": round(float(np.percentile(lat32, 99)) * 1e3, 3),
            "batch1_p50": round(float(np.percentile(lat1, 50)) * 1e3, 3),
            "batch1_p99": round(float(np.percentile(lat1, 99)) * 1e3, 3),
        },
        "stages_ms": {"batch1": stages1, f"batch{B}": stages32},
        "preprocess_s": round(t_pre, 2),
        "bit_exact_check": bool(ok),
        "verified_queries": verified,
        "roofline": {"bound": "hbm", "achieved": round(achievedB, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achievedB / HBM_PEAK_GBS, 4), "traffic": trafficB,
                     "in_answer": None if ia_msB is None else {
                         "ms_per_pass": round(ia_msB, 3),
                         "achieved": round(db_bytes / (ia_msB * 1e-3) / 1e9, 1),
                         "frac": round(db_bytes / (ia_msB * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "note": "the same pass inside the timed answers' shape (streamed or chunked launches "
                                 "beside the collapse, pack and chains), HIP events on its stream, 10 answers"},
                     "traffic_profiled_kernel": traffic_kernel,
                     "kernel": f"{knameB} (database pass, batch {min(B, 32)}; {dbB_ms:.3f} ms per launch)"},
        "batched_db_pass": {"kernel": knameB, "queries": min(B, 32), "ms": round(dbB_ms, 3),
                            "hbm_GBs": round(achievedB, 1),
                            "int8_mfma_TOPs": round(2 * int8_macs(min(B, 32), R, Cc, knameB) / (dbB_ms * 1e-3) / 1e12,
                                                    1),
                            "int8_mfma_frac_of_5000": round(2 * int8_macs(min(B, 32), R, Cc, knameB)
                                                            / (dbB_ms * 1e-3) / 5.0e15, 4)},
        "batch1_db_pass": {"kernel": kname, "ms": round(dbpass_ms, 3), "hbm_GBs": round(achieved, 1),
                           "in_answer_ms": None if ia_ms1 is None else round(ia_ms1, 3),
                           "frac_of_hbm_peak": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "tuning": tuning,
        "client": client,
        "max_batch_fitted": fitted,
        "graph_replay": graph,
        "horner_step_us": {"batch1": round(st1[4] * 1e3 / max(1, geo["Da"] - 1), 2),
                           f"batch{B}": round(st32[4] * 1e3 / max(1, geo["Da"] - 1), 2)},
        "ntt_throughput": ntt,
        "host_wire": wire,
        "cpu_baseline": cpu,
    }
    if serve is not None:
        k, per = serve[0]["partitions"], serve[0]["slots_per_partition"]
        out["serving"] = {"loop": f"ipir_loop (continuous batching, {k} partition{'s' if k > 1 else ''} of "
                                  f"{per} slots)", "rates": serve}
    for h in hs:
        cl.destroy(h)
    dq.free()
    dout.free()
    srv.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
