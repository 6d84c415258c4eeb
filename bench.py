"""Benchmark: FP64-equivalent TFLOP/s of the SIP block-contraction hot path.

Headline workload: the whole unrestricted (T) triples sum of the triples
configuration -- o = 30 (segments 20, 10), v = 240 (12 segments of 20), every
block tuple (A, B, C, I, J, K), 1.2093e15 flop of W contractions -- with W
kept on chip: one workgroup per 5^6 box of W runs the twelve contractions of
the box (six GEMMs over d and l) on FP64 MFMA, adds them in LDS, divides by
the orbital-energy denominators and reduces the energy; W is never written.
The box slots are cut into `--steps` equal launches, so the K timed steps are
exactly one sweep of the whole sum (no sampling); the operands are packed
once before the timed region (its time is reported beside).

Beside it, in the same JSON line: the split-precision path against the plain
path on the triples tuples (each shape class), the ring term with its fused
permute-in-load batch (the earlier headline), the block-sparse ladder, the
serving stream, permutes, block operations, graphs, the autotuner, the
single-contraction latency, and the reference's CPU path timed on bounded
samples of the triples and ring workloads.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from aces4_amd import sipgpu as sg  # noqa: E402
from aces4_amd import workloads as wl  # noqa: E402

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak
HBM_PEAK_GBS = 8000.0    # MI355X HBM3E peak
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "ring_gemm_traffic.txt")
PROFILE_TRIPLES_TRAFFIC = os.path.join(ROOT, "profiles", "triples_box_traffic.txt")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--ring-steps", type=int, default=20)
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="budget for the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--latency-samples", type=int, default=200)
    p.add_argument("--no-split", action="store_true")
    p.add_argument("--no-sparse", action="store_true")
    p.add_argument("--no-serve", action="store_true")
    p.add_argument("--no-triples", action="store_true")
    p.add_argument("--no-graphs", action="store_true")
    p.add_argument("--no-traffic", action="store_true",
                   help="take roofline.traffic from the committed profile instead of measuring it")
    p.add_argument("--traffic-probe", action="store_true",
                   help=argparse.SUPPRESS)  # the child run that the counters are read from
    return p.parse_args()


_T0 = time.time()


def progress(msg):
    """One line per finished leg on stderr (stdout carries only the JSON)."""
    print("[bench %6.1f s] %s" % (time.time() - _T0, msg), file=sys.stderr, flush=True)


def cpu_info():
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                 if l.startswith("model name")][0]
    except Exception:
        model = "unknown"
    return os.cpu_count(), model


def cpu_baseline(w, budget_s, gpu_result, steps_run):
    """The reference's own kernels (oracle/_ref) when present, else the C
    restatement: per block pair TMP = W*T (tensor_block_contract, 8 OpenMP
    threads as MAX_OMP_THREADS) then R += TMP (tensor_block_add), over a
    seeded random sample of the workload's block pairs until `budget_s`.
    Also checks one GPU result block (after `steps_run` accumulating steps)
    against the C restatement."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from checker import Oracle, RefLib
    from cases import rel_fro
    ref = RefLib() if RefLib.available() else None
    kind = "reference" if ref else "port"
    o = Oracle()
    D, L, R, ptrn = w["D"], w["L"], w["R"], w["ptrn"]
    rng = np.random.default_rng(4)
    order = rng.permutation(len(w["entries"]))
    mnk = wl.mnk(w)
    out = D.host.copy()
    flops = 0.0
    npairs = 0
    t0 = time.perf_counter()
    for idx in order:
        db, lb, rb = w["entries"][idx]
        if ref is not None:
            tmp, e = ref.contract(ptrn, L.block(lb), L.dims(lb), R.block(rb), R.dims(rb),
                                  D.dims(db), nthreads=8)
            blk = out[D.offset[db]:D.offset[db] + tmp.size]
            ref.add(D.dims(db), blk, tmp)
        else:
            tmp, e = o.contract(ptrn, L.block(lb), L.dims(lb), R.block(rb), R.dims(rb),
                                D.dims(db))
            out[D.offset[db]:D.offset[db] + tmp.size] += tmp
        m, n, k = mnk[idx]
        flops += 2.0 * m * n * k
        npairs += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    db0 = next(iter(D.block_ids()))
    want = np.zeros(int(np.prod(D.dims(db0))))
    for db, lb, rb in w["entries"]:
        if db == db0:
            tmp, _ = o.contract(ptrn, L.block(lb), L.dims(lb), R.block(rb), R.dims(rb),
                                D.dims(db))
            want += tmp
    check = rel_fro(D.block(db0, gpu_result), steps_run * want)
    return dict(value=flops / dt / 1e12, unit="TFLOP/s", cores=8 if ref else 1, kind=kind,
                check_rel_err=check,
                sample="%d of %d block pairs (seeded random), %.2e flop in %.2f s; "
                       "TMP = W*T then R += TMP per pair; %s" % (
                           npairs, len(w["entries"]), flops, dt,
                           "reference tensor_dil_omp.F90 built with -DNO_BLAS (its own GEMM), "
                           "nthreads=8" if ref else "C restatement, 1 thread"))


def traffic_from_profile(path=PROFILE_TRAFFIC, key="hbm_bytes_per_launch"):
    """HBM bytes per launch (or per `key` unit) measured by rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), if
    recorded."""
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def latency_sample(w, n):
    """Per-contraction latency (enqueue -> complete) of single unbatched
    contractions of the workload's block pairs, one at a time, after one
    untimed contraction of every block shape."""
    D, L, R, ptrn = w["D"], w["L"], w["R"], w["ptrn"]
    rng = np.random.default_rng(9)
    scratch = sg.DeviceArray(max(int(np.prod(D.dims(db))) for db, _, _ in w["entries"]))
    # steady state: every block shape is planned (offset tables uploaded,
    # split workspace sized) once before the timed samples
    seen = set()
    for db, lb, rb in w["entries"]:
        key = (tuple(L.dims(lb)), tuple(R.dims(rb)), tuple(D.dims(db)))
        if key not in seen:
            seen.add(key)
            sg.contract(ptrn, L.addr(lb), L.dims(lb), R.addr(rb), R.dims(rb), scratch.ptr,
                        D.dims(db))
    sg.synchronize()
    lat = []
    for idx in rng.integers(0, len(w["entries"]), size=n):
        db, lb, rb = w["entries"][idx]
        t0 = time.perf_counter()
        sg.contract(ptrn, L.addr(lb), L.dims(lb), R.addr(rb), R.dims(rb), scratch.ptr,
                    D.dims(db))
        sg.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
    scratch.free()
    lat = np.array(lat)
    return dict(p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                samples=int(n), mode="single contraction, unbatched, host enqueue to completion")


def permute_bench(reps=20):
    """Achieved HBM GB/s of the permute kernels (16 B moved per element):
    a 64^4 block (134 MB: input and output together fit the 256 MB
    last-level cache, so repeated runs read part of the input from it) and a
    120^4 block (1.66 GB, HBM only) with the full index reversal, the ring block shape
    30x15x30x15 with perm (3,4,1,2), and all 144 ring T blocks permuted one
    launch per block and as one batched launch, timed with HIP events."""
    out = {}
    for name, dims, perm in (("64^4 reversal", [64, 64, 64, 64], [4, 3, 2, 1]),
                             ("120^4 reversal", [120, 120, 120, 120], [4, 3, 2, 1]),
                             ("ring block (3,4,1,2)", [30, 15, 30, 15], [3, 4, 1, 2])):
        n = int(np.prod(dims))
        din = sg.DeviceArray.from_host(wl.uniform(11, n))
        dout = sg.DeviceArray(n)
        e0, e1 = sg.Event(), sg.Event()
        sg.permute(dims, [1] + perm, din.ptr, dout.ptr)
        sg.synchronize()
        e0.record()
        for _ in range(reps):
            sg.permute(dims, [1] + perm, din.ptr, dout.ptr)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        out[name] = {"dims": dims, "perm": perm, "ms": ms, "GBps": 16.0 * n / (ms * 1e-3) / 1e9}
        din.free()
        dout.free()
    # every T(a,i,c,k) block of the ring workload (v segments 30 x 4, 15 x 2;
    # o segments 15, 15: 144 blocks, 20.25 M elements) permuted (3,4,1,2):
    # block by block, and as one batched launch
    vs, os_ = [30, 30, 30, 30, 15, 15], [15, 15]
    shapes = [[a, i, c, k] for a in vs for i in os_ for c in vs for k in os_]
    sizes = [int(np.prod(d)) for d in shapes]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(offs[-1])
    din = sg.DeviceArray.from_host(wl.uniform(12, n))
    dout = sg.DeviceArray(n)
    perm = [1, 3, 4, 1, 2]
    ents = [sg.permutation(d, perm, din.ptr + 8 * int(offs[q]), dout.ptr + 8 * int(offs[q]))
            for q, d in enumerate(shapes)]
    b = sg.PermuteBatch(ents)
    e0, e1 = sg.Event(), sg.Event()

    def timed(fn):
        fn()
        sg.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_ms(e1) / reps

    def one_by_one():
        for q, d in enumerate(shapes):
            sg.permute(d, perm, din.ptr + 8 * int(offs[q]), dout.ptr + 8 * int(offs[q]))
    ms_each = timed(one_by_one)
    ms_batch = timed(b.execute)
    out["ring T blocks (3,4,1,2), 144 blocks"] = {
        "blocks": len(shapes), "elements": n,
        "block_by_block": {"ms": ms_each, "GBps": 16.0 * n / (ms_each * 1e-3) / 1e9},
        "batched": {"ms": ms_batch, "GBps": 16.0 * n / (ms_batch * 1e-3) / 1e9,
                    "workgroups": b.workgroups}}
    b.destroy()
    din.free()
    dout.free()
    return out


def rocblas_dgemm_ceiling(n=8192, reps=10):
    """FP64 TFLOP/s of rocBLAS's dense DGEMM (n x n x n, column-major, no
    transposes) on the null stream, the library's own ceiling for the FP64
    matrix cores on this GPU: wall time of `reps` calls between device
    synchronisations after two warm-up calls.  None when rocBLAS is not
    loadable or fails."""
    import ctypes
    try:
        rb = ctypes.CDLL("/opt/rocm/lib/librocblas.so")
    except OSError:
        return None
    h = ctypes.c_void_p()
    if rb.rocblas_create_handle(ctypes.byref(h)) != 0:
        return None
    bufs = [sg.DeviceArray(n * n) for _ in range(3)]
    try:
        for b in bufs[:2]:
            sg.fill_random(b.ptr, n * n, 21, 1, "uniform", -1.0, 1.0)
        bufs[2].zero()
        alpha, beta = ctypes.c_double(1.0), ctypes.c_double(0.0)
        none = 111  # rocblas_operation_none
        rb.rocblas_dgemm.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                     ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.c_int]

        def call():
            return rb.rocblas_dgemm(h, none, none, n, n, n, ctypes.byref(alpha), bufs[0].ptr, n,
                                    bufs[1].ptr, n, ctypes.byref(beta), bufs[2].ptr, n)
        sg.synchronize()
        for _ in range(2):
            if call() != 0:
                return None
        sg.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        sg.synchronize()
        wall = time.perf_counter() - t0
        return 2.0 * n ** 3 * reps / wall / 1e12
    finally:
        for b in bufs:
            b.free()
        rb.rocblas_destroy_handle(h)


def blockops_bench(reps=10):
    """Achieved HBM GB/s of the block kernels (algorithmic bytes: slice and
    insert 16 B per slice element, add 24 B, scale 16 B, fill 8 B, norm2 8 B
    per element), timed with HIP events on the library stream.  The slices
    are cut from a 120^4 block (1.66 GB): a half-extent slice (inner run of
    60 elements) and a 15^4 one (inner run of 15)."""
    out = {}
    te = [120, 120, 120, 120]
    nt = int(np.prod(te))
    tens = sg.DeviceArray(nt)
    sg.fill_random(tens.ptr, nt, 21, 0)
    big = sg.DeviceArray(nt)
    sg.fill_random(big.ptr, nt, 22, 0)
    res = sg.DeviceArray(1)

    def timed(name, fn, nbytes, **info):
        fn()
        sg.synchronize()
        e0, e1 = sg.Event(), sg.Event()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = dict(info, ms=ms, GBps=gbs, frac=gbs / HBM_PEAK_GBS)

    for se, eb in (([60, 120, 60, 120], [30, 0, 60, 0]), ([15, 15, 15, 15], [100, 3, 50, 7])):
        ns = int(np.prod(se))
        slc = sg.DeviceArray(ns)
        timed("slice %s of 120^4" % "x".join(map(str, se)),
              lambda: sg.block_slice(tens.ptr, te, slc.ptr, se, eb), 16.0 * ns,
              slice_ext=se, ext_beg=eb)
        timed("insert %s into 120^4" % "x".join(map(str, se)),
              lambda: sg.block_insert(tens.ptr, te, slc.ptr, se, eb), 16.0 * ns,
              slice_ext=se, ext_beg=eb)
        slc.free()
    n = nt
    timed("add (scaled) 120^4", lambda: sg.block_add(big.ptr, tens.ptr, [n], 0.5), 24.0 * n)
    timed("scale 120^4", lambda: sg.block_scale(big.ptr, [n], 1.0000001), 16.0 * n)
    timed("init 120^4", lambda: sg.block_init(big.ptr, [n], 0.25), 8.0 * n)
    timed("norm2 120^4", lambda: sg.block_norm2(tens.ptr, [n], res.ptr), 8.0 * n)
    for d in (tens, big, res):
        d.free()
    return out


def split_bench(w, fp64_batch, reps=20, moduli=(14, 13, 12)):
    """The split-precision path (INT8 MFMA over pairwise-coprime moduli + CRT)
    on the same batch: FP64-equivalent TFLOP/s (the batch's 2MNK flop over the
    whole execution: exponents, residue conversion, INT8 GEMMs,
    reconstruction), and the worst per-block relative Frobenius error against
    the plain FP64 path."""
    from cases import rel_fro
    D = w["D"]
    zeros = np.zeros_like(D.host)
    D.dev.upload(zeros)
    fp64_batch.execute()
    sg.synchronize()
    ref = D.dev.download()
    out = {}
    for nm in moduli:
        b, tres = sg.tune(wl.contractions_for(w), moduli=nm)  # planned by the autotuner
        D.dev.upload(zeros)
        b.execute()
        sg.synchronize()
        got = D.dev.download()
        err = max(rel_fro(D.block(db, got), D.block(db, ref)) for db in D.block_ids())
        b.execute()
        sg.synchronize()
        e0, e1 = sg.Event(), sg.Event()
        e0.record()
        for _ in range(reps):
            b.execute()
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        out["moduli_%d" % nm] = {"moduli": nm, "bits": b.bits, "ms": ms,
                                 "tflops_fp64_equiv": b.flops / (ms * 1e-3) / 1e12,
                                 "max_block_rel_err_vs_fp64": err,
                                 "tuned": {"tile": tres["tile"], "merge": tres["merge"],
                                           "untuned_ms": tres["default_ms"]}}
        b.destroy()
    return out


def sparse_bench(reps=5, moduli=14):
    """Block-sparse ladder (o=40, v=200, uneven segments 12..40, 4 irreps):
    one grouped FP64 launch and the split path over all surviving block
    pairs, data generated in HBM.  Split checked against FP64 on every
    result block."""
    from cases import rel_fro
    w = wl.sparse_ladder()
    for k in ("D", "L", "R"):
        w[k].to_device()
    D = w["D"]
    out = {"pairs": len(w["entries"]), "result_blocks": len(D.offset),
           "flop": wl.flops(w), "dense_flop": w["dense_flops"],
           "surviving_fraction": wl.flops(w) / w["dense_flops"],
           "v_segments": w["v_segs"], "o_segments": w["o_segs"]}
    ref = None
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from checker import Oracle
    o = Oracle()
    # one result block (seeded choice) against the oracle's sum over its pairs
    ids = list(D.block_ids())
    rb = ids[int(np.random.default_rng(11).integers(len(ids)))]
    want = np.zeros(int(np.prod(D.dims(rb))))
    for db, vb, tb in w["entries"]:
        if db == rb:
            tmp, _ = o.contract(w["ptrn"], w["L"].download_block(vb), w["L"].dims(vb),
                                w["R"].download_block(tb), w["R"].dims(tb), D.dims(rb))
            want += tmp
    out["oracle_block"] = list(rb)
    for name, nm in (("fp64", None), ("split", moduli)):
        b, tres = sg.tune(wl.contractions_for(w), moduli=nm)  # planned by the autotuner
        D.dev.zero()
        b.execute()
        sg.synchronize()
        got = D.dev.download()
        oracle_err = rel_fro(D.download_block(rb, got), want)
        if ref is None:
            ref = got
            err = None
        else:
            err = max(rel_fro(D.download_block(db, got), D.download_block(db, ref))
                      for db in D.block_ids())
        e0, e1 = sg.Event(), sg.Event()
        e0.record()
        for _ in range(reps):
            b.execute()
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        out[name] = {"ms": ms, "tflops_fp64_equiv": b.flops / (ms * 1e-3) / 1e12,
                     "moduli": b.moduli, "bits": b.bits, "tuned": tune_summary(tres),
                     "rel_err_vs_oracle_on_oracle_block": oracle_err}
        if err is not None:
            out[name]["max_block_rel_err_vs_fp64"] = err
        b.destroy()
    for k in ("D", "L", "R"):
        w[k].dev.free()
    return out


def serving_bench(label, budget_gb, n, rate, seg_lo, seg_hi, max_batch=256, moduli=0):
    """Continuous batching of a Poisson stream of single block contractions
    (ladder / ring / hole-hole mix) through the device block cache: the
    stream is replayed once to warm the cache, then timed against its
    arrival schedule.  Latency = completion - scheduled arrival.  `moduli`
    > 0: every request on the split-precision path."""
    mix = wl.serving_mix(n_requests=n, rate=rate, seg_lo=seg_lo, seg_hi=seg_hi, moduli=moduli)
    cache = sg.Cache(int(budget_gb * (1 << 30)))
    ws = wl.populate_cache(cache, mix)
    srv = sg.Server(cache, max_batch=max_batch)
    srv.replay(mix["requests"], np.zeros(n))
    c0, s0 = cache.stats(), srv.stats()
    lat, wall = srv.replay(mix["requests"], mix["arrivals"])
    c1, s1 = cache.stats(), srv.stats()
    srv.destroy()
    cache.destroy()
    batches = s1["batches"] - s0["batches"]
    return {
        "scenario": label, "requests": n, "offered_rate_per_s": rate,
        "precision": "split, %d moduli" % moduli if moduli else "fp64",
        "segments": mix["segs"], "terms": mix["terms"],
        "working_set_gb": ws / (1 << 30), "budget_gb": budget_gb,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
        "wall_s": wall, "tflops": float(mix["flops"].sum()) / wall / 1e12,
        "batches": batches, "mean_batch": n / max(batches, 1),
        "cache": {k: c1[k] - c0[k] for k in ("hits", "misses", "evictions", "spills",
                                             "bytes_h2d", "bytes_d2h")},
    }


def graph_bench(reps=20):
    """CC iterations replayed as graphs against the same work enqueued
    eagerly: the ladder of the CPU-runnable configuration (o = 10, v = 40,
    256 block pairs of 100 x 100 x 100, one grouped launch, plus the energy
    sum over its result), and one triples tuple (twelve batches and the
    fused energy).  Per-iteration time between events on the launch stream,
    averaged over `reps` back-to-back iterations."""
    from aces4_amd import triples as tp
    s = sg.Stream()
    res = {}

    def timed(fn):
        fn(s)
        s.synchronize()
        e0, e1 = sg.Event(), sg.Event()
        e0.record(s)
        for _ in range(reps):
            fn(s)
        e1.record(s)
        e1.synchronize()
        return e0.elapsed_ms(e1) / reps

    def compare(name, it, flops):
        eager = timed(it)
        g = sg.Graph.capture(it, s)
        replay = timed(lambda st: g.launch(st))
        res[name] = {"eager_ms": eager, "graph_ms": replay, "graph_nodes": g.nodes,
                     "tflops_graph": flops / (replay * 1e-3) / 1e12, "speedup": eager / replay}
        g.destroy()

    w = wl.ladder()
    b = wl.batch_for(w)
    D = w["D"]
    out = sg.DeviceArray(1)
    nd = len(D.host)
    eps = sg.DeviceArray.from_host(np.linspace(0.5, 1.5, nd))  # one weight per element

    def ladder_it(st):
        b.execute(st)
        sg.energy_sum([nd], D.dev.ptr, D.dev.ptr, [eps.ptr], [1], 1.0, out.ptr, s=st)

    compare("ladder_iteration", ladder_it, b.flops)
    b.destroy()
    w3 = tp.triples()
    tp.to_device(w3)
    t = (0, 1, 2, 0, 0, 0)
    nw = int(np.prod(tp.tuple_dims(w3, t)))
    wbuf = sg.DeviceArray(nw)
    parts = [sg.DeviceArray(nw) for _ in range(5)]
    te = tp.TupleEnergy(w3, t, wbuf.ptr, out.ptr, parts=[p.ptr for p in parts], tune=True,
                        two_streams=True)
    compare("triples_tuple", lambda st: te.execute(st), te.flops)
    te.destroy()
    wbuf.free()
    for p in parts:
        p.free()
    for k in ("V1", "V2", "T"):
        w3[k].dev.free()
    s.destroy()
    return res


# chunks the traffic probe runs (of the bench's chunk size, spread over the sweep;
# indices past the last chunk wrap round)
PROBE_LAUNCHES = (0, 5, 10, 15)


def traffic_probe(steps):
    """Child process under rocprofv3 --pmc: the calibration copy (1 GiB read,
    1 GiB written by copy_kernel, via sipgpu_measure_peaks), then launches of
    the headline kernel over the same slot chunks as the bench."""
    from aces4_amd import triples as tp
    sg.init(0)
    sg.measure_peaks(1 << 30)
    w = tp.triples()
    tp.to_device(w)
    ft = tp.FusedTriples(w)
    ft.pack()
    chunk = -(-ft.slots // steps)
    for i in PROBE_LAUNCHES:
        i %= steps
        ft.run(i * chunk, max(0, min(chunk, ft.slots - i * chunk)))
    sg.synchronize()
    ft.destroy()
    print("traffic probe done", flush=True)


def measure_traffic(steps, timeout_s=240):
    """HBM bytes per launch of triples_box_kernel, measured now: two child
    runs of this script under rocprofv3 --pmc (FETCH_SIZE, then WRITE_SIZE),
    each counter scaled by what it reads for the calibration copy of known
    size in the same run.  Returns (bytes per launch, details) or (None,
    reason) when the profiler is unavailable or fails."""
    import csv
    import shutil
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None, "rocprofv3 not found"
    known = float(1 << 30)  # bytes the calibration copy reads, and writes
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="traffic_", dir="/tmp")
        cmd = [prof, "--pmc", counter, "--kernel-include-regex", "triples_box_kernel|copy_kernel",
               "--output-format", "csv", "-d", d, "-o", "t", "--",
               sys.executable, os.path.abspath(__file__), "--traffic-probe", "--steps", str(steps)]
        env = dict(os.environ, TMPDIR="/tmp")
        try:
            r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True,
                               timeout=timeout_s)
        except subprocess.TimeoutExpired:
            shutil.rmtree(d, ignore_errors=True)
            return None, "rocprofv3 %s run timed out" % counter
        rows = []
        for root, _, files in os.walk(d):
            for f in files:
                if f.endswith("counter_collection.csv"):
                    rows += list(csv.DictReader(open(os.path.join(root, f))))
        shutil.rmtree(d, ignore_errors=True)
        copy = [float(x["Counter_Value"]) for x in rows if "copy_kernel" in x["Kernel_Name"]]
        box = [float(x["Counter_Value"]) for x in rows if "triples_box_kernel" in x["Kernel_Name"]]
        if r.returncode != 0 and not (copy and box):
            return None, "rocprofv3 %s run failed (%d)" % (counter, r.returncode)
        if not copy or len(box) < len(PROBE_LAUNCHES):
            return None, "rocprofv3 %s: calibration or kernel dispatches missing" % counter
        scale = known / (1024.0 * float(np.median(copy)))  # counter reads KiB
        # bytes per step: a step's boxes may take several kernel launches
        per_step = float(np.sum(box)) / len(PROBE_LAUNCHES)
        out[counter] = {"kib_per_dispatch": [float(b) for b in box], "calibration_kib": copy[0],
                        "scale": scale, "bytes_per_step": per_step * 1024.0 * scale}
    total = out["FETCH_SIZE"]["bytes_per_step"] + out["WRITE_SIZE"]["bytes_per_step"]
    return total, out


def triples_fused_bench(steps, warmup):
    """The headline: the whole unrestricted (T) sum with W on chip.  The box
    slots are cut into `steps` equal launches, timed between device
    synchronisations (value) and each between HIP events on its stream
    (roofline); `warmup` launches of the same chunks run first.  The
    operand packing runs once before the timed region and is timed apart."""
    from aces4_amd import triples as tp
    w = tp.triples()
    tp.to_device(w)
    ft = tp.FusedTriples(w)
    out = sg.DeviceArray(1)
    e0, e1 = sg.Event(), sg.Event()
    e0.record()
    ft.pack()
    e1.record()
    e1.synchronize()
    pack_ms = e0.elapsed_ms(e1)
    chunk = -(-ft.slots // steps)
    ranges = [(i * chunk, max(0, min(chunk, ft.slots - i * chunk))) for i in range(steps)]
    for i in range(warmup):
        ft.run(*ranges[i % steps])
    sg.synchronize()
    evs = [sg.Event() for _ in range(steps + 1)]
    t0 = time.perf_counter()
    evs[0].record()
    for i in range(steps):
        ft.run(*ranges[i])
        evs[i + 1].record()
    sg.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = np.array([evs[i].elapsed_ms(evs[i + 1]) for i in range(steps)])
    out.zero()
    ft.energy(out.ptr)
    sg.synchronize()
    energy = float(out.download()[0])
    v, o = sum(w["v_segs"]), sum(w["o_segs"])
    flops = ft.flops  # 12 terms x 2 v^3 o^3 (v or o): every box is full here
    # sipgpu_triples_run makes one kernel launch per tp.LAUNCH_BOXES slots:
    # the roofline divides a step by its launches (one at 12 steps or more)
    kernels = -(-chunk // tp.LAUNCH_BOXES)
    per_launch = flops * chunk / ft.slots / kernels
    res = {"wall_s": wall, "flops": flops, "energy": energy, "pack_ms": pack_ms,
           "slots": ft.slots, "boxes_per_launch": chunk / kernels, "boxes_per_step": chunk,
           "kernels_per_step": kernels,
           "launch_ms": launch_ms / kernels, "flop_per_launch": per_launch,
           "panel_bytes": ft.panel_bytes,
           "mfma_flop_issued_per_box": ft.mfma_flop_per_box,
           "algorithmic_flop_per_box": flops / ft.slots,
           "v_segments": w["v_segs"], "o_segments": w["o_segs"], "o": o, "v": v}
    return w, ft, res


def triples_tuple_bench(w, ft, tuples, moduli=(14,), reps=2):
    """Per tuple, for one tuple of every occupied-segment shape class: the
    fused kernel's energy against the twelve-batch FP64 path (each batch one
    GEMM task; oracle-pinned at small sizes), and the twelve-batch path's
    split-precision form (INT8 MFMA over `moduli` moduli) timed against the
    fused FP64 tuple, with its energy error."""
    from aces4_amd import triples as tp
    nmax = max(int(np.prod(tp.tuple_dims(w, t))) for t in tuples)
    wbuf = sg.DeviceArray(nmax)
    parts = [sg.DeviceArray(nmax) for _ in range(5)]
    out = sg.DeviceArray(1)
    flops = 0.0
    fused_e, fused_ms = [], 0.0
    for t in tuples:
        # six GEMMs of 2 |W| (v + o) flop each
        flops += 12.0 * np.prod(tp.tuple_dims(w, t)) * (sum(w["v_segs"]) + sum(w["o_segs"]))
    # fused tuples (tuple mode)
    for t in tuples:
        out.zero()
        ft.tuple(t, out.ptr)
        sg.synchronize()
        fused_e.append(float(out.download()[0]))
    e0, e1 = sg.Event(), sg.Event()
    e0.record()
    for _ in range(reps):
        for t in tuples:
            ft.tuple(t, out.ptr)
    e1.record()
    e1.synchronize()
    fused_ms = e0.elapsed_ms(e1) / reps
    res = {"tuples": [list(t) for t in tuples], "flop": flops,
           "fused_fp64": {"ms": fused_ms, "tflops": flops / (fused_ms * 1e-3) / 1e12}}
    for nm in (None,) + tuple(moduli):
        tes = [tp.TupleEnergy(w, t, wbuf.ptr, out.ptr, moduli=nm, parts=[p.ptr for p in parts],
                              tune=True, two_streams=True) for t in tuples]
        errs = []
        for te, ef in zip(tes, fused_e):
            out.zero()
            te.execute()
            sg.synchronize()
            e = float(out.download()[0])
            errs.append(abs(e - ef) / abs(ef))
        e0.record()
        for _ in range(reps):
            for te in tes:
                te.execute()
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        key = "batched_fp64" if nm is None else "batched_split_%d" % nm
        res[key] = {"ms": ms, "tflops": flops / (ms * 1e-3) / 1e12,
                    "max_tuple_energy_rel_err_vs_fused_fp64": max(errs),
                    "vs_fused_fp64": fused_ms / ms}
        for te in tes:
            te.destroy()
    wbuf.free()
    out.free()
    for p in parts:
        p.free()
    return res


def triples_cpu_baseline(w, budget_s):
    """The reference's CPU path on a bounded sample of the triples workload:
    the fourteen contractions of one 20^6 tuple's first permutation
    (term 1 over the 12 d segments, term 2 over the 2 l segments), each
    TMP = A*B then W += TMP, with the reference's own kernels (oracle/_ref,
    -DNO_BLAS, 8 threads) and with the reference's BLAS build restated on
    the host (oracle/blas_path.py: its permutes, numpy's OpenBLAS dgemm) at 1
    and at all host threads; each leg stops after `budget_s` seconds."""
    from aces4_amd import triples as tp
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from checker import Oracle, RefLib
    import blas_path
    o = Oracle()
    tup = (0, 1, 2, 0, 0, 0)
    dims = tp.tuple_dims(w, tup)
    specs = tp.term_labels()[:2]
    pairs = []
    for ptrn, prs in tp.tuple_entries(w, tup, specs):
        for ln, lb, rn, rb in prs:
            L, R = w[ln], w[rn]
            pairs.append((ptrn, L.download_block(lb), L.dims(lb), R.download_block(rb), R.dims(rb)))
    ncpu = os.cpu_count() or 1
    mn = float(np.prod(dims))

    def flop(ad, bd):
        # |A| |B| = M K N K and M N = |W|
        return 2.0 * mn * np.sqrt(float(np.prod(ad)) * float(np.prod(bd)) / mn)

    total = sum(flop(ad, bd) for _, _, ad, _, bd in pairs)
    out = {"sample": "the %d contractions of tuple %s's first permutation (%.3g flop), "
                     "TMP = A*B then W += TMP each" % (len(pairs), list(tup), total)}

    def leg(name, contract, add, cores):
        W = np.zeros(int(mn))
        t0 = time.perf_counter()
        fl, n = 0.0, 0
        for ptrn, A, ad, B, bd in pairs:
            tmp = contract(ptrn, A, ad, B, bd)
            add(W, tmp)
            fl += flop(ad, bd)
            n += 1
            if time.perf_counter() - t0 > budget_s:
                break
        dt = time.perf_counter() - t0
        out[name] = {"value": fl / dt / 1e12, "unit": "TFLOP/s", "cores": cores,
                     "contractions": n, "seconds": dt}

    if RefLib.available():
        ref = RefLib()
        leg("reference_no_blas",
            lambda p, A, ad, B, bd: ref.contract(p, A, ad, B, bd, dims, nthreads=8)[0],
            lambda W, t: ref.add(dims, W, t), 8)
    for th in (1, ncpu):
        with blas_path.blas_threads(th):
            leg("blas_path_%d_threads" % th,
                lambda p, A, ad, B, bd: blas_path.contract(o, p, A, ad, B, bd, dims),
                lambda W, t: np.add(W, t, out=W), th)
    return out


def tune_summary(res):
    out = dict(res)
    out["path"] = sg.PATHS[res["path"]]
    out["speedup_vs_untuned"] = res["default_ms"] / res["best_ms"] if res["best_ms"] else None
    return out


def tune_report(w):
    """Tunes workload w's batch and reports the choice and both timings."""
    b, res = sg.tune(wl.contractions_for(w))
    out = tune_summary(res)
    out["flop"] = b.flops
    out["tflops_tuned"] = b.flops / (res["best_ms"] * 1e-3) / 1e12
    b.destroy()
    for k in ("D", "L", "R"):
        w[k].dev.free()
    return out


GEMM_KERNEL = {"tables": "gemm_f64_kernel", "packed": "gemm_f64_packed_kernel"}
PATH_NAME = {"tables": "plain FP64 MFMA, grouped batch, permutations fused into the GEMM loads",
             "packed": "plain FP64 MFMA, operands packed into K-contiguous panels once per "
                       "step (pack_kernel), then one grouped GEMM over the panels"}


def ring_bench(steps, warmup, cpu_seconds, no_cpu, no_split, latency_samples):
    """The ring term (o = 30 in 15, 15; v = 150 in 30, 30, 30, 30, 15, 15):
    the tuned grouped batch timed per step, its GEMM alone per launch, the
    split path on the same batch, single-contraction latency, and the
    reference's CPU path on the same block pairs."""
    w = wl.ring()
    # the batch is planned by the autotuner (a planning step, not timed)
    batch, tres = sg.tune(wl.contractions_for(w))
    flops = batch.flops
    for _ in range(warmup):
        batch.execute()
    sg.synchronize()
    evs = [sg.Event() for _ in range(steps + 1)]
    t0 = time.perf_counter()
    evs[0].record()
    for i in range(steps):
        batch.execute()
        evs[i + 1].record()
    sg.synchronize()
    wall = time.perf_counter() - t0
    step_ms = np.array([evs[i].elapsed_ms(evs[i + 1]) for i in range(steps)])
    got = w["D"].dev.download()
    # the GEMM kernel by itself: operands prepared once, then `steps`
    # launches of the product stage alone, each between events on its stream
    batch.execute_stages(sg.STAGE_PREPARE)
    batch.execute_stages(sg.STAGE_PRODUCT)
    sg.synchronize()
    pevs = [sg.Event() for _ in range(steps + 1)]
    pevs[0].record()
    for i in range(steps):
        batch.execute_stages(sg.STAGE_PRODUCT)
        pevs[i + 1].record()
    sg.synchronize()
    launch_ms = np.array([pevs[i].elapsed_ms(pevs[i + 1]) for i in range(steps)])
    achieved = flops / (float(launch_ms.mean()) * 1e-3) / 1e12
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    res = {"workload": "ring R(a,i,b,j) += W(k,b,c,j) T(a,i,c,k)", "o_segments": [15, 15],
           "v_segments": [30, 30, 30, 30, 15, 15], "block_pairs": batch.npairs,
           "result_blocks": batch.ntasks, "flop_per_step": flops, "path": PATH_NAME[batch.kind],
           "steps": steps, "tflops": flops * steps / wall / 1e12,
           "ms_per_step": wall / steps * 1e3,
           "roofline": {"bound": "mfma", "kernel": GEMM_KERNEL[batch.kind],
                        "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                        "frac": achieved / FP64_PEAK_TFLOPS,
                        "avg_launch_ms": float(launch_ms.mean()),
                        "avg_step_ms": float(step_ms.mean()),
                        "traffic": traffic_from_profile()},
           "tuning": tune_summary(tres)}
    if not no_split:
        split = split_bench(w, batch)
        ok = [v for v in split.values() if v["max_block_rel_err_vs_fp64"] <= 1e-12]
        best = max(ok, key=lambda v: v["tflops_fp64_equiv"]) if ok else None
        step_rate = flops / (float(step_ms.mean()) * 1e-3) / 1e12
        split["best_within_1e-12"] = best["moduli"] if best else None
        bt = sg.Batch(wl.contractions_for(w), tol=1e-12)  # moduli from the error target
        split["moduli_for_target_1e-12"] = {"moduli": bt.moduli, "bits": bt.bits,
                                            "rule": "fewest moduli with 16 * 2^-P <= 1e-12"}
        bt.destroy()
        split["speedup_vs_plain"] = best["tflops_fp64_equiv"] / step_rate if best else None
        split["speedup_vs_plain_gemm_only"] = best["tflops_fp64_equiv"] / achieved if best else None
        res["split"] = split
    if latency_samples > 0:
        res["latency"] = latency_sample(w, latency_samples)
    if not no_cpu:
        res["cpu_baseline"] = cpu_baseline(w, cpu_seconds, got, warmup + steps)
    batch.destroy()
    for k in ("D", "L", "R"):
        w[k].dev.free()
    return res


def main():
    a = parse()
    if a.traffic_probe:
        traffic_probe(a.steps)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: this build runs the contraction path on a single "
              "MI355X (multi-GPU sharding is a design note only).", file=sys.stderr)
        sys.exit(2)
    sg.init(0)
    # the ladder of the CPU-runnable configuration, tuned first: its choice
    # is remembered for the graph leg
    tuning = {"ladder_cfg1": tune_report(wl.ladder())}
    progress("ladder tuned")

    # headline: the whole (T) sum with W on chip
    w5, ft, tri = triples_fused_bench(a.steps, a.warmup)
    progress("whole (T) sum: %.1f TFLOP/s" % (tri["flops"] / tri["wall_s"] / 1e12))
    launch_ms = tri.pop("launch_ms")
    avg_launch_s = float(launch_ms.mean()) / 1e3
    achieved = tri["flop_per_launch"] / avg_launch_s / 1e12
    per_box = traffic_from_profile(PROFILE_TRIPLES_TRAFFIC, "hbm_bytes_per_box")
    traffic = per_box * tri["boxes_per_launch"] if per_box else None
    kernels = tri["kernels_per_step"]
    traffic_src = {"source": "profiles/triples_box_traffic.txt (committed rocprofv3 --pmc runs)"}
    shape_classes = [(0, 1, 2, i, j, k) for i in (0, 1) for j in (0, 1) for k in (0, 1)]
    tri_tuples = None if a.no_triples else triples_tuple_bench(w5, ft, shape_classes)
    progress("triples tuples")
    tri_cpu = None if a.no_cpu else triples_cpu_baseline(w5, a.cpu_seconds / 2)
    progress("triples CPU baseline")
    ft.destroy()
    for k in ("V1", "V2", "T"):
        w5[k].dev.free()

    ring = ring_bench(a.ring_steps, 5, a.cpu_seconds, a.no_cpu, a.no_split, a.latency_samples)
    tuning["ring"] = ring.pop("tuning")
    progress("ring")
    sparse = None if a.no_sparse else sparse_bench()
    progress("sparse ladder")
    serving = None
    if not a.no_serve:
        serving = [serving_bench("resident (HBM budget 200 GB)", 200.0, 3000, 2000.0, 8, 64),
                   serving_bench("resident (HBM budget 200 GB), split precision", 200.0, 3000,
                                 2000.0, 8, 64, moduli=14),
                   serving_bench("LRU spill (budget 60% of working set)", 8.5, 1500, 4000.0,
                                 8, 40)]
        # offered rate swept to saturation: throughput against p50 / p99
        sweep = []
        for rate in (4000.0, 8000.0, 16000.0, 64000.0):
            r = serving_bench("resident, rate sweep", 200.0, 3000, rate, 8, 64)
            sweep.append({k: r[k] for k in ("offered_rate_per_s", "p50_ms", "p99_ms", "tflops",
                                            "wall_s", "mean_batch")})
        serving.append({"scenario": "resident (HBM budget 200 GB), offered rate swept to "
                                    "saturation, plain FP64", "points": sweep})
        progress("serving")
    graphs = None if a.no_graphs else graph_bench()
    perm = permute_bench()
    blockops = blockops_bench()
    progress("graphs, permutes, block operations")
    cores, model = cpu_info()
    mfma_meas, copy_meas = sg.measure_peaks(1 << 30)
    if not a.no_traffic:
        sg.synchronize()
        live, info = measure_traffic(a.steps)
        progress("traffic: %s" % ("%.3e bytes per launch" % live if live else info))
        if live is not None:
            traffic = live / kernels
            traffic_src = {"source": "measured in this run: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                                     "child runs of launches %s of the same chunk size, scaled by the "
                                     "calibration copy" % (list(PROBE_LAUNCHES),), "counters": info}
        else:
            traffic_src["live_measurement"] = info
    clock = sg.measured_clock_mhz()
    dgemm = rocblas_dgemm_ceiling()
    read_gbs = blockops.get("norm2 120^4", {}).get("GBps")
    best = max([x for x in (mfma_meas, dgemm) if x], default=None)
    measured = {"mfma_f64_tflops": mfma_meas, "clock_mhz": clock,
                "rocblas_dgemm_8192_tflops": dgemm,
                "fp64_matrix_best_tflops": best,
                "copy_gbs": copy_meas, "copy_bytes": 2 * (1 << 30),
                "read_stream_gbs": read_gbs,
                "lds_read_bytes_per_clock_per_cu": sg.measure_lds(),
                "note": "MFMA-only FP64 loop over all CUs; rocBLAS dense DGEMM 8192^3 on the same GPU; "
                        "fp64_matrix_best_tflops is the larger of the two (frac_of_measured divides by "
                        "it); streaming copy of 1 GiB (read+write); read stream = the norm2 kernel over "
                        "a 120^4 block; 16-byte LDS reads, two workgroups per CU"}
    cpu = None
    if tri_cpu:
        ref = tri_cpu.get("reference_no_blas")
        pick = ref or tri_cpu["blas_path_%d_threads" % (os.cpu_count() or 1)]
        cpu = {"value": pick["value"], "unit": "TFLOP/s", "cores": pick["cores"],
               "kind": "reference" if ref else "port", "sample": tri_cpu["sample"],
               "host": "%s, %d logical cores" % (model, cores), "legs": tri_cpu}
    res = {
        "metric": "FP64-equiv TFLOP/s, split vs plain FP64; p50/p99 per-contraction latency",
        "value": tri["flops"] / tri["wall_s"] / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": tri["wall_s"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (SplitMix64-seeded V ~ N(0,0.1^2), T ~ N(0,1e-2^2), "
                "e_occ ~ U(-2,-0.3), e_vir ~ U(0.1,3))",
        "config": {"workload": "(T) triples energy, whole unrestricted sum: W(abcijk) = sum over "
                               "6 index-pair permutations of V(bcdk)T(adij) - V(lcjk)T(abil), "
                               "E = sum W*W/D",
                   "o": tri["o"], "v": tri["v"], "o_segments": tri["o_segments"],
                   "v_segments": tri["v_segments"], "block_tuples": 12 ** 3 * 2 ** 3,
                   "flop": tri["flops"], "energy": tri["energy"],
                   "path": "plain FP64 MFMA, W kept on chip (5^6 boxes in LDS, six GEMMs per box, "
                           "energy reduced in the epilogue); operands packed once into MFMA-order "
                           "panels",
                   "pack_ms": tri["pack_ms"], "slots": tri["slots"],
                   "boxes_per_step": tri["boxes_per_step"],
                   "kernel_launches_per_step": kernels,
                   "tflops_including_pack": tri["flops"] / (tri["wall_s"] + tri["pack_ms"] * 1e-3)
                   / 1e12},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "triples_box_kernel", "avg_launch_ms": avg_launch_s * 1e3,
                     "flop_per_launch": tri["flop_per_launch"],
                     "mfma_flop_issued_per_box": tri["mfma_flop_issued_per_box"],
                     "algorithmic_flop_per_box": tri["algorithmic_flop_per_box"],
                     "measured_ceiling": measured,
                     "frac_of_measured": achieved / best if best else None},
        "triples_tuples": tri_tuples,
        "ring": ring,
        "permute": perm,
        "block_ops": blockops,
        "sparse_ladder": sparse,
        "serving": serving,
        "graphs": graphs,
        "tuning": tuning,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
