"""Benchmark of the k-mer counting hot path on one MI355X.

Workload (default): seeded 100 Mbp genome (GC 41%, 10% of the sequence as 1-5 kbp copies at 2%
divergence), 30x coverage of 150 bp reads (20M reads, 1% substitutions), k=31, minimizer 11,
512 partitions, full table (min-count 1). Reads are generated on the host, uploaded once, and
resident in HBM before the timed region. A step = one pass over the batch: 2-bit packing,
minimizer scan + superkmer partitioning (exact two-pass sizing), and the count stage that
produces every partition's (k-mer, count) table in HBM.

Prints one JSON line (rank 0). `value` = k-mers counted per second (N / step time), N = sum over
reads of (len - k + 1). `roofline` covers the count kernel: algorithmic bytes N*W + D_out*(W+4)
(read every key once, write the table once) over its HIP-event time. `cpu_baseline` times the
reference's own CPU path (cpu_funcs.cpp built from source into oracle/_ref) on a bounded sample of
the same workload on this host.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12  # bytes/s, MI355X HBM3E


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(wl, chars, offs, threads, sample_reads):
    """The reference CPU pipeline (GenSuperkmerCPU + KmerCountingCPU with its MSD radix sort),
    -cpuonly -im semantics, on the first `sample_reads` reads."""
    path = os.path.join(ROOT, "oracle", "_ref", "libref_cpu64.so" if wl.k <= 32 else "libref_cpu128.so")
    kind = "reference"
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    L.ref_cpu_pipeline.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    L.ref_cpu_pipeline.restype = ctypes.c_int64
    n = min(sample_reads, len(offs) - 1)
    so = np.ascontiguousarray(offs[: n + 1])
    t1 = np.zeros(1); t2 = np.zeros(1)
    t0 = time.perf_counter()
    distinct = L.ref_cpu_pipeline(chars.ctypes.data, so.ctypes.data, n, wl.k, wl.p, wl.parts, threads, 2000,
                                  t1.ctypes.data, t2.ctypes.data)
    wall = time.perf_counter() - t0
    lens = np.diff(so.astype(np.int64))
    nk = int(np.maximum(lens - wl.k + 1, 0).sum())
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": nk / wall, "unit": "k-mers/s", "cores": threads, "kind": kind,
            "host": {"nproc": os.cpu_count(), "model": model},
            "phase1_s": float(t1[0]), "phase2_s": float(t2[0]),
            "sample": f"first {n} reads of the workload ({nk} k-mers), {threads} threads, batches of 2000 reads, "
                      f"in-memory stores; phase1 {t1[0]:.2f}s phase2 {t2[0]:.2f}s",
            "distinct": int(distinct), "seconds": wall}


def predictor_path(args, c, wl, chars, offs, N, plain_ms, plain_st):
    """Learned sizing: train the load predictor on counter outputs of seeded genomes, predict this
    batch's per-partition superkmers and distinct k-mers from a 1% read sample, then time the
    predictor-sized path (one persistent phase-1 pass into predicted regions, count with predicted
    table regions and sweep sizes; exact fallback on overflow) against the exact path."""
    from rapidgkc_amd import native, predictor as pr
    full = args.predictor == "full"
    n_gen, glen = (200, 1_000_000) if full else (60, 500_000)
    t0 = time.perf_counter()
    X, T, _ = pr.make_dataset(c, n_genomes=n_gen, k=wl.k, p=wl.p, parts=wl.parts, genome_len=glen, log=log)
    t_data = time.perf_counter() - t0
    n_tr = int(0.8 * len(X))
    model = pr.Predictor(c, X.shape[1], wl.parts, seed=0)
    losses = model.fit(X[:n_tr], T[:n_tr], epochs=200 if full else 150, batch=16, lr=3e-4, log=log)
    m_skm, m_dist = model.calibrate(X[n_tr:], T[n_tr:])
    t_train = time.perf_counter() - t0 - t_data
    x, (sample_skm, sample_kmers, kshare), _ = pr.features(c, chars, offs, wl.k, wl.p, wl.parts, 0.01, seed=12345)
    skm_hat, dist_hat = model.predict_counts(x, N, sample_skm, sample_kmers, kshare)
    naive = pr.naive_skm(dict(N=N, sample_skm=sample_skm, sample_kmers=sample_kmers, kshare=kshare))
    cap_b, cap_r = pr.capacities(skm_hat, kshare, N, wl.k, margin=m_skm)
    # exact per-partition counts of this batch (to score the predictions)
    c.set_reads(chars, offs)
    _, skm_ex, _ = c.superkmers(wl.k, wl.p, wl.parts)
    rows_ex, _ = c.count(1, 0xFFFFFFFF)
    fallbacks = overflow_p1 = 0
    times = []
    for i in range(args.warmup + args.steps):
        ts = time.perf_counter()
        try:
            c.superkmers_sized(wl.k, wl.p, wl.parts, cap_b, cap_r)
        except native.RgkcError as e:
            if e.code != -4:
                raise
            overflow_p1 += 1
            c.superkmers(wl.k, wl.p, wl.parts)
        c.set_sizing(dist_hat, margin=m_dist)
        c.count(wl.min_count, wl.max_count)
        stp = c.stats()
        c.set_sizing(None)
        fallbacks += stp["fallbacks"]
        if i >= args.warmup:
            times.append((time.perf_counter() - ts, stp))
    ms = 1e3 * float(np.mean([t for t, _ in times]))
    stp = times[-1][1]
    res = {
        "ms_per_step": ms, "value": N / (ms * 1e-3), "speedup_vs_exact": plain_ms / ms,
        "stages_ms": {k: float(np.mean([s[k] for _, s in times])) for k in
                      ("pack_ms", "segments_ms", "p1_count_ms", "p1_emit_ms", "plan_ms", "expand_ms", "count_ms",
                       "total_ms")},
        "table_rows_alloc": int(stp["table_rows_alloc"]), "table_rows_alloc_exact": int(plain_st["table_rows_alloc"]),
        "phase1_overflows": overflow_p1, "count_fallbacks": int(fallbacks),
        "skm_rel_err": pr.relative_error(skm_hat, skm_ex.astype(np.float64)),
        "skm_rel_err_sample_scaling": pr.relative_error(naive, skm_ex.astype(np.float64)),
        "distinct_rel_err": pr.relative_error(dist_hat, rows_ex.astype(np.float64)),
        "margins": {"superkmers": m_skm, "distinct": m_dist},
        "heads": {"superkmers": "network" if model.use_net[0] else "sample scaling",
                  "distinct": "network" if model.use_net[1] else "instance bound",
                  "heldout_log_err_net_vs_baseline": model.val_err},
        "train": {"genomes": n_tr, "calibration_genomes": len(X) - n_tr, "genome_bp": glen, "data_s": t_data,
                  "train_s": t_train,
                  "loss_first": losses[0], "loss_last": losses[-1], "params": model.net.n_params},
    }
    model.net.close()
    log(f"[bench] predictor path: {ms:.2f} ms/step (exact {plain_ms:.2f}), skm err {res['skm_rel_err']:.3f}, "
        f"distinct err {res['distinct_rel_err']:.3f}, overflows {overflow_p1}, fallbacks {fallbacks}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="genome_100m_30x")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the workload's genome and reads")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-sample-reads", type=int, default=20_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--predictor", choices=["off", "quick", "full"], default="full",
                    help="learned sizing: train on seeded genomes (quick: 60 x 0.5 Mbp, full: 200 x 1 Mbp) and time "
                         "the predictor-sized path beside the exact one")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the counting path does not span GPUs.",
              file=sys.stderr)
        sys.exit(2)

    from rapidgkc_amd import native, synth
    wl = synth.WORKLOADS[args.workload]
    t0 = time.perf_counter()
    chars, offs = wl.make_reads(args.scale, threads=min(16, os.cpu_count() or 8))
    n_reads = len(offs) - 1
    lens = np.diff(offs.astype(np.int64))
    N = int(np.maximum(lens - wl.k + 1, 0).sum())
    log(f"[bench] {wl.name}: {n_reads} reads, {int(offs[-1])} bases, N={N} k-mers (generated in "
        f"{time.perf_counter() - t0:.1f}s)")

    c = native.Counter(0)
    c.set_reads(chars, offs)
    log("[bench] reads resident in HBM")
    for i in range(args.warmup):
        c.run(wl.k, wl.p, wl.parts, wl.min_count, wl.max_count)
        log(f"[bench] warmup {i} done: {c.stats()['total_ms']:.2f} ms device")
    per = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        c.run(wl.k, wl.p, wl.parts, wl.min_count, wl.max_count)
        st = c.stats()
        per.append((time.perf_counter() - ts, st))
        log(f"[bench] step {i}: wall {1e3 * per[-1][0]:.2f} ms, device {st['total_ms']:.2f} ms "
            f"(pack {st['pack_ms']:.2f} seg {st['segments_ms']:.2f} p1c {st['p1_count_ms']:.2f} "
            f"p1e {st['p1_emit_ms']:.2f} group+plan {st['plan_ms']:.2f} expand {st['expand_ms']:.2f} "
            f"count {st['count_ms']:.2f})")
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1e3 * elapsed / args.steps
    st = per[-1][1]
    assert st["n_kmers"] == N, (st["n_kmers"], N)
    W = 8 if wl.k <= 32 else 16
    count_ms = float(np.mean([s["count_ms"] for _, s in per]))
    rows = int(st["rows"])
    alg_bytes = N * W + rows * (W + 4)
    achieved = alg_bytes / (count_ms * 1e-3)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"count_traffic_{wl.name}.json")
    if os.path.exists(tpath) and args.scale == 1.0:
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    stage = {k: float(np.mean([s[k] for _, s in per])) for k in
             ("pack_ms", "segments_ms", "p1_count_ms", "p1_emit_ms", "plan_ms", "expand_ms", "count_ms", "total_ms")}
    out = {
        "metric": "k-mers counted/sec on 1 MI355X vs host CPU path; % of 8 TB/s HBM roofline",
        "value": N / (ms_per_step * 1e-3),
        "unit": "k-mers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64" if W == 8 else "u128",
        "data": "synthetic (seeded genome and reads generated on the host; reads resident in HBM)",
        "config": {"workload": wl.name, "genome_bp": int(wl.genome_len * args.scale), "reads": n_reads,
                   "read_len": wl.read_len, "error_rate": wl.err, "k": wl.k, "p": wl.p, "parts": wl.parts,
                   "min_count": wl.min_count, "kmers": N, "distinct": int(st["distinct"]), "rows": rows,
                   "superkmers": int(st["n_records"]), "store_bytes": int(st["store_bytes"])},
        "stages_ms": stage,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM, "traffic": traffic,
                     "kernel": "k_count_items", "alg_bytes_per_launch": alg_bytes, "kernel_ms": count_ms,
                     "count_stage_ms": count_ms + stage["plan_ms"] + stage["expand_ms"],
                     "count_stage_frac": alg_bytes / ((count_ms + stage["plan_ms"] + stage["expand_ms"]) * 1e-3) /
                     PEAK_HBM},
        "cpu_baseline": None,
    }
    if args.predictor != "off":
        out["predictor_path"] = predictor_path(args, c, wl, chars, offs, N, ms_per_step, st)
    c.close()
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"[bench] cpu baseline: reference CPU path, {threads} threads, {args.cpu_sample_reads} reads")
        out["cpu_baseline"] = cpu_baseline(wl, chars, offs, threads, args.cpu_sample_reads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
