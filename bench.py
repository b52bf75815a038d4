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

Beside the headline, the same line carries:
  - `sort_rows_ms`: rgkc_sort_rows after a count, buffers already allocated (`sort_rows_first_ms`:
    the first call, which also allocates them) (each partition's rows in ascending key order,
    the reference's table order), and `upload`: the pinned host-to-device copy of the reads with
    the rate including it (never `value`);
  - `other_configs`: the other workloads through the same exact path (uniform 1 Mbp with min-count
    2 and a one-thread CPU run for per-core normalisation, mixed short + long reads, k=55 tandem
    repeats), the mixed workload through the batched path (rgkc_stream_*: uneven batches, one
    pinned copy and one hipGraph replay each) with capacities from a 1% sample, and the 1 Gbp
    learned-sizing workload at full size, exact and predictor-sized (the headline's trained model).
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


def host_info():
    """nproc and the CPU model as lscpu reports it (with /proc/cpuinfo's fields as a fallback)."""
    info = {"nproc": os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            key, _, val = ln.partition(":")
            key = key.strip()
            if key in ("Model name", "Vendor ID", "CPU family", "Model", "Socket(s)", "Core(s) per socket",
                       "Thread(s) per core", "CPU max MHz"):
                info[key.lower().replace(" ", "_").replace("(s)", "s")] = val.strip()
    except (OSError, ValueError):
        pass
    if "model_name" not in info:
        try:
            with open("/proc/cpuinfo") as f:
                info["model_name"] = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
    info["model"] = info.get("model_name", "")
    return info


def pinned_copy(a):
    """A copy of numpy array `a` in pinned host memory (hipHostMalloc), and a closer."""
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
    hip.hipHostFree.argtypes = [ctypes.c_void_p]
    p = ctypes.c_void_p()
    if hip.hipHostMalloc(ctypes.byref(p), max(a.nbytes, 16), 0) != 0:
        raise RuntimeError("hipHostMalloc failed")
    buf = (ctypes.c_uint8 * max(a.nbytes, 16)).from_address(p.value)
    v = np.frombuffer(buf, dtype=a.dtype, count=a.size)
    np.copyto(v, a)
    return v, (lambda: hip.hipHostFree(p))


def run_steps(c, wl, steps, warmup, tag):
    """`warmup` then `steps` timed passes of the exact path on the resident batch; returns the mean
    wall time per step and the per-step stats."""
    for i in range(warmup):
        c.run(wl.k, wl.p, wl.parts, wl.min_count, wl.max_count)
        log(f"[bench] {tag} warmup {i} done: {c.stats()['total_ms']:.2f} ms device")
    per = []
    t_start = time.perf_counter()
    for i in range(steps):
        ts = time.perf_counter()
        c.run(wl.k, wl.p, wl.parts, wl.min_count, wl.max_count)
        st = c.stats()
        per.append((time.perf_counter() - ts, st))
        log(f"[bench] {tag} step {i}: wall {1e3 * per[-1][0]:.2f} ms, device {st['total_ms']:.2f} ms "
            f"(pack {st['pack_ms']:.2f} seg {st['segments_ms']:.2f} p1c {st['p1_count_ms']:.2f} "
            f"p1e {st['p1_emit_ms']:.2f} group+plan {st['plan_ms']:.2f} expand {st['expand_ms']:.2f} "
            f"count {st['count_ms']:.2f})")
    return 1e3 * (time.perf_counter() - t_start) / steps, per


STAGES = ("pack_ms", "segments_ms", "p1_count_ms", "p1_emit_ms", "plan_ms", "expand_ms", "count_ms", "total_ms")


def summarize(wl, N, ms_per_step, per):
    """Stage times, roofline of the count kernel and of the count stage for one workload."""
    st = per[-1][1]
    W = 8 if wl.k <= 32 else 16
    count_ms = float(np.mean([s["count_ms"] for _, s in per]))
    rows = int(st["rows"])
    alg_bytes = N * W + rows * (W + 4)
    stage = {k: float(np.mean([s[k] for _, s in per])) for k in STAGES}
    cs_ms = count_ms + stage["plan_ms"] + stage["expand_ms"]
    return st, stage, {"bound": "hbm", "achieved": alg_bytes / (count_ms * 1e-3) / 1e9, "peak": PEAK_HBM / 1e9,
                       "unit": "GB/s", "frac": alg_bytes / (count_ms * 1e-3) / PEAK_HBM, "kernel": "k_count_items",
                       "alg_bytes_per_launch": alg_bytes, "kernel_ms": count_ms, "count_stage_ms": cs_ms,
                       "count_stage_frac": alg_bytes / (cs_ms * 1e-3) / PEAK_HBM}


def tune_for(c, wl, N, mode):
    """The autotuner's choice for this batch (rapidgkc_amd/autotune.py): tuned on the resident
    batch (mode "on"), taken from its cache ("cache", tuning when there is none), or the library's
    defaults ("off"). The choice stays applied on the counter."""
    from rapidgkc_amd import autotune
    if mode == "off":
        autotune.apply(c, {})
        return {"choice": "defaults"}
    choice = autotune.cached(wl.k, wl.p, wl.parts, N) if mode == "cache" else None
    if choice is not None:
        autotune.apply(c, choice)
        return {"choice": choice, "source": "cache"}
    choice, times = autotune.tune(c, wl.k, wl.p, wl.parts, N, wl.min_count, wl.max_count, reps=2, log=log)
    return {"choice": choice, "source": "tuned on this batch", "times_ms": times}


def other_config(c, name, steps, warmup, cpu_threads, cpu_sample_reads, one_core=False, tune="on"):
    """One more workload through the exact path, with its CPU baseline on a bounded sample."""
    from rapidgkc_amd import synth
    wl = synth.WORKLOADS[name]
    chars, offs = wl.make_reads(1.0, threads=min(16, os.cpu_count() or 8))
    N = int(np.maximum(np.diff(offs.astype(np.int64)) - wl.k + 1, 0).sum())
    c.set_reads(chars, offs)
    tuning = tune_for(c, wl, N, tune)
    ms, per = run_steps(c, wl, steps, warmup, name)
    st, stage, roof = summarize(wl, N, ms, per)
    res = {"workload": name, "path": "exact", "tuning": tuning, "k": wl.k, "p": wl.p, "parts": wl.parts, "min_count": wl.min_count,
           "reads": len(offs) - 1, "kmers": N, "distinct": int(st["distinct"]), "rows": int(st["rows"]),
           "value": N / (ms * 1e-3), "unit": "k-mers/s", "ms_per_step": ms, "steps": steps, "stages_ms": stage,
           "roofline": roof, "sweep_retries": int(st["sweep_retries"]), "ovf_keys": int(st.get("ovf_keys", 0)),
           "cpu_baseline": None}
    if cpu_threads:
        res["cpu_baseline"] = cpu_baseline(wl, chars, offs, cpu_threads, cpu_sample_reads)
        if one_core:
            res["cpu_baseline_1core"] = cpu_baseline(wl, chars, offs, 1, cpu_sample_reads)
    return res, (wl, chars, offs, N)


def stream_config(c, wl, chars, offs, N, steps, n_batches=64):
    """The batched path on a workload: its reads cut into uneven batches, each appended with one
    pinned copy and one hipGraph replay (rgkc_stream_add), then counted. Region capacities come from
    a 1% read sample scaled to the batch (the predictor's superkmer baseline) with a 50% margin; an
    overflow would fall back to exact sizing and is reported."""
    from rapidgkc_amd import native, predictor as pr
    x, (sample_skm, sample_kmers, kshare), _ = pr.features(c, chars, offs, wl.k, wl.p, wl.parts, 0.01, seed=777)
    skm_hat = pr.naive_skm(dict(N=N, sample_skm=sample_skm, sample_kmers=sample_kmers, kshare=kshare))
    cap_b, cap_r = pr.capacities(skm_hat, kshare, N, wl.k, margin=0.5)
    rng = np.random.default_rng(4242)
    n = len(offs) - 1
    cuts = np.sort(rng.choice(np.arange(1, n), n_batches - 1, replace=False))
    bounds = [0] + cuts.tolist() + [n]
    max_reads = max(b - a for a, b in zip(bounds[:-1], bounds[1:]))
    max_bases = max(int(offs[b] - offs[a]) for a, b in zip(bounds[:-1], bounds[1:]))
    batches = [(chars[int(offs[a]):int(offs[b])], (offs[a:b + 1] - offs[a]).astype(np.uint64))
               for a, b in zip(bounds[:-1], bounds[1:])]
    times, overflows = [], 0
    for i in range(steps + 1):
        ts = time.perf_counter()
        c.stream_begin(wl.k, wl.p, wl.parts, max_reads, max_bases, cap_b, cap_r, True)
        for bc, bo in batches:
            c.stream_add(bc, bo)
        try:
            c.stream_end()
        except native.RgkcError as e:
            if e.code != -4:
                raise
            overflows += 1
            c.set_reads(chars, offs)
            c.superkmers(wl.k, wl.p, wl.parts)
        rows, _ = c.count(wl.min_count, wl.max_count)
        if i > 0:
            times.append(time.perf_counter() - ts)
    ms = 1e3 * float(np.mean(times))
    st = c.stats()
    log(f"[bench] {wl.name} batched path: {ms:.2f} ms per pass over {n_batches} batches")
    return {"workload": wl.name, "path": "batched (rgkc_stream_*, hipGraph replay per batch)", "batches": n_batches,
            "max_batch_reads": int(max_reads), "ms_per_pass": ms, "value_incl_host_copies": N / (ms * 1e-3),
            "unit": "k-mers/s", "passes": steps, "phase1_overflows": overflows, "distinct": int(st["distinct"]),
            "note": "includes each batch's pinned host-to-device copy (the loader hands host buffers over)"}


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
    return {"value": nk / wall, "unit": "k-mers/s", "cores": threads, "kind": kind,
            "host": host_info(),
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
    # with sample scaling as the superkmer estimate, its error is the sampling error: a margin per
    # partition from its sample count; with the network, the calibrated margin
    cap_b, cap_r = pr.capacities(skm_hat, kshare, N, wl.k,
                                 margin=m_skm if model.use_net[0] else pr.sample_margin(sample_skm))
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
    log(f"[bench] predictor path: {ms:.2f} ms/step (exact {plain_ms:.2f}), skm err {res['skm_rel_err']:.3f}, "
        f"distinct err {res['distinct_rel_err']:.3f}, overflows {overflow_p1}, fallbacks {fallbacks}")
    return res, model, m_skm, m_dist


def predicted_pass(c, wl, cap_b, cap_r, dist_hat, m_dist):
    """One pass of the predictor-sized path: phase 1 into predicted regions (exact sizing if one
    overflows), the count with predicted table regions (partitions that overflow are recounted)."""
    from rapidgkc_amd import native
    over = 0
    try:
        c.superkmers_sized(wl.k, wl.p, wl.parts, cap_b, cap_r)
    except native.RgkcError as e:
        if e.code not in (-4, -3):  # a region too small, or no room for the regions: exact sizing
            raise
        over = 1
        c.superkmers(wl.k, wl.p, wl.parts)
    c.set_sizing(dist_hat, margin=m_dist)
    c.count(wl.min_count, wl.max_count)
    st = c.stats()
    c.set_sizing(None)
    return st, over


def learned_config(c, model, m_skm, m_dist, name, steps, warmup, cpu_threads, cpu_sample_reads):
    """A workload through the exact path and the predictor-sized path at full size (the learned
    sizing workload: the exact path's instance-bounded table regions do not fit the device, so it
    counts partition groups into a temporary table and compacts them; the predicted regions hold
    the table directly)."""
    from rapidgkc_amd import predictor as pr, synth
    wl = synth.WORKLOADS[name]
    t0 = time.perf_counter()
    chars, offs = wl.make_reads(1.0, threads=min(16, os.cpu_count() or 8))
    N = int(np.maximum(np.diff(offs.astype(np.int64)) - wl.k + 1, 0).sum())
    log(f"[bench] {name}: {len(offs) - 1} reads, N={N} ({time.perf_counter() - t0:.1f}s)")
    x, (sample_skm, sample_kmers, kshare), _ = pr.features(c, chars, offs, wl.k, wl.p, wl.parts, 0.01, seed=31337)
    skm_hat, dist_hat = model.predict_counts(x, N, sample_skm, sample_kmers, kshare)
    # phase-1 regions: sample scaling is the superkmer estimate here; its error is the sampling
    # error, so the margin follows each partition's sample count
    cap_b, cap_r = pr.capacities(skm_hat, kshare, N, wl.k, margin=pr.sample_margin(sample_skm))
    c.set_reads(chars, offs)
    tuning = tune_for(c, wl, N, "on")
    ms, per = run_steps(c, wl, steps, warmup, name)
    st, stage, roof = summarize(wl, N, ms, per)
    rows_ex = int(st["rows"])
    res = {"workload": name, "path": "exact", "k": wl.k, "p": wl.p, "parts": wl.parts, "reads": len(offs) - 1,
           "kmers": N, "distinct": int(st["distinct"]), "rows": rows_ex, "value": N / (ms * 1e-3), "unit": "k-mers/s",
           "ms_per_step": ms, "steps": steps, "stages_ms": stage, "roofline": roof,
           "count_groups": int(st["count_groups"]), "table_rows_alloc": int(st["table_rows_alloc"]),
           "tuning": tuning, "cpu_baseline": None}
    for _ in range(warmup):
        predicted_pass(c, wl, cap_b, cap_r, dist_hat, m_dist)
    times, sts, overs = [], [], 0
    for _ in range(steps):
        ts = time.perf_counter()
        stp, over = predicted_pass(c, wl, cap_b, cap_r, dist_hat, m_dist)
        times.append(time.perf_counter() - ts)
        sts.append(stp)
        overs += over
    pms = 1e3 * float(np.mean(times))
    stp = sts[-1]
    assert int(stp["rows"]) == rows_ex, (stp["rows"], rows_ex)
    res["predictor_path"] = {
        "ms_per_step": pms, "value": N / (pms * 1e-3), "speedup_vs_exact": ms / pms,
        "stages_ms": {k: float(np.mean([s_[k] for s_ in sts])) for k in STAGES},
        "table_rows_alloc": int(stp["table_rows_alloc"]), "table_rows_alloc_exact_bound": N,
        "count_groups": int(stp["count_groups"]), "phase1_overflows": overs,
        "partitions_recounted": int(stp["fallbacks"]),
        "margins": {"superkmers": "sample count: 0.05 + 6 / sqrt(sample superkmers)", "distinct": m_dist}}
    log(f"[bench] {name}: exact {ms:.1f} ms ({st['count_groups']} groups), predicted {pms:.1f} ms "
        f"({stp['count_groups']} groups, {stp['fallbacks']} partitions recounted)")
    if cpu_threads:
        res["cpu_baseline"] = cpu_baseline(wl, chars, offs, cpu_threads, cpu_sample_reads)
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
    ap.add_argument("--tune", choices=["on", "cache", "off"], default="on",
                    help="autotuner: tune launch and tile choices on each batch, use its cache, or library defaults")
    ap.add_argument("--configs", choices=["all", "none"], default="all",
                    help="also measure the other workloads (uniform 1 Mbp, mixed reads exact and batched, k=55)")
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
    # the upload from pinned host memory (reported beside, never part of `value`)
    pch, free_ch = pinned_copy(chars)
    pof, free_of = pinned_copy(offs)
    tu = time.perf_counter()
    c.set_reads(pch, pof)
    upload_ms = 1e3 * (time.perf_counter() - tu)
    free_ch(); free_of()
    log(f"[bench] reads resident in HBM (pinned upload {upload_ms:.1f} ms)")
    tuning = tune_for(c, wl, N, args.tune)
    ms_per_step, per = run_steps(c, wl, args.steps, args.warmup, wl.name)
    st, stage, roof = summarize(wl, N, ms_per_step, per)
    assert st["n_kmers"] == N, (st["n_kmers"], N)
    W = 8 if wl.k <= 32 else 16
    rows = int(st["rows"])
    tpath = os.path.join(ROOT, "profiles", f"count_traffic_{wl.name}.json")
    traffic = None
    if os.path.exists(tpath) and args.scale == 1.0:
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roof = {**{k: roof[k] for k in ("bound", "achieved", "peak", "unit", "frac")}, "traffic": traffic,
            **{k: roof[k] for k in ("kernel", "alg_bytes_per_launch", "kernel_ms", "count_stage_ms", "count_stage_frac")}}
    # the reference's table order: each partition's rows ascending. The first call after a count
    # also allocates the sort's second row buffer (kept for later calls); the second call is the
    # sort alone.
    ts = time.perf_counter()
    c.sort_rows()
    sort_first_ms = 1e3 * (time.perf_counter() - ts)
    ts = time.perf_counter()
    c.sort_rows()
    sort_ms = 1e3 * (time.perf_counter() - ts)
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
        "roofline": roof,
        "sort_rows_ms": sort_ms,
        "sort_rows_first_ms": sort_first_ms,
        "tuning": tuning,
        "upload": {"pinned_h2d_ms": upload_ms, "bytes": int(chars.nbytes + offs.nbytes),
                   "value_incl_upload": N / ((ms_per_step + upload_ms) * 1e-3)},
        "cpu_baseline": None,
    }
    model = None
    if args.predictor != "off":
        out["predictor_path"], model, m_skm, m_dist = predictor_path(args, c, wl, chars, offs, N, ms_per_step, st)
    threads = 0 if args.no_cpu else (args.cpu_threads or min(16, os.cpu_count() or 1))
    if threads:
        log(f"[bench] cpu baseline: reference CPU path, {threads} threads, {args.cpu_sample_reads} reads")
        out["cpu_baseline"] = cpu_baseline(wl, chars, offs, threads, args.cpu_sample_reads)
    if args.configs == "all" and args.scale == 1.0:
        del chars, offs
        osteps = max(1, min(args.steps, 5))
        others = []
        r1, _ = other_config(c, "uniform_1m", max(args.steps, 10), args.warmup, threads, 200_000, one_core=bool(threads),
                             tune=args.tune)
        others.append(r1)
        r3, (wl3, ch3, of3, n3) = other_config(c, "mixed_50m", osteps, 1, threads, 2_000_000, tune=args.tune)
        others.append(r3)
        others.append(stream_config(c, wl3, ch3, of3, n3, osteps))
        del ch3, of3
        r4, _ = other_config(c, "tandem_200m_k55", osteps, 1, threads, 2_000_000, tune=args.tune)
        others.append(r4)
        if model is not None:
            others.append(learned_config(c, model, m_skm, m_dist, "genome_1g_20x", min(args.steps, 3), 1, threads,
                                         2_000_000))
        out["other_configs"] = others
    if model is not None:
        model.net.close()
    c.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
