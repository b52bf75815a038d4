"""Autotuner for the launch and tile choices of the counting path, measured on the batch itself.

Tuned (all results are identical in every setting; only the time changes):
  - `inst_cap`: k-mer instances planned per count sweep (the LDS table's work item). Small sweeps
    pay the per-sweep fixed work more often; large ones overflow the table and are split.
  - `count_wg_per_cu`: resident workgroups per CU of the persistent count kernel (1..4; its 39 KB
    LDS table allows four).
  - `p1_lists`: whether phase 1 keeps superkmer lists for its emit pass (on: one scan of the reads
    and a list region of ~1.25 B per k-mer; off: the emit pass scans the reads again).

The search is coordinate-wise: phase-1 lists on the phase-1 time, then the sweep size and the
workgroups per CU on the count stage's time (grouping + plan + expansion + count, HIP events),
each candidate timed `reps` times on the resident batch, the fastest kept. Results are cached per
(device, k, p, parts, log2 of the batch's k-mers) in a text file next to the library, one line per
shape, which the standalone CLI (host/rgkc_main.cpp) reads as well:

    gfx950:k31:p11:parts512:n31 inst_cap=3328 count_wg_per_cu=4 p1_lists=1
"""
from __future__ import annotations

import math
import os

from . import native

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_cache.txt")


def shape_key(k: int, p: int, parts: int, n_kmers: int) -> str:
    return f"gfx950:k{k}:p{p}:parts{parts}:n{max(0, int(math.log2(max(n_kmers, 1))))}"


def load(path: str = CACHE) -> dict:
    out = {}
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) < 2:
                    continue
                out[parts[0]] = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in parts[1:] if "=" in kv}
    except OSError:
        pass
    return out


def save(key: str, choice: dict, path: str = CACHE):
    cache = load(path)
    cache[key] = choice
    try:
        with open(path, "w") as f:
            for kk in sorted(cache):
                f.write(kk + " " + " ".join(f"{a}={b}" for a, b in sorted(cache[kk].items())) + "\n")
    except OSError:
        pass


def apply(counter: native.Counter, choice: dict):
    counter.set_tuning(int(choice.get("inst_cap", 0)))
    counter.set_count_workgroups(int(choice.get("count_wg_per_cu", 0)))
    counter.set_p1_lists(0 if choice.get("p1_lists", 1) else 0xFFFFFFFF)


def _count_ms(counter, cmin, cmax, reps):
    best = None
    for _ in range(reps):
        counter.count(cmin, cmax)
        st = counter.stats()
        ms = st["plan_ms"] + st["expand_ms"] + st["count_ms"]
        best = ms if best is None else min(best, ms)
    return best


def _p1_ms(counter, k, p, parts, reps):
    best = None
    for _ in range(reps):
        counter.superkmers(k, p, parts)
        st = counter.stats()
        ms = st["p1_count_ms"] + st["p1_emit_ms"]
        best = ms if best is None else min(best, ms)
    return best


def tune(counter: native.Counter, k: int, p: int, parts: int, n_kmers: int, cmin: int = 1,
         cmax: int = 0xFFFFFFFF, reps: int = 2, persist: bool = True, log=None):
    """Tune on the counter's resident read batch; returns (choice, {setting: ms}). The counter is
    left with the choice applied and a fresh phase 1 done."""
    times = {}
    choice = {"inst_cap": 0, "count_wg_per_cu": 0, "p1_lists": 1}
    # phase 1: superkmer lists on / off
    best_p1 = None
    for lists in (1, 0):
        counter.set_p1_lists(0 if lists else 0xFFFFFFFF)
        ms = _p1_ms(counter, k, p, parts, reps)
        times[f"p1_lists={lists}"] = ms
        if best_p1 is None or ms < best_p1:
            best_p1, choice["p1_lists"] = ms, lists
    counter.set_p1_lists(0 if choice["p1_lists"] else 0xFFFFFFFF)
    counter.superkmers(k, p, parts)
    # count stage: sweep size, then workgroups per CU
    caps = (2560, 3072, 3328, 3584, 4096) if k <= 32 else (768, 1024, 1280, 1536)
    best = None
    for cap in caps:
        counter.set_tuning(cap)
        ms = _count_ms(counter, cmin, cmax, reps)
        times[f"inst_cap={cap}"] = ms
        if best is None or ms < best:
            best, choice["inst_cap"] = ms, cap
    counter.set_tuning(choice["inst_cap"])
    best = None
    for wg in (2, 3, 4):
        counter.set_count_workgroups(wg)
        ms = _count_ms(counter, cmin, cmax, reps)
        times[f"count_wg_per_cu={wg}"] = ms
        if best is None or ms < best:
            best, choice["count_wg_per_cu"] = ms, wg
    apply(counter, choice)
    if log:
        log(f"[autotune] {shape_key(k, p, parts, n_kmers)}: {choice} ({', '.join(f'{a} {b:.2f}' for a, b in times.items())})")
    if persist:
        save(shape_key(k, p, parts, n_kmers), choice)
    return choice, times


def cached(k: int, p: int, parts: int, n_kmers: int):
    """The tuned choice for this shape, or None."""
    return load().get(shape_key(k, p, parts, n_kmers))
