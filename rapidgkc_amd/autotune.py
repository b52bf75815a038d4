"""Autotuner for the count stage's sweep size (k-mer instances planned per LDS sweep).

Small sweeps waste the fixed per-sweep work (table clear, 4096-bin scans); large ones overflow the
LDS table more often and get split and retried. The best value depends on the data's
distinct/instance ratio and on k (u64 vs u128 tables), so it is measured on the batch itself:
each candidate runs the count stage on the resident store (phase 1 is done once), timed with the
library's HIP events, and the fastest is kept. Results are cached per (device, k, p, parts) in a
JSON file next to the library.
"""
from __future__ import annotations

import json
import os

from . import native

_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_cache.json")


def _load():
    try:
        with open(_CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _key(k, p, parts):
    return f"gfx950:k{k}:p{p}:parts{parts}"


def tune_inst_cap(counter: native.Counter, k: int, p: int, parts: int, candidates=None, reps: int = 2,
                  persist: bool = True):
    """Time the count stage for each candidate on the counter's resident batch (phase 1 must have
    run with the same k, p, parts). Returns (best, {candidate: ms})."""
    if candidates is None:
        candidates = (2048, 3072, 3328, 3584, 4096) if k <= 32 else (768, 1024, 1280, 1536, 2048)
    times = {}
    for cand in candidates:
        counter.set_tuning(cand)
        best = None
        for _ in range(reps):
            counter.count(1, 0xFFFFFFFF)
            st = counter.stats()
            ms = st["plan_ms"] + st["expand_ms"] + st["count_ms"]
            best = ms if best is None else min(best, ms)
        times[int(cand)] = best
    choice = min(times, key=times.get)
    counter.set_tuning(choice)
    if persist:
        cache = _load()
        cache[_key(k, p, parts)] = {"inst_cap": choice, "ms": times}
        try:
            with open(_CACHE, "w") as f:
                json.dump(cache, f, indent=1, sort_keys=True)
        except OSError:
            pass
    return choice, times


def cached_inst_cap(k: int, p: int, parts: int):
    """The tuned value for this shape, or None."""
    v = _load().get(_key(k, p, parts))
    return None if v is None else int(v["inst_cap"])
