"""GPU parity on the five benchmark workloads at their benchmark sizes.

Each workload runs through the HIP path once (pack, minimizer scan, partition store, count) and is
checked against the oracle on the same seeded reads:
  - every partition's superkmer, k-mer and byte counters equal the oracle's store counters;
  - sampled partitions (the largest, the smallest non-empty, the one all-filtered windows route to,
    and evenly spaced ones) hold the oracle's record multiset byte for byte, and their table rows
    equal the oracle's table of the oracle's store (cpu_funcs.cpp:586-651 restated), in ascending
    key order after rgkc_sort_rows;
  - the counts of all partitions sum to N = sum(len - k + 1) when every row is kept, partitions
    are key-disjoint (the row total equals the distinct total), and the distinct total equals the
    reference's own CPU pipeline (cpu_funcs.cpp built into oracle/_ref) where that was built.
uniform_1m is compared partition by partition in full. All five run at full size (genome_1g_20x:
1 Gbp, 133.3 M reads, 16 G k-mers, counted in partition groups sized to the device memory).
"""
import time

import numpy as np
import pytest

import oracle_api as oa
from rapidgkc_amd import synth

pytestmark = pytest.mark.gpu

CONFIGS = [
    # workload, scale, sampled partitions (None: all)
    ("uniform_1m", 1.0, None),
    ("genome_100m_30x", 1.0, 16),
    ("mixed_50m", 1.0, 16),
    ("tandem_200m_k55", 1.0, 16),
    ("genome_1g_20x", 1.0, 12),
]


def say(msg):
    print(msg, flush=True)


def pick_parts(kc, parts, n, p):
    """Deterministic sample: largest, smallest non-empty, the partition of the all-filtered
    minimizer (~mask % parts), then evenly spaced partitions."""
    if n is None or n >= parts:
        return list(range(parts))
    kc = np.asarray(kc)
    nz = np.flatnonzero(kc)
    mask = (1 << (2 * p)) - 1
    pick = [int(np.argmax(kc)), int(nz[np.argmin(kc[nz])]), int((~mask & 0xFFFFFFFF) % parts)]
    for i in np.linspace(0, parts - 1, n).astype(int):
        if len(pick) >= n:
            break
        if int(i) not in pick:
            pick.append(int(i))
    return sorted(set(pick))


@pytest.mark.parametrize("cfg", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_config_parity(cfg):
    from rapidgkc_amd import native
    name, scale, n_sample = cfg
    wl = synth.WORKLOADS[name]
    t0 = time.perf_counter()
    chars, offs = wl.make_reads(scale, threads=16)
    lens = np.diff(offs.astype(np.int64))
    N = int(np.maximum(lens - wl.k + 1, 0).sum())
    say(f"[gpu] config {name} scale {scale}: {len(offs) - 1} reads, {int(offs[-1])} bases, N={N}")
    c = native.Counter(0)
    c.set_reads(chars, offs)
    po, sc, kc = c.superkmers(wl.k, wl.p, wl.parts)
    gst = c.copy_store(int(po[-1]))
    rows, distinct = c.count(wl.min_count, wl.max_count)
    c.sort_rows()
    keys, cnts, ro = c.table(rows)
    st = c.stats()
    c.close()
    say(f"[gpu]   device {st['total_ms']:.1f} ms, distinct {distinct}, rows {int(rows.sum())}, "
        f"{time.perf_counter() - t0:.1f}s")
    assert st["n_kmers"] == N
    assert int(kc.sum()) == N
    # key-disjoint partitions: every row is a distinct key of the batch
    if wl.min_count <= 1:
        assert int(rows.sum()) == distinct
        assert int(cnts.astype(np.uint64).sum()) == N, "counts do not sum to the k-mer instances"
    sample = pick_parts(kc, wl.parts, n_sample, wl.p)
    pb, ps, pk, kept = oa.store_parallel(chars, offs, wl.k, wl.p, wl.parts, keep=sample)
    say(f"[gpu]   oracle store ({len(sample)} partitions kept), {time.perf_counter() - t0:.1f}s")
    assert np.array_equal(sc, ps), "superkmer counters differ"
    assert np.array_equal(kc, pk), "k-mer counters differ"
    assert np.array_equal(np.diff(po), pb), "partition byte sizes differ"
    tabs = oa.count_parts_parallel(kept, wl.k, wl.min_count, wl.max_count)
    say(f"[gpu]   oracle tables, {time.perf_counter() - t0:.1f}s")
    for i in sample:
        a = oa.records_multiset(gst[int(po[i]):int(po[i + 1])])
        b = oa.records_multiset(kept[i])
        assert a == b, f"partition {i}: record multisets differ"
        okk, occ, _, onk = tabs[i]
        assert onk == int(kc[i])
        gk, gc = keys[ro[i]:ro[i + 1]], cnts[ro[i]:ro[i + 1]]
        assert len(gk) == len(okk), f"partition {i}: {len(gk)} rows, oracle {len(okk)}"
        assert np.array_equal(gk, okk), f"partition {i}: keys differ (ascending order)"
        assert np.array_equal(gc, occ), f"partition {i}: counts differ"
    say(f"[gpu]   {len(sample)} partitions equal, {time.perf_counter() - t0:.1f}s")
    ref = oa.ref_pipeline_distinct(chars, offs, wl.k, wl.p, wl.parts)
    if ref is not None:
        say(f"[gpu]   reference CPU pipeline distinct {ref}, {time.perf_counter() - t0:.1f}s")
        assert ref == distinct
