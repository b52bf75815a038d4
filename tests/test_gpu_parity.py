"""GPU parity: the gfx950 path against the oracle on the same seeded inputs.

Integer work, so the bar is bit-exact:
  - phase 1: per-partition byte/record/k-mer counters equal; each partition holds the same
    multiset of superkmer records byte for byte (neighbouring-base bits included) -- the record
    order inside a partition is not defined by the reference either (skmstore.hpp:62-99);
  - count: each partition's table holds the same (k-mer, count) rows after the [min, max] filter;
    rows are compared after ordering by key (the GPU groups rows by minimizer sub-bucket).
"""
import zlib

import numpy as np
import pytest

import oracle_api as oa
from rapidgkc_amd import synth

pytestmark = pytest.mark.gpu


def _native():
    from rapidgkc_amd import native
    return native


def say(msg):
    print(msg, flush=True)


def split_records(st):
    """Records of a partition byte stream: a record ends at a non-full byte that follows a full
    byte (cpu_funcs.cpp:516-525)."""
    if len(st) == 0:
        return []
    full = st >= 0xC0
    end = np.zeros(len(st), dtype=bool)
    end[1:] = (~full[1:]) & full[:-1]
    ends = np.flatnonzero(end) + 1
    starts = np.concatenate([[0], ends[:-1]])
    return sorted(st[s:e].tobytes() for s, e in zip(starts, ends))


def check_store(c, chars, offs, k, p, parts):
    po, sc, kc = c.superkmers(k, p, parts)
    st, opo, ops, opk = oa.store(chars, offs, k, p, parts)
    assert np.array_equal(sc, ops), "skm_cnt differs"
    assert np.array_equal(kc, opk), "kmer_cnt differs"
    assert np.array_equal(po, opo), "partition byte sizes differ"
    gst = c.copy_store(int(po[-1]))
    for i in range(parts):
        a = split_records(gst[po[i]:po[i + 1]])
        b = split_records(st[opo[i]:opo[i + 1]])
        assert a == b, f"partition {i}: record multisets differ"
    return st, opo


def check_table(c, st, opo, k, parts, cmin, cmax):
    rows, distinct = c.count(cmin, cmax)
    keys, cnts, ro = c.table(rows)
    ok_, oc_, oro = oa.count_store(st, opo, k, cmin, cmax)
    assert np.array_equal(rows, np.diff(oro)), "rows per partition differ"
    for i in range(parts):
        a = oa.sort_table(keys[ro[i]:ro[i + 1]], cnts[ro[i]:ro[i + 1]])
        assert np.array_equal(a[0], ok_[oro[i]:oro[i + 1]]), f"partition {i}: keys differ"
        assert np.array_equal(a[1], oc_[oro[i]:oro[i + 1]]), f"partition {i}: counts differ"
    # unfiltered distinct equals the brute-force distinct count
    return rows, distinct


CASES = [
    # name, genome bp, reads, read len, err, k, p, parts, cmin, cmax
    ("k31_p11_512", 200_000, 20_000, 150, 0.01, 31, 11, 512, 1, 0xFFFFFFFF),
    ("k31_p11_64_min2", 100_000, 12_000, 150, 0.005, 31, 11, 64, 2, 0xFFFFFFFF),
    ("k21_p7_31", 100_000, 8_000, 120, 0.01, 21, 7, 31, 1, 1000),
    ("k25_p9_512", 150_000, 10_000, 150, 0.005, 25, 9, 512, 1, 0xFFFFFFFF),
    ("k55_p13_128", 150_000, 8_000, 250, 0.01, 55, 13, 128, 1, 0xFFFFFFFF),
    ("k33_p11_64", 80_000, 5_000, 150, 0.01, 33, 11, 64, 1, 0xFFFFFFFF),
    ("k32_p11_64", 80_000, 5_000, 150, 0.01, 32, 11, 64, 2, 100),
    # more than 512 partitions: the phase-1 kernels' full-size partition tables
    ("k31_p11_1000", 100_000, 8_000, 150, 0.01, 31, 11, 1000, 1, 0xFFFFFFFF),
    ("k55_p13_700", 80_000, 5_000, 150, 0.01, 55, 13, 700, 1, 0xFFFFFFFF),
]


SPLIT_CASES = [
    # name, genome bp, reads, k, p, sweep size: one partition and a short minimizer make a few
    # sub-buckets hold far more distinct k-mers than one table; the huge sweep size makes the
    # first sweeps span many sub-buckets, so both split kinds (sub-bucket ranges, hash slices) run
    ("k31_p7_1part", 400_000, 30_000, 31, 7, 1 << 22),
    ("k55_p7_1part", 300_000, 12_000, 55, 7, 1 << 22),
]


@pytest.mark.parametrize("case", SPLIT_CASES, ids=[c[0] for c in SPLIT_CASES])
def test_count_split_paths(case):
    name, glen, nr, k, p, sweep = case
    say(f"[gpu] split case {name}")
    g = synth.genome(zlib.crc32(name.encode()) & 0xFFFF, glen, gc=0.45)
    chars, offs = synth.short_reads(g, 9, nr, 150, 0.01)
    c = _native().Counter(0)
    c.set_reads(chars, offs)
    c.set_tuning(sweep)
    st, opo = check_store(c, chars, offs, k, p, 1)
    rows, distinct = check_table(c, st, opo, k, 1, 1, 0xFFFFFFFF)
    s = c.stats()
    say(f"[gpu] sweeps {s['sweeps']} retries {s['sweep_retries']}")
    assert s["sweep_retries"] > 0, "the case did not exercise the split paths"
    _, _, bd = oa.bruteforce(chars, offs, k)
    assert distinct == bd
    # other sweep sizes cut the sub-bucket ranges at other places (ranges that start or end on
    # empty sub-buckets included): the same table every time
    # heavy sub-buckets (more instances than a work item) are read from expanded keys or, with
    # expansion off, from their records: the same table either way
    expanded = 0
    for alt in (3_000, 12_000, 33_000):
        for mode in (0, 1):
            c.set_tuning(alt)
            c.set_heavy_expand(mode)
            rows2, distinct2 = c.count(1, 0xFFFFFFFF)
            assert distinct2 == bd and np.array_equal(rows2, rows), f"sweep size {alt}, expand mode {mode}"
            expanded += c.stats()["heavy_groups"] if mode == 0 else 0
    assert expanded > 0, "no sub-bucket was expanded into keys"
    keys2, cnts2, ro2 = c.table(rows2)
    for i in range(1):
        a2 = oa.sort_table(keys2[ro2[i]:ro2[i + 1]], cnts2[ro2[i]:ro2[i + 1]])
        ok_, oc_, oro = oa.count_store(st, opo, k, 1, 0xFFFFFFFF)
        assert np.array_equal(a2[0], ok_[oro[i]:oro[i + 1]]) and np.array_equal(a2[1], oc_[oro[i]:oro[i + 1]])
    c.close()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_short_reads(case):
    name, glen, nr, rl, err, k, p, parts, cmin, cmax = case
    say(f"[gpu] case {name}")
    g = synth.genome(zlib.crc32(name.encode()) & 0xFFFF, glen, gc=0.45)
    chars, offs = synth.short_reads(g, 5, nr, rl, err)
    c = _native().Counter(0)
    c.set_reads(chars, offs)
    st, opo = check_store(c, chars, offs, k, p, parts)
    rows, distinct = check_table(c, st, opo, k, parts, cmin, cmax)
    _, _, bd = oa.bruteforce(chars, offs, k)
    assert distinct == bd
    c.close()


def test_long_and_mixed_reads():
    say("[gpu] case long+short mixed, k25 p9")
    k, p, parts = 25, 9, 256
    g = synth.genome(77, 300_000, gc=0.5, tandem_frac=0.05)
    short = synth.short_reads(g, 78, 6000, 150, 0.005)
    lr = synth.long_reads(g, 79, 60, mean=6000, sd=4000, err=0.05, max_len=40000)
    chars, offs = synth.interleave_csr(short, lr, 80)
    c = _native().Counter(0)
    c.set_reads(chars, offs)
    st, opo = check_store(c, chars, offs, k, p, parts)
    check_table(c, st, opo, k, parts, 1, 0xFFFFFFFF)
    c.close()


def test_p1_list_and_rescan_paths():
    """The emit pass writes records from the count pass's superkmer lists; a workgroup whose list
    overflows (a tiny list capacity here) is rescanned, and lists can be switched off: all three
    give the oracle's store."""
    k, p, parts = 25, 9, 64
    g = synth.genome(91, 200_000, gc=0.5, tandem_frac=0.05)
    short = synth.short_reads(g, 92, 5000, 150, 0.005)
    lr = synth.long_reads(g, 93, 30, mean=5000, sd=3000, err=0.05, max_len=30000)
    chars, offs = synth.interleave_csr(short, lr, 94)
    c = _native().Counter(0)
    c.set_reads(chars, offs)
    for cap in (0, 64, 0xFFFFFFFF):
        say(f"[gpu] case list cap {cap:#x}, k25 p9")
        c.set_p1_lists(cap)
        st, opo = check_store(c, chars, offs, k, p, parts)
        if cap == 64:
            assert c.stats()["p1_rescans"] > 0, "a 64-entry list must overflow and be rescanned"
    check_table(c, st, opo, k, parts, 1, 0xFFFFFFFF)
    c.close()


def test_edge_reads():
    """Reads shorter than k (dropped), exactly k, N and lower-case bases, homopolymers whose
    p-mers are all filtered (minimizer = mask), a tandem repeat (one long superkmer)."""
    say("[gpu] case edge reads")
    k, p, parts = 31, 11, 31
    rng = np.random.default_rng(5)
    reads = [b"ACGT" * 3, b"A" * 30, bytes(rng.choice(list(b"ACGT"), 31)), b"A" * 200, b"AC" * 150,
             bytes(rng.choice(list(b"ACGTNacgtn"), 300)), b"T" * 31 + b"G" * 40, b"", b"ACGTACGTAC" * 40]
    reads += [bytes(rng.choice(list(b"ACGT"), int(rng.integers(20, 400)))) for _ in range(500)]
    offs = np.zeros(len(reads) + 1, dtype=np.uint64)
    np.cumsum([len(r) for r in reads], out=offs[1:])
    chars = np.frombuffer(b"".join(reads), dtype=np.uint8).copy()
    c = _native().Counter(0)
    c.set_reads(chars, offs)
    st, opo = check_store(c, chars, offs, k, p, parts)
    check_table(c, st, opo, k, parts, 1, 0xFFFFFFFF)
    c.close()


def test_golden_fixture_on_gpu():
    """The reference's own outputs (tests/golden) reproduced by the GPU path."""
    import os
    say("[gpu] case golden fixture")
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_small.npz"))
    for tag in ("k31", "k55"):
        k, p, parts = (int(x) for x in z[f"{tag}_kpp"])
        c = _native().Counter(0)
        c.set_reads(z["chars"], z["offs"])
        po, sc, kc = c.superkmers(k, p, parts)
        assert np.array_equal(po, z[f"{tag}_part_offs"])
        assert np.array_equal(sc, z[f"{tag}_skm_cnt"]) and np.array_equal(kc, z[f"{tag}_kmer_cnt"])
        gst = c.copy_store(int(po[-1]))
        rst = z[f"{tag}_store"]
        for i in range(parts):
            assert split_records(gst[po[i]:po[i + 1]]) == split_records(rst[po[i]:po[i + 1]])
        rows, _ = c.count(1, 0xFFFFFFFF)
        keys, cnts, ro = c.table(rows)
        kk = z[f"{tag}_keys"]
        if k > 32:
            kk = kk.view(oa.key_dtype(k)).reshape(-1)
        rro = z[f"{tag}_row_offs"]
        for i in range(parts):
            a = oa.sort_table(keys[ro[i]:ro[i + 1]], cnts[ro[i]:ro[i + 1]])
            assert np.array_equal(a[0], kk[rro[i]:rro[i + 1]]) and np.array_equal(a[1], z[f"{tag}_counts"][rro[i]:rro[i + 1]])
        c.close()


@pytest.mark.parametrize("k,p,parts", [(31, 11, 64), (55, 13, 32), (21, 7, 31)])
def test_host_store_count(k, p, parts):
    """kmc_counting_GPU_streams on reference-format partitions built on the host (the oracle's
    store, byte-identical to GenSuperkmerCPU's): the GPU recovers record boundaries and minimizers
    from the bytes alone."""
    say(f"[gpu] case host store k{k} p{p} parts{parts}")
    g = synth.genome(91, 120_000, gc=0.43)
    chars, offs = synth.short_reads(g, 92, 9000, 150, 0.01)
    st, opo, ops, opk = oa.store(chars, offs, k, p, parts)
    c = _native().Counter(0)
    rows, distinct = c.count_host_store(st, opo, parts, k, p, 1, 0xFFFFFFFF)
    keys, cnts, ro = c.table(rows)
    ok_, oc_, oro = oa.count_store(st, opo, k)
    assert np.array_equal(rows, np.diff(oro))
    for i in range(parts):
        a = oa.sort_table(keys[ro[i]:ro[i + 1]], cnts[ro[i]:ro[i + 1]])
        assert np.array_equal(a[0], ok_[oro[i]:oro[i + 1]]) and np.array_equal(a[1], oc_[oro[i]:oro[i + 1]])
    c.close()


@pytest.mark.parametrize("use_graph", [True, False])
def test_stream_batches(use_graph):
    """Batched input: short and long reads interleaved into uneven batches, appended into regions
    sized exactly (no slack) by the oracle's store sizes, with and without the captured graph.
    The partitions hold the oracle's records and count to the oracle's table; capacities one
    record short overflow."""
    say(f"[gpu] case stream batches graph={use_graph}")
    k, p, parts = 25, 9, 64
    g = synth.genome(55, 200_000, gc=0.5)
    short = synth.short_reads(g, 56, 8000, 150, 0.005)
    lr = synth.long_reads(g, 57, 40, mean=5000, sd=3000, err=0.05, max_len=30000)
    chars, offs = synth.interleave_csr(short, lr, 58)
    st, opo, ops, opk = oa.store(chars, offs, k, p, parts)
    cap_b = np.diff(opo).astype(np.uint64)
    cap_r = ops.astype(np.uint64)
    rng = np.random.default_rng(9)
    cuts = np.sort(rng.choice(np.arange(1, len(offs) - 1), 11, replace=False))
    bounds = [0] + cuts.tolist() + [len(offs) - 1]
    max_reads = max(b - a for a, b in zip(bounds[:-1], bounds[1:]))
    max_bases = max(int(offs[b] - offs[a]) for a, b in zip(bounds[:-1], bounds[1:]))
    c = _native().Counter(0)
    c.stream_begin(k, p, parts, max_reads, max_bases, cap_b, cap_r, use_graph)
    for a, b in zip(bounds[:-1], bounds[1:]):
        bo = (offs[a:b + 1] - offs[a]).astype(np.uint64)
        c.stream_add(chars[int(offs[a]):int(offs[b])], bo)
    pb, sc, kc = c.stream_end()
    assert np.array_equal(pb, cap_b) and np.array_equal(sc, ops) and np.array_equal(kc, opk)
    gst = c.copy_store(int(pb.sum()))
    go = np.concatenate([[0], np.cumsum(pb)]).astype(np.int64)
    for i in range(parts):
        assert split_records(gst[go[i]:go[i + 1]]) == split_records(st[opo[i]:opo[i + 1]])
    check_table(c, st, opo, k, parts, 1, 0xFFFFFFFF)
    # one partition's capacity one record short: overflow reported
    cap_r2 = cap_r.copy()
    j = int(np.argmax(cap_r2))
    cap_r2[j] -= 1
    c.stream_begin(k, p, parts, max_reads, max_bases, cap_b, cap_r2, use_graph)
    for a, b in zip(bounds[:-1], bounds[1:]):
        c.stream_add(chars[int(offs[a]):int(offs[b])], (offs[a:b + 1] - offs[a]).astype(np.uint64))
    with pytest.raises(_native().RgkcError) as e:
        c.stream_end()
    assert e.value.code == -4
    c.close()


def test_sized_single_batch():
    """Phase 1 into capacity-sized regions (the predictor-sized path) in one persistent pass:
    exact capacities reproduce the oracle's store and table; one byte short overflows."""
    say("[gpu] case sized single batch")
    k, p, parts = 31, 11, 256
    g = synth.genome(61, 150_000, gc=0.42)
    chars, offs = synth.short_reads(g, 62, 15_000, 150, 0.01)
    st, opo, ops, opk = oa.store(chars, offs, k, p, parts)
    c = _native().Counter(0)
    c.set_reads(chars, offs)
    # superkmer lists (runs reserved per tile), a tiny list (tiles rescanned by the persistent
    # emit) and no lists (the persistent emit alone): the same store
    for lcap in (64, 0xFFFFFFFF, 0):
        c.set_p1_lists(lcap)
        pb, sc, kc = c.superkmers_sized(k, p, parts, np.diff(opo), ops)
        assert np.array_equal(pb, np.diff(opo)) and np.array_equal(sc, ops) and np.array_equal(kc, opk), lcap
        if lcap == 64:
            assert c.stats()["p1_rescans"] > 0
        gst = c.copy_store(int(pb.sum()))
        for i in range(parts):
            assert split_records(gst[opo[i]:opo[i + 1]]) == split_records(st[opo[i]:opo[i + 1]]), (lcap, i)
    check_table(c, st, opo, k, parts, 2, 0xFFFFFFFF)
    cap = np.diff(opo).astype(np.uint64)
    cap[int(np.argmax(cap))] -= 1
    with pytest.raises(_native().RgkcError) as e:
        c.superkmers_sized(k, p, parts, cap, ops)
    assert e.value.code == -4
    c.close()


@pytest.mark.parametrize("k,p,gc", [(31, 11, 0.3), (55, 13, 0.3), (21, 7, 0.5)])
def test_sort_rows_levels(k, p, gc):
    """rgkc_sort_rows: one partition of millions of rows (several radix levels: an AT-rich genome
    puts far more than one LDS sort of rows under the all-A digit) and a 64-partition table; the
    rows equal the oracle's table row for row, unsorted on the host."""
    say(f"[gpu] case sort rows k{k} p{p} gc{gc}")
    g = synth.genome(31 + k, 2_000_000, gc=gc)
    chars, offs = synth.short_reads(g, 32 + k, 150_000, 150, 0.01)
    c = _native().Counter(0)
    c.set_reads(chars, offs)
    for parts in (1, 64):
        st, opo = check_store(c, chars, offs, k, p, parts)
        rows, _ = c.count(1, 0xFFFFFFFF)
        c.sort_rows()
        keys, cnts, ro = c.table(rows)
        ok_, oc_, oro = oa.count_store(st, opo, k)
        assert np.array_equal(rows, np.diff(oro))
        assert np.array_equal(keys, ok_), f"parts {parts}: keys not in the oracle's order"
        assert np.array_equal(cnts, oc_), f"parts {parts}: counts differ"
    c.close()


def test_p1_list_order_branch():
    """Tiles with more superkmers than the list emit can put in store order (k=21, p=7 over long
    reads with 5% errors: ~4000 superkmers per 256-segment tile) are written in list order; the
    store is the oracle's either way."""
    say("[gpu] case list-order tiles k21 p7")
    k, p, parts = 21, 7, 128
    g = synth.genome(101, 300_000, gc=0.5)
    lr = synth.long_reads(g, 102, 60, mean=8000, sd=3000, err=0.05, max_len=30000)
    c = _native().Counter(0)
    c.set_reads(*lr)
    st, opo = check_store(c, lr[0], lr[1], k, p, parts)
    assert c.stats()["p1_list_order"] > 0, "no tile took the list-order branch"
    check_table(c, st, opo, k, parts, 1, 0xFFFFFFFF)
    c.close()


@pytest.mark.parametrize("k,p,limit_mb", [(31, 11, 100), (55, 13, 150)])
def test_count_in_groups(k, p, limit_mb):
    """The count stage under a device memory limit: partitions are counted in groups one after
    another, each into a temporary table whose rows are compacted into the final one (exact sizing),
    and into predicted regions directly (learned sizing), where the partitions whose predicted region
    is too small are counted again alone. Every partition's table equals the oracle's."""
    say(f"[gpu] case count in groups k={k} limit {limit_mb} MB")
    parts = 64
    g = synth.genome(71, 400_000, gc=0.45)
    chars, offs = synth.short_reads(g, 72, 40_000, 150, 0.01)
    c = _native().Counter(0)
    c.set_reads(chars, offs)
    st, opo = check_store(c, chars, offs, k, p, parts)
    c.set_memory_limit(limit_mb << 20)
    rows, _ = check_table(c, st, opo, k, parts, 1, 0xFFFFFFFF)
    s1 = c.stats()
    say(f"[gpu]   exact: {s1['count_groups']} groups, {s1['table_rows_alloc']} rows allocated for {int(rows.sum())}")
    assert s1["count_groups"] > 1
    assert s1["table_rows_alloc"] < s1["n_kmers"], "the final table was not compacted"
    # predicted regions: half the true distinct k-mers for every other partition (those overflow)
    pred = rows.astype(np.float64) * np.where(np.arange(parts) % 2 == 0, 0.4, 1.1)
    c.set_sizing(pred, margin=0.0)
    check_table(c, st, opo, k, parts, 1, 0xFFFFFFFF)
    s2 = c.stats()
    say(f"[gpu]   predicted: {s2['count_groups']} groups, {s2['fallbacks']} partitions counted again")
    assert s2["fallbacks"] >= 1
    c.set_sizing(None)
    # the row order of the reference's table on both tables
    c.set_sizing(pred, margin=0.0)
    c.count(1, 0xFFFFFFFF)
    c.sort_rows()
    keys, cnts, ro = c.table(rows)
    ok_, oc_, oro = oa.count_store(st, opo, k, 1, 0xFFFFFFFF)
    for i in range(parts):
        assert np.array_equal(keys[ro[i]:ro[i + 1]], ok_[oro[i]:oro[i + 1]]), f"partition {i}: order differs"
    c.close()


def test_autotuned_settings_keep_parity():
    """The autotuner's search (phase-1 lists, sweep size, count workgroups per CU) on a batch, then
    the tuned choice and the search's extreme settings each give the oracle's store and table."""
    from rapidgkc_amd import autotune
    say("[gpu] case autotuner")
    k, p, parts = 31, 11, 128
    g = synth.genome(81, 300_000, gc=0.42)
    chars, offs = synth.short_reads(g, 82, 30_000, 150, 0.01)
    c = _native().Counter(0)
    c.set_reads(chars, offs)
    n_kmers = int(np.maximum(np.diff(offs.astype(np.int64)) - k + 1, 0).sum())
    choice, times = autotune.tune(c, k, p, parts, n_kmers, persist=False)
    say(f"[gpu]   tuned {choice}")
    assert set(choice) == {"inst_cap", "count_wg_per_cu", "p1_lists"} and len(times) >= 8
    for setting in (choice, {"inst_cap": 2560, "count_wg_per_cu": 1, "p1_lists": 0},
                    {"inst_cap": 4096, "count_wg_per_cu": 4, "p1_lists": 1}):
        autotune.apply(c, setting)
        st, opo = check_store(c, chars, offs, k, p, parts)
        check_table(c, st, opo, k, parts, 1, 0xFFFFFFFF)
    autotune.apply(c, {})
    c.close()
