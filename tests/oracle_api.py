"""ctypes access to the oracle (oracle/liboracle.so) and to the reference's own CPU code built
into oracle/_ref (when present). Test infrastructure only: used as the checker."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = ctypes.c_void_p
i32, i64, u32, u64 = ctypes.c_int, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64

_or = None
_ref = {}


def _p(a):
    return None if a is None else a.ctypes.data


def oracle():
    global _or
    if _or is None:
        L = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
        L.or_pmer_valid.argtypes = [u32, i32]; L.or_pmer_valid.restype = i32
        L.or_pack_read.argtypes = [P, i32, P]; L.or_pack_read.restype = i32
        L.or_kmer_minimizers.argtypes = [P, i32, i32, i32, P]; L.or_kmer_minimizers.restype = i32
        L.or_skm_offsets.argtypes = [P, i32, i32, i32, P, P]; L.or_skm_offsets.restype = i32
        L.or_partition.argtypes = [u32, u32]; L.or_partition.restype = u32
        L.or_skm_bytes.argtypes = [i32, i32, i32]; L.or_skm_bytes.restype = i32
        L.or_encode_skm.argtypes = [P, i32, i32, i32, P]; L.or_encode_skm.restype = None
        L.or_revcomp64.argtypes = [u64, i32]; L.or_revcomp64.restype = u64
        L.or_store_sizes.argtypes = [P, P, i64, i32, i32, i32, P, P, P]; L.or_store_sizes.restype = i32
        L.or_store_fill.argtypes = [P, P, i64, i32, i32, i32, P, P]; L.or_store_fill.restype = i32
        L.or_extract_kmers.argtypes = [P, u64, i32, P]; L.or_extract_kmers.restype = i64
        L.or_sort_keys.argtypes = [P, i64, i32]; L.or_sort_keys.restype = None
        L.or_rle.argtypes = [P, i64, i32, u32, u32, P, P, P]; L.or_rle.restype = i64
        L.or_count_partition.argtypes = [P, u64, i32, u32, u32, P, P, P, P]; L.or_count_partition.restype = i64
        L.or_bruteforce_count.argtypes = [P, P, i64, i32, u32, u32, P, P, P]; L.or_bruteforce_count.restype = i64
        _or = L
    return _or


def ref(width):
    """The reference's cpu_funcs.cpp built with 64-bit (width=8) or 128-bit (width=16) keys, or None."""
    if width not in _ref:
        path = os.path.join(ROOT, "oracle", "_ref", "libref_cpu64.so" if width == 8 else "libref_cpu128.so")
        if not os.path.exists(path):
            _ref[width] = None
        else:
            L = ctypes.CDLL(path)
            L.ref_gen_skm_offs.argtypes = [P, i32, i32, i32, P, P]; L.ref_gen_skm_offs.restype = i32
            L.ref_read_compression.argtypes = [P, i32, P]; L.ref_read_compression.restype = None
            L.ref_gen_superkmers.argtypes = [P, P, i64, i32, i32, i32, ctypes.POINTER(P), P, P, P]
            L.ref_gen_superkmers.restype = i32
            L.ref_free.argtypes = [P]; L.ref_free.restype = None
            L.ref_extract_sorted.argtypes = [P, u64, u64, u64, i32, P]; L.ref_extract_sorted.restype = i64
            L.ref_count_partition.argtypes = [P, u64, u64, u64, i32]; L.ref_count_partition.restype = i64
            L.ref_cpu_pipeline.argtypes = [P, P, i64, i32, i32, i32, i32, i32, P, P]; L.ref_cpu_pipeline.restype = i64
            _ref[width] = L
    return _ref[width]


def key_width(k):
    return 8 if k <= 32 else 16


def key_dtype(k):
    """Keys as numpy rows: uint64 for k <= 32, (lo, hi) uint64 pairs viewed as a structured scalar."""
    return np.uint64 if k <= 32 else np.dtype([("lo", "<u8"), ("hi", "<u8")])


# ---- oracle conveniences -------------------------------------------------------------------
def minimizers(read: bytes, k, p):
    a = np.frombuffer(read, dtype=np.uint8).copy()
    mm = np.zeros(max(len(read), 1), dtype=np.uint32)
    n = oracle().or_kmer_minimizers(_p(a), len(read), k, p, _p(mm))
    return mm[:n]


def skm_offsets(read: bytes, k, p):
    a = np.frombuffer(read, dtype=np.uint8).copy()
    offs = np.zeros(len(read) + 2, dtype=np.int32)
    mms = np.zeros(len(read) + 1, dtype=np.uint32)
    n = oracle().or_skm_offsets(_p(a), len(read), k, p, _p(offs), _p(mms))
    return offs[: n + 1], mms[:n]


def pack(read: bytes):
    a = np.frombuffer(read, dtype=np.uint8).copy()
    out = np.zeros(len(read) // 3 + 1, dtype=np.uint8)
    oracle().or_pack_read(_p(a), len(read), _p(out))
    return out


def store(chars, offs, k, p, parts):
    """Partition store of a read batch: (bytes, part_offs[parts+1], skm_cnt, kmer_cnt)."""
    O = oracle()
    pb = np.zeros(parts, dtype=np.uint64); ps = np.zeros(parts, dtype=np.uint64); pk = np.zeros(parts, dtype=np.uint64)
    rc = O.or_store_sizes(_p(chars), _p(offs), len(offs) - 1, k, p, parts, _p(pb), _p(ps), _p(pk))
    assert rc == 0, rc
    po = np.zeros(parts + 1, dtype=np.uint64)
    np.cumsum(pb, out=po[1:])
    st = np.zeros(max(int(po[-1]), 1), dtype=np.uint8)
    rc = O.or_store_fill(_p(chars), _p(offs), len(offs) - 1, k, p, parts, _p(po), _p(st))
    assert rc == 0, rc
    return st[: int(po[-1])], po, ps, pk


def count_partition(skms, k, cmin=1, cmax=0xFFFFFFFF):
    """(keys, counts, distinct_unfiltered, n_kmers) of one partition's bytes."""
    n_max = len(skms) * 3 + 1
    keys = np.zeros(n_max, dtype=key_dtype(k))
    cnts = np.zeros(n_max, dtype=np.uint32)
    d = np.zeros(1, dtype=np.int64); nk = np.zeros(1, dtype=np.int64)
    buf = np.ascontiguousarray(skms, dtype=np.uint8)
    rows = oracle().or_count_partition(_p(buf), len(buf), k, cmin, cmax, _p(keys), _p(cnts), _p(d), _p(nk))
    assert rows >= 0
    return keys[:rows], cnts[:rows], int(d[0]), int(nk[0])


def count_store(st, po, k, cmin=1, cmax=0xFFFFFFFF):
    """Per-partition tables concatenated in partition order, plus per-partition row offsets."""
    ks, cs, ro = [], [], [0]
    for i in range(len(po) - 1):
        kk, cc, _, _ = count_partition(st[po[i]:po[i + 1]], k, cmin, cmax)
        ks.append(kk); cs.append(cc); ro.append(ro[-1] + len(kk))
    return np.concatenate(ks), np.concatenate(cs), np.array(ro, dtype=np.int64)


def bruteforce(chars, offs, k, cmin=1, cmax=0xFFFFFFFF):
    nk = int(sum(max(0, int(l) - k + 1) for l in np.diff(offs)))
    keys = np.zeros(max(nk, 1), dtype=key_dtype(k))
    cnts = np.zeros(max(nk, 1), dtype=np.uint32)
    d = np.zeros(1, dtype=np.int64)
    rows = oracle().or_bruteforce_count(_p(chars), _p(offs), len(offs) - 1, k, cmin, cmax, _p(keys), _p(cnts), _p(d))
    return keys[:rows], cnts[:rows], int(d[0])


def sort_table(keys, cnts):
    """Rows sorted by key (for multiset comparison of tables in different row orders)."""
    if keys.dtype == np.uint64:
        o = np.argsort(keys, kind="stable")
    else:
        o = np.lexsort((keys["lo"], keys["hi"]))
    return keys[o], cnts[o]


# ---- large batches: the oracle over read chunks on several threads ---------------------------
# The oracle's C functions hold no shared state and ctypes releases the GIL while they run, so
# chunks of a batch are processed on Python threads in parallel. A partition's records are those
# of every chunk (a record never spans reads), so per-chunk stores concatenated in chunk order hold
# the whole batch's record multiset.
def _chunk_bounds(n, parts):
    step = max(1, (n + parts - 1) // parts)
    return [(a, min(a + step, n)) for a in range(0, n, step)]


def store_parallel(chars, offs, k, p, parts, keep=(), threads=16):
    """Per-partition (bytes, superkmers, k-mers) of the whole batch, plus the record bytes of the
    partitions in `keep` as {part: uint8 array}."""
    from concurrent.futures import ThreadPoolExecutor
    O = oracle()
    keep = sorted(set(int(x) for x in keep))
    offs = np.ascontiguousarray(offs, dtype=np.uint64)
    chars = np.ascontiguousarray(chars, dtype=np.uint8)

    def one(bounds):
        a, b = bounds
        o = offs[a:b + 1]
        pb = np.zeros(parts, dtype=np.uint64); ps = np.zeros(parts, dtype=np.uint64)
        pk = np.zeros(parts, dtype=np.uint64)
        rc = O.or_store_sizes(_p(chars), _p(o), b - a, k, p, parts, _p(pb), _p(ps), _p(pk))
        assert rc == 0, rc
        po = np.zeros(parts + 1, dtype=np.uint64)
        np.cumsum(pb, out=po[1:])
        st = np.zeros(max(int(po[-1]), 1), dtype=np.uint8)
        rc = O.or_store_fill(_p(chars), _p(o), b - a, k, p, parts, _p(po), _p(st))
        assert rc == 0, rc
        return pb, ps, pk, {i: st[int(po[i]):int(po[i + 1])].copy() for i in keep}

    with ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(one, _chunk_bounds(len(offs) - 1, threads * 4)))
    pb = sum(r[0] for r in res); ps = sum(r[1] for r in res); pk = sum(r[2] for r in res)
    kept = {i: np.concatenate([r[3][i] for r in res]) for i in keep}
    return pb, ps, pk, kept


def count_parts_parallel(part_bytes, k, cmin=1, cmax=0xFFFFFFFF, threads=16):
    """{part: (keys ascending, counts, distinct unfiltered, k-mers)} for {part: record bytes}."""
    from concurrent.futures import ThreadPoolExecutor
    items = sorted(part_bytes.items())
    with ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(lambda kv: count_partition(kv[1], k, cmin, cmax), items))
    return {kv[0]: r for kv, r in zip(items, res)}


def records_multiset(st):
    """Sorted records of a partition byte stream: a record ends at a non-full byte that follows a
    full byte (cpu_funcs.cpp:516-525)."""
    if len(st) == 0:
        return []
    full = st >= 0xC0
    end = np.zeros(len(st), dtype=bool)
    end[1:] = (~full[1:]) & full[:-1]
    ends = np.flatnonzero(end) + 1
    starts = np.concatenate([[0], ends[:-1]])
    b = st.tobytes()
    return sorted(b[s:e] for s, e in zip(starts.tolist(), ends.tolist()))


def ref_pipeline_distinct(chars, offs, k, p, parts, threads=16, batch=2000):
    """The reference's own CPU pipeline (GenSuperkmerCPU + KmerCountingCPU, cpu_funcs.cpp built into
    oracle/_ref) on the batch: its distinct k-mer total, or None where oracle/_ref was not built."""
    L = ref(key_width(k))
    if L is None:
        return None
    t1 = np.zeros(1); t2 = np.zeros(1)
    offs = np.ascontiguousarray(offs, dtype=np.uint64)
    return int(L.ref_cpu_pipeline(_p(chars), _p(offs), len(offs) - 1, k, p, parts, threads, batch, _p(t1), _p(t2)))
