"""The no-allocation row sort (rgkc_sort_table_noalloc) against rgkc_sort_rows, the entry point it
is a variant of: the same per-partition ascending table, bit for bit, on two calls in a row (the
second with a smaller table), in a caller-owned device table with gaps between partitions. The
device buffers are the caller's (allocated here with hipMalloc); the entry point allocates nothing."""
import ctypes

import numpy as np
import pytest

from rapidgkc_amd import synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


class DevMem:
    """Caller-owned device memory through the HIP runtime (test harness only)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 16)) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def download(self, ptr, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


CASES = [
    # seed, genome bp, reads, error rate, k, p, parts: the first has every partition above the LDS
    # leaf size (4096 rows), so the radix levels run in more than one round of the fixed areas
    (21, 2_500_000, 200_000, 0.002, 31, 11, 512),
    (22, 300_000, 40_000, 0.01, 55, 13, 64),
    (23, 60_000, 6_000, 0.01, 31, 11, 16),
]


def test_noalloc_sort_matches_sort_rows():
    from rapidgkc_amd import native
    mem = DevMem()
    c = native.Counter(0)
    try:
        prev_rows = None
        for seed, glen, nreads, err, k, p, parts in CASES:
            g = synth.genome(seed, glen, gc=0.45)
            chars, offs = synth.short_reads(g, seed + 1, nreads, 150, err)
            c.set_reads(chars, offs)
            c.superkmers(k, p, parts)
            rows, _ = c.count()
            keys_u, cnts_u, ro = c.table(rows)  # rows as counted (sweep order)
            c.sort_rows()
            keys_s, cnts_s, _ = c.table(rows)   # the existing entry point's order
            total = int(rows.sum())
            say(f"[gpu] noalloc sort: k={k} parts={parts} rows={total}")
            assert total <= (64 << 20) // (12 if k <= 32 else 20)
            if prev_rows is not None:
                assert total < prev_rows, "cases run from larger to smaller tables"
            prev_rows = total
            # the caller's table: partition i at row_base[i], 7 unused rows between partitions
            gap = 7
            row_base = ro[:-1].astype(np.uint64) + np.arange(parts, dtype=np.uint64) * gap
            n_all = total + gap * parts
            kk = np.zeros(n_all, dtype=keys_u.dtype)
            cc = np.full(n_all, 0xDEADBEEF, dtype=np.uint32)
            for i in range(parts):
                b, n = int(row_base[i]), int(rows[i])
                kk[b:b + n] = keys_u[ro[i]:ro[i + 1]]
                cc[b:b + n] = cnts_u[ro[i]:ro[i + 1]]
            dk, dc = mem.upload(kk), mem.upload(cc)
            c.sort_table_noalloc(k, dk, dc, row_base, rows)
            gk, gc = mem.download(dk, kk), mem.download(dc, cc)
            for i in range(parts):
                b, n = int(row_base[i]), int(rows[i])
                assert np.array_equal(gk[b:b + n], keys_s[ro[i]:ro[i + 1]]), f"partition {i}: keys differ"
                assert np.array_equal(gc[b:b + n], cnts_s[ro[i]:ro[i + 1]]), f"partition {i}: counts differ"
                assert np.all(gc[b + n:b + n + gap] == 0xDEADBEEF), f"rows after partition {i} were written"
            mem.free()
        # a table beyond the fixed area is refused before any device work
        big = np.array([(64 << 20) // 12 + 1], dtype=np.uint64)
        with pytest.raises(native.RgkcError) as e:
            c.sort_table_noalloc(31, 0, 0, np.zeros(1, dtype=np.uint64), big)
        assert e.value.code == -2
    finally:
        mem.free()
        c.close()
