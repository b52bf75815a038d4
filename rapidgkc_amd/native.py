"""ctypes binding of the C ABI (include/rgkc.h) exported by librgkc.so, the gfx950 path.

This is the host-side mirror of the reference's GPU interface for the counting path:
``superkmers`` is GenSuperkmerGPU (gpu_skmgen.h:54-58), ``count`` is kmc_counting_GPU_streams
(gpu_kmercounting.h:14-17). There is no CPU fallback: without the library or a device every call
raises.
"""
from __future__ import annotations

import atexit
import ctypes
import sys
import weakref
import os

import numpy as np


_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

P = ctypes.c_void_p
i32, i64, u32, u64 = ctypes.c_int, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("pack_ms", "segments_ms", "p1_count_ms", "p1_emit_ms", "index_ms", "plan_ms", "count_ms", "total_ms")] + \
               [(n, ctypes.c_uint64) for n in
                ("n_reads", "n_bases", "n_kmers", "n_records", "store_bytes", "distinct", "rows", "sweeps",
                 "sweep_retries", "buckets", "items", "table_rows_alloc")] + \
               [("fallbacks", ctypes.c_uint32), ("err_bits", ctypes.c_uint32), ("p1_rescans", ctypes.c_uint32),
                ("heavy_groups", ctypes.c_uint32), ("expand_ms", ctypes.c_float), ("p1_list_order", ctypes.c_uint32),
                ("ovf_keys", ctypes.c_uint64), ("count_groups", ctypes.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


EXPORTS = ["rgkc_device_reset", "rgkc_device_count", "rgkc_create", "rgkc_destroy", "rgkc_last_error",
           "rgkc_set_reads", "rgkc_superkmers", "rgkc_copy_store", "rgkc_count", "rgkc_count_host_store",
           "rgkc_copy_table", "rgkc_sort_rows", "rgkc_run", "rgkc_get_stats", "rgkc_set_tuning", "rgkc_set_sizing",
           "rgkc_sample_features", "rgkc_mlp_create", "rgkc_mlp_destroy", "rgkc_mlp_num_params", "rgkc_mlp_set_params",
           "rgkc_mlp_get_params", "rgkc_mlp_forward", "rgkc_mlp_grad", "rgkc_mlp_train_step", "rgkc_stream_begin", "rgkc_stream_add", "rgkc_stream_end", "rgkc_superkmers_sized", "rgkc_set_p1_lists", "rgkc_set_heavy_expand",
           "rgkc_sort_table_noalloc", "rgkc_set_memory_limit", "rgkc_store_minimizer_length",
           "rgkc_set_count_workgroups"]


_LIVE = weakref.WeakSet()


@atexit.register
def _close_live():
    """Destroy the handles still open at interpreter exit while the HIP runtime is alive (networks
    first, they belong to a counter's device)."""
    objs = list(_LIVE)
    for o in sorted(objs, key=lambda o: 0 if isinstance(o, Mlp) else 1):
        try:
            o.close()
        except Exception:
            pass


def lib_path():
    return os.path.join(_HERE, "librgkc.so")


def lib():
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        L.rgkc_device_reset.argtypes = [i32]; L.rgkc_device_reset.restype = ctypes.c_size_t
        L.rgkc_device_count.argtypes = []; L.rgkc_device_count.restype = i32
        L.rgkc_create.argtypes = [i32, ctypes.POINTER(P)]; L.rgkc_create.restype = i32
        L.rgkc_destroy.argtypes = [P]; L.rgkc_destroy.restype = i32
        L.rgkc_last_error.argtypes = [P]; L.rgkc_last_error.restype = ctypes.c_char_p
        L.rgkc_set_reads.argtypes = [P, P, P, i64]; L.rgkc_set_reads.restype = i32
        L.rgkc_superkmers.argtypes = [P, i32, i32, i32, P, P, P]; L.rgkc_superkmers.restype = i32
        L.rgkc_copy_store.argtypes = [P, P]; L.rgkc_copy_store.restype = i32
        L.rgkc_count.argtypes = [P, u32, u32, P, P]; L.rgkc_count.restype = i32
        L.rgkc_count_host_store.argtypes = [P, P, P, i32, P, i32, i32, i32, u32, u32, P, P]
        L.rgkc_count_host_store.restype = i32
        L.rgkc_copy_table.argtypes = [P, P, P]; L.rgkc_copy_table.restype = i32
        L.rgkc_sort_rows.argtypes = [P]; L.rgkc_sort_rows.restype = i32
        L.rgkc_sort_table_noalloc.argtypes = [P, i32, P, P, P, P, i32]; L.rgkc_sort_table_noalloc.restype = i32
        L.rgkc_run.argtypes = [P, i32, i32, i32, u32, u32]; L.rgkc_run.restype = i32
        L.rgkc_get_stats.argtypes = [P, ctypes.POINTER(Stats)]; L.rgkc_get_stats.restype = i32
        L.rgkc_set_tuning.argtypes = [P, u32]; L.rgkc_set_tuning.restype = i32
        L.rgkc_set_memory_limit.argtypes = [P, u64]; L.rgkc_set_memory_limit.restype = i32
        L.rgkc_set_count_workgroups.argtypes = [P, u32]; L.rgkc_set_count_workgroups.restype = i32
        L.rgkc_store_minimizer_length.argtypes = [P, P, i32, i32, P]; L.rgkc_store_minimizer_length.restype = i32
        L.rgkc_set_p1_lists.argtypes = [P, u32]; L.rgkc_set_p1_lists.restype = i32
        L.rgkc_set_heavy_expand.argtypes = [P, i32]; L.rgkc_set_heavy_expand.restype = i32
        L.rgkc_set_sizing.argtypes = [P, P, i32, ctypes.c_float]; L.rgkc_set_sizing.restype = i32
        L.rgkc_superkmers_sized.argtypes = [P, i32, i32, i32, P, P, P, P, P]; L.rgkc_superkmers_sized.restype = i32
        L.rgkc_stream_begin.argtypes = [P, i32, i32, i32, u64, u64, P, P, i32]; L.rgkc_stream_begin.restype = i32
        L.rgkc_stream_add.argtypes = [P, P, P, i64]; L.rgkc_stream_add.restype = i32
        L.rgkc_stream_end.argtypes = [P, P, P, P]; L.rgkc_stream_end.restype = i32
        L.rgkc_sample_features.argtypes = [P, i32, i32, i32, P, P, P, P]; L.rgkc_sample_features.restype = i32
        L.rgkc_mlp_create.argtypes = [P, i32, i32, i32, i32, ctypes.POINTER(P)]; L.rgkc_mlp_create.restype = i32
        L.rgkc_mlp_destroy.argtypes = [P]; L.rgkc_mlp_destroy.restype = i32
        L.rgkc_mlp_num_params.argtypes = [P, P]; L.rgkc_mlp_num_params.restype = i32
        L.rgkc_mlp_set_params.argtypes = [P, P]; L.rgkc_mlp_set_params.restype = i32
        L.rgkc_mlp_get_params.argtypes = [P, P]; L.rgkc_mlp_get_params.restype = i32
        L.rgkc_mlp_forward.argtypes = [P, P, i32, P]; L.rgkc_mlp_forward.restype = i32
        L.rgkc_mlp_grad.argtypes = [P, P, P, P, i32, P, P]; L.rgkc_mlp_grad.restype = i32
        L.rgkc_mlp_train_step.argtypes = [P, P, P, P, i32, ctypes.c_float, ctypes.c_float, P]
        L.rgkc_mlp_train_step.restype = i32
        _LIB = L
    return _LIB


def key_dtype(k):
    return np.uint64 if k <= 32 else np.dtype([("lo", "<u8"), ("hi", "<u8")])


def store_minimizer_length(store: np.ndarray, part_offs: np.ndarray, k: int) -> int:
    """The minimizer length a store was built with, recovered from its records (host only)."""
    st = np.ascontiguousarray(store, dtype=np.uint8)
    po = np.ascontiguousarray(part_offs, dtype=np.uint64)
    out = np.zeros(1, dtype=np.int32)
    rc = lib().rgkc_store_minimizer_length(st.ctypes.data if len(st) else None, po.ctypes.data, len(po) - 1, k,
                                           out.ctypes.data)
    if rc != 0:
        raise RgkcError(rc, "the minimizer length could not be recovered from the records")
    return int(out[0])


class RgkcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rgkc error {code}: {msg}")
        self.code = code


class Counter:
    """One device context: a resident read batch, its superkmer store and its (k-mer, count) table."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = P()
        rc = self._L.rgkc_create(device, ctypes.byref(h))
        if rc != 0:
            raise RgkcError(rc, "cannot create a context (no HIP device?)")
        self._h = h
        _LIVE.add(self)
        self.k = self.parts = None

    def close(self):
        if getattr(self, "_h", None):
            self._L.rgkc_destroy(self._h)
            self._h = None

    def __del__(self, _finalizing=sys.is_finalizing):  # bound early: module globals may be gone at exit
        if _finalizing():
            return  # the atexit hook has closed it, or the runtime is going away
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise RgkcError(rc, self._L.rgkc_last_error(self._h).decode())

    def set_reads(self, chars: np.ndarray, offs: np.ndarray):
        chars = np.ascontiguousarray(chars, dtype=np.uint8)
        offs = np.ascontiguousarray(offs, dtype=np.uint64)
        self._chk(self._L.rgkc_set_reads(self._h, chars.ctypes.data, offs.ctypes.data, len(offs) - 1))

    def superkmers(self, k: int, p: int, parts: int):
        po = np.zeros(parts + 1, dtype=np.uint64)
        sc = np.zeros(parts, dtype=np.uint64)
        kc = np.zeros(parts, dtype=np.uint64)
        self._chk(self._L.rgkc_superkmers(self._h, k, p, parts, po.ctypes.data, sc.ctypes.data, kc.ctypes.data))
        self.k, self.parts = k, parts
        return po, sc, kc

    def copy_store(self, total_bytes: int):
        out = np.zeros(max(int(total_bytes), 1), dtype=np.uint8)
        self._chk(self._L.rgkc_copy_store(self._h, out.ctypes.data))
        return out[: int(total_bytes)]

    def count(self, cmin: int = 1, cmax: int = 0xFFFFFFFF):
        rows = np.zeros(self.parts, dtype=np.uint64)
        d = np.zeros(1, dtype=np.uint64)
        self._chk(self._L.rgkc_count(self._h, cmin, cmax, rows.ctypes.data, d.ctypes.data))
        return rows, int(d[0])

    def count_host_store(self, store: np.ndarray, part_offs: np.ndarray, parts_total: int, k: int, p: int,
                         cmin: int = 1, cmax: int = 0xFFFFFFFF, part_ids=None):
        """Count partitions given as reference-format superkmer bytes (partition-major)."""
        store = np.ascontiguousarray(store, dtype=np.uint8)
        part_offs = np.ascontiguousarray(part_offs, dtype=np.uint64)
        n = len(part_offs) - 1
        ids = np.ascontiguousarray(np.arange(n) if part_ids is None else part_ids, dtype=np.int32)
        rows = np.zeros(n, dtype=np.uint64)
        d = np.zeros(1, dtype=np.uint64)
        self._chk(self._L.rgkc_count_host_store(self._h, store.ctypes.data if len(store) else None, part_offs.ctypes.data,
                                                n, ids.ctypes.data, parts_total, k, p, cmin, cmax, rows.ctypes.data,
                                                d.ctypes.data))
        self.k, self.parts = k, n
        return rows, int(d[0])

    def table(self, rows: np.ndarray):
        n = int(rows.sum())
        keys = np.zeros(max(n, 1), dtype=key_dtype(self.k))
        cnts = np.zeros(max(n, 1), dtype=np.uint32)
        self._chk(self._L.rgkc_copy_table(self._h, keys.ctypes.data, cnts.ctypes.data))
        ro = np.zeros(len(rows) + 1, dtype=np.int64)
        np.cumsum(rows, out=ro[1:])
        return keys[:n], cnts[:n], ro

    def sort_rows(self):
        """Order each partition's table rows by ascending key on the device (the reference's table
        order, cpu_funcs.cpp:614-640)."""
        self._chk(self._L.rgkc_sort_rows(self._h))

    def sort_table_noalloc(self, k: int, keys_dev: int, counts_dev: int, row_base, row_counts):
        """Order the rows of a caller-owned device table (device addresses keys_dev, counts_dev) by
        ascending key within each partition, without allocating device memory."""
        rb = np.ascontiguousarray(row_base, dtype=np.uint64)
        rc = np.ascontiguousarray(row_counts, dtype=np.uint64)
        self._chk(self._L.rgkc_sort_table_noalloc(self._h, k, keys_dev, counts_dev, rb.ctypes.data, rc.ctypes.data,
                                                  len(rb)))

    def run(self, k: int, p: int, parts: int, cmin: int = 1, cmax: int = 0xFFFFFFFF):
        self.k, self.parts = k, parts
        self._chk(self._L.rgkc_run(self._h, k, p, parts, cmin, cmax))

    def stats(self) -> dict:
        s = Stats()
        self._chk(self._L.rgkc_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def set_tuning(self, inst_per_sweep: int = 0):
        self._chk(self._L.rgkc_set_tuning(self._h, inst_per_sweep))

    def set_count_workgroups(self, per_cu: int = 0):
        """Resident workgroups per CU of the persistent count kernel (0: four)."""
        self._chk(self._L.rgkc_set_count_workgroups(self._h, per_cu))

    def set_memory_limit(self, nbytes: int = 0):
        """Device memory the count stage plans with (0: what is free); a small limit makes it count
        partition groups one after another, compacting rows when instance-sized regions do not fit."""
        self._chk(self._L.rgkc_set_memory_limit(self._h, int(nbytes)))

    def set_heavy_expand(self, mode: int = 0):
        """0: heavy sub-buckets are expanded into keys once (when memory allows), 1: never."""
        self._chk(self._L.rgkc_set_heavy_expand(self._h, mode))

    def set_p1_lists(self, list_cap: int = 0):
        """Superkmer list entries kept per phase-1 workgroup (0 default, 0xFFFFFFFF: no lists)."""
        self._chk(self._L.rgkc_set_p1_lists(self._h, list_cap))

    def superkmers_sized(self, k, p, parts, cap_bytes, cap_recs):
        """Phase 1 into capacity-sized regions; raises RgkcError(-4) if a capacity is too small."""
        cb = np.ascontiguousarray(cap_bytes, dtype=np.uint64)
        cr = np.ascontiguousarray(cap_recs, dtype=np.uint64)
        pb = np.zeros(parts, dtype=np.uint64); sc = np.zeros(parts, dtype=np.uint64); kc = np.zeros(parts, dtype=np.uint64)
        self.k, self.parts = k, parts
        self._chk(self._L.rgkc_superkmers_sized(self._h, k, p, parts, cb.ctypes.data, cr.ctypes.data, pb.ctypes.data,
                                                sc.ctypes.data, kc.ctypes.data))
        return pb, sc, kc

    def stream_begin(self, k, p, parts, max_reads, max_bases, cap_bytes, cap_recs, use_graph=True):
        """Batched input: regions of the given capacities; batches appended by stream_add."""
        cb = np.ascontiguousarray(cap_bytes, dtype=np.uint64)
        cr = np.ascontiguousarray(cap_recs, dtype=np.uint64)
        self._chk(self._L.rgkc_stream_begin(self._h, k, p, parts, int(max_reads), int(max_bases), cb.ctypes.data,
                                            cr.ctypes.data, 1 if use_graph else 0))
        self.k, self.parts = k, parts

    def stream_add(self, chars, offs):
        chars = np.ascontiguousarray(chars, dtype=np.uint8)
        offs = np.ascontiguousarray(offs, dtype=np.uint64)
        self._chk(self._L.rgkc_stream_add(self._h, chars.ctypes.data, offs.ctypes.data, len(offs) - 1))

    def stream_end(self):
        """Returns (bytes, superkmers, k-mers) per partition; raises RgkcError(-4) on overflow."""
        pb = np.zeros(self.parts, dtype=np.uint64); sc = np.zeros(self.parts, dtype=np.uint64)
        kc = np.zeros(self.parts, dtype=np.uint64)
        self._chk(self._L.rgkc_stream_end(self._h, pb.ctypes.data, sc.ctypes.data, kc.ctypes.data))
        return pb, sc, kc

    def set_sizing(self, distinct_pred=None, margin: float = 0.1):
        """Size the next count from predicted distinct k-mers per partition (None: exact sizing)."""
        if distinct_pred is None:
            self._chk(self._L.rgkc_set_sizing(self._h, None, 0, 0.0))
            return
        d = np.ascontiguousarray(distinct_pred, dtype=np.float64)
        self._chk(self._L.rgkc_set_sizing(self._h, d.ctypes.data, len(d), margin))

    def sample_features(self, k: int, p: int, parts: int):
        """Phase 1 on the resident batch (a read sample): minimizer-prefix histogram and
        per-partition superkmer / k-mer / byte counts."""
        nb = 1 << (2 * min(p, 8))
        hist = np.zeros(nb, dtype=np.uint32)
        sc = np.zeros(parts, dtype=np.uint64); kc = np.zeros(parts, dtype=np.uint64); bc = np.zeros(parts, dtype=np.uint64)
        self._chk(self._L.rgkc_sample_features(self._h, k, p, parts, hist.ctypes.data, sc.ctypes.data, kc.ctypes.data,
                                               bc.ctypes.data))
        self.k, self.parts = k, parts
        return hist, sc, kc, bc


class Mlp:
    """The load predictor's network on the device: y = W3 relu(W2 relu(W1 x + b1) + b2) + b3 with BF16
    weights (FP32 master copy), FP32 accumulation, masked MSE, Adam."""

    def __init__(self, counter: Counter, n_in: int, h1: int, h2: int, n_out: int):
        self._c = counter
        self._L = counter._L
        h = P()
        self._L.rgkc_mlp_create(counter._h, n_in, h1, h2, n_out, ctypes.byref(h)) == 0 or self._err()
        self._h = h
        _LIVE.add(self)
        self.dims = (n_in, h1, h2, n_out)
        n = np.zeros(1, dtype=np.uint64)
        self._L.rgkc_mlp_num_params(self._h, n.ctypes.data)
        self.n_params = int(n[0])

    def _err(self):
        raise RgkcError(-1, self._L.rgkc_last_error(self._c._h).decode())

    def _chk(self, rc):
        if rc != 0:
            raise RgkcError(rc, self._L.rgkc_last_error(self._c._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.rgkc_mlp_destroy(self._h)
            self._h = None

    def __del__(self, _finalizing=sys.is_finalizing):  # bound early: module globals may be gone at exit
        if _finalizing():
            return  # the atexit hook has closed it, or the runtime is going away
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.n_params
        self._chk(self._L.rgkc_mlp_set_params(self._h, w.ctypes.data))

    def get_params(self):
        w = np.zeros(self.n_params, dtype=np.float32)
        self._chk(self._L.rgkc_mlp_get_params(self._h, w.ctypes.data))
        return w

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.dims[3]), dtype=np.float32)
        self._chk(self._L.rgkc_mlp_forward(self._h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def grad(self, x, t, mask=None):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._chk(self._L.rgkc_mlp_grad(self._h, x.ctypes.data, t.ctypes.data, None if m is None else m.ctypes.data,
                                        x.shape[0], g.ctypes.data, loss.ctypes.data))
        return g, float(loss[0])

    def train_step(self, x, t, mask=None, lr=1e-3, wd=0.0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._chk(self._L.rgkc_mlp_train_step(self._h, x.ctypes.data, t.ctypes.data,
                                              None if m is None else m.ctypes.data, x.shape[0], lr, wd,
                                              loss.ctypes.data))
        return float(loss[0])
