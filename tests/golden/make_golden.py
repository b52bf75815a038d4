"""Writes tests/golden/golden_small.npz from the reference's own CPU code (cpu_funcs.cpp built into
oracle/_ref by oracle/Makefile). Inputs are a small seeded read set; outputs are the reference's
partition stores (GenSuperkmerCPU) and per-partition (k-mer, count) tables (extract_kmers_cpu +
sort_kmers + the run-length loop of KmerCountingCPU, cpu_funcs.cpp:586-651).

Run in the development container (needs oracle/_ref): python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle_api as oa  # noqa: E402
from rapidgkc_amd import synth  # noqa: E402


def ref_store(R, chars, offs, k, p, parts):
    ptr = ctypes.c_void_p()
    po = np.zeros(parts + 1, dtype=np.uint64); sc = np.zeros(parts, dtype=np.uint64); kc = np.zeros(parts, dtype=np.uint64)
    assert R.ref_gen_superkmers(chars.ctypes.data, offs.ctypes.data, len(offs) - 1, k, p, parts, ctypes.byref(ptr),
                                po.ctypes.data, sc.ctypes.data, kc.ctypes.data) == 0
    st = np.ctypeslib.as_array((ctypes.c_uint8 * int(po[-1])).from_address(ptr.value)).copy()
    R.ref_free(ptr)
    return st, po, sc, kc


def ref_tables(R, st, po, sc, kc, k):
    keys, cnts, ro = [], [], [0]
    for i in range(len(po) - 1):
        seg = np.ascontiguousarray(st[po[i]:po[i + 1]])
        n = int(kc[i])
        if n == 0:
            ro.append(ro[-1]); continue
        rk = np.zeros(n, dtype=oa.key_dtype(k))
        R.ref_extract_sorted(seg.ctypes.data, len(seg), int(sc[i]), n, k, rk.ctypes.data)
        chg = np.ones(n, dtype=bool); chg[1:] = rk[1:] != rk[:-1]
        s = np.flatnonzero(chg)
        keys.append(rk[s]); cnts.append(np.diff(np.append(s, n)).astype(np.uint32)); ro.append(ro[-1] + len(s))
    return np.concatenate(keys), np.concatenate(cnts), np.array(ro, dtype=np.int64)


def main():
    g = synth.genome(101, 8000, gc=0.42, tandem_frac=0.0)
    chars, offs = synth.short_reads(g, 102, 600, 120, 0.01)
    out = {"chars": chars, "offs": offs}
    for tag, (k, p, parts) in {"k31": (31, 11, 64), "k55": (55, 13, 64)}.items():
        R = oa.ref(oa.key_width(k))
        assert R is not None, "oracle/_ref not built"
        st, po, sc, kc = ref_store(R, chars, offs, k, p, parts)
        keys, cnts, ro = ref_tables(R, st, po, sc, kc, k)
        out[f"{tag}_kpp"] = np.array([k, p, parts], dtype=np.int64)
        out[f"{tag}_store"] = st
        out[f"{tag}_part_offs"] = po
        out[f"{tag}_skm_cnt"] = sc
        out[f"{tag}_kmer_cnt"] = kc
        out[f"{tag}_keys"] = keys.view(np.uint64).reshape(-1) if k > 32 else keys
        out[f"{tag}_counts"] = cnts
        out[f"{tag}_row_offs"] = ro
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **out)
    print("wrote", os.path.join(HERE, "golden_small.npz"), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
