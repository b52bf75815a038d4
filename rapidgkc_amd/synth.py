"""Seeded synthetic genomes and read sets for the k-mer counting workloads.

Thin numpy wrapper over ``librgkc_synth.so`` (``csrc/synth.cpp``). Reads come back in CSR form:
``chars`` (uint8, concatenated bases) and ``offs`` (uint64, n_reads + 1), the layout of the
reference's ``PinnedCSR`` (gpu_skmgen.h:29-42).
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "librgkc_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i64, u64, f64, i32 = ctypes.c_int64, ctypes.c_uint64, ctypes.c_double, ctypes.c_int
        L.sg_genome.argtypes = [u64, i64, f64, f64, i32, i32, f64, f64, i32, i32, P]
        L.sg_genome.restype = i64
        L.sg_short_reads.argtypes = [P, i64, u64, i64, i32, f64, P, P, i32]
        L.sg_short_reads.restype = i32
        L.sg_lognormal_lengths.argtypes = [u64, i64, f64, f64, i32, i32, P]
        L.sg_lognormal_lengths.restype = i64
        L.sg_long_reads.argtypes = [P, i64, u64, i64, P, f64, f64, f64, f64, P, P, i32]
        L.sg_long_reads.restype = i32
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def genome(seed: int, length: int, gc: float = 0.5, inter_frac: float = 0.0, rep_min: int = 1000,
           rep_max: int = 5000, rep_div: float = 0.02, tandem_frac: float = 0.0, unit_min: int = 2,
           unit_max: int = 200) -> np.ndarray:
    out = np.empty(length, dtype=np.uint8)
    lib().sg_genome(seed, length, gc, inter_frac, rep_min, rep_max, rep_div, tandem_frac, unit_min, unit_max,
                    _ptr(out))
    return out


def short_reads(g: np.ndarray, seed: int, n_reads: int, read_len: int, sub_rate: float, threads: int = 8):
    chars = np.empty(n_reads * read_len, dtype=np.uint8)
    offs = np.empty(n_reads + 1, dtype=np.uint64)
    rc = lib().sg_short_reads(_ptr(g), len(g), seed, n_reads, read_len, sub_rate, _ptr(chars), _ptr(offs), threads)
    if rc != 0:
        raise ValueError("genome shorter than the read length")
    return chars, offs


def long_reads(g: np.ndarray, seed: int, n_reads: int, mean: float = 10000.0, sd: float = 6000.0,
               err: float = 0.05, split=(0.5, 0.25, 0.25), min_len: int = 100, max_len: int = 100000,
               threads: int = 8):
    """Log-normal lengths with the given mean and sd (natural-log parameters derived from them)."""
    sigma = math.sqrt(math.log(1.0 + (sd / mean) ** 2))
    mu = math.log(mean) - 0.5 * sigma * sigma
    lens = np.empty(n_reads, dtype=np.int32)
    max_len = min(max_len, len(g) // 2)
    lib().sg_lognormal_lengths(seed, n_reads, mu, sigma, min_len, max_len, _ptr(lens))
    offs = np.zeros(n_reads + 1, dtype=np.uint64)
    np.cumsum(lens, out=offs[1:])
    chars = np.empty(int(offs[-1]), dtype=np.uint8)
    lib().sg_long_reads(_ptr(g), len(g), seed, n_reads, _ptr(lens), err, split[0], split[1], split[2],
                        _ptr(chars), _ptr(offs), threads)
    return chars, offs


def concat_csr(parts):
    """Concatenate CSR read sets [(chars, offs), ...] in order."""
    chars = np.concatenate([c for c, _ in parts])
    offs = [np.zeros(1, dtype=np.uint64)]
    base = 0
    for c, o in parts:
        offs.append(o[1:] + np.uint64(base))
        base += int(o[-1])
    return chars, np.concatenate(offs)


def interleave_csr(a, b, seed: int):
    """Interleave two read sets in a seeded random order (uneven mixing of short and long reads)."""
    (ca, oa), (cb, ob) = a, b
    na, nb = len(oa) - 1, len(ob) - 1
    rng = np.random.default_rng(seed)
    order = rng.permutation(na + nb)
    lens = np.concatenate([np.diff(oa), np.diff(ob)]).astype(np.int64)
    starts = np.concatenate([oa[:-1].astype(np.int64), ob[:-1].astype(np.int64) + len(ca)])
    allc = np.concatenate([ca, cb])
    lens_o = lens[order]
    offs = np.zeros(na + nb + 1, dtype=np.uint64)
    np.cumsum(lens_o, out=offs[1:])
    idx = np.repeat(starts[order] - offs[:-1].astype(np.int64), lens_o) + np.arange(int(offs[-1]), dtype=np.int64)
    return allc[idx], offs


@dataclass
class Workload:
    """One named k-mer counting workload (genome + reads + counting parameters)."""
    name: str
    k: int
    p: int
    parts: int
    min_count: int = 1
    max_count: int = 0xFFFFFFFF
    genome_len: int = 1_000_000
    gc: float = 0.5
    inter_frac: float = 0.0
    rep_div: float = 0.02
    tandem_frac: float = 0.0
    n_reads: int = 200_000
    read_len: int = 150
    err: float = 0.005
    seed: int = 1
    long_cov: float = 0.0
    notes: dict = field(default_factory=dict)

    def make_genome(self, scale: float = 1.0) -> np.ndarray:
        return genome(self.seed, max(int(self.genome_len * scale), 10_000), self.gc, self.inter_frac,
                      rep_div=self.rep_div, tandem_frac=self.tandem_frac)

    def make_reads(self, scale: float = 1.0, threads: int = 8):
        g = self.make_genome(scale)
        n = max(int(self.n_reads * scale), 1)
        short = short_reads(g, self.seed + 100, n, self.read_len, self.err, threads)
        if self.long_cov <= 0:
            return short
        n_long = max(1, int(self.long_cov * len(g) / 10000.0))
        lr = long_reads(g, self.seed + 200, n_long, threads=threads)
        return interleave_csr(short, lr, self.seed + 300)


# The five workloads of the k-mer counting hot path.
WORKLOADS = {
    # 1 Mbp uniform genome, 200k x 150 bp, 0.5% substitutions; k=31, p=11; min-count 2.
    "uniform_1m": Workload("uniform_1m", k=31, p=11, parts=512, min_count=2, genome_len=1_000_000, gc=0.5,
                           n_reads=200_000, read_len=150, err=0.005, seed=1),
    # 100 Mbp, GC 41%, 10% as 1-5 kbp copies at 2% divergence; 30x of 150 bp (20M reads), 1% errors.
    "genome_100m_30x": Workload("genome_100m_30x", k=31, p=11, parts=512, genome_len=100_000_000, gc=0.41,
                                inter_frac=0.10, rep_div=0.02, n_reads=20_000_000, read_len=150, err=0.01,
                                seed=2),
    # 50 Mbp; 20x of 150 bp at 0.5% + log-normal long reads (mean 10 kbp, sd 6 kbp, 5% sub/ins/del
    # 50/25/25) at 2x; interleaved; k=25, p=9.
    "mixed_50m": Workload("mixed_50m", k=25, p=9, parts=512, genome_len=50_000_000, gc=0.5, n_reads=6_666_667,
                          read_len=150, err=0.005, seed=3, long_cov=2.0),
    # 200 Mbp, GC 40%, 25% tandem repeats (unit 2-200 bp); 10x of 250 bp, 1% errors; k=55, p=13.
    "tandem_200m_k55": Workload("tandem_200m_k55", k=55, p=13, parts=512, genome_len=200_000_000, gc=0.40,
                                tandem_frac=0.25, n_reads=8_000_000, read_len=250, err=0.01, seed=4),
    # 1 Gbp, 30% repeats; 20x of 150 bp; k=31, p=11 (learned sizing workload).
    "genome_1g_20x": Workload("genome_1g_20x", k=31, p=11, parts=512, genome_len=1_000_000_000, gc=0.41,
                              inter_frac=0.30, n_reads=133_333_333, read_len=150, err=0.01, seed=5),
}
