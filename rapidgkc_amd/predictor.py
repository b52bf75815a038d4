"""Learned load predictor: sizes the count stage from a 1% read sample.

Inputs (from the sample, computed on the device by phase 1 with the histogram hook and by
counting the sample):
  - 4^8-bin histogram of minimizer prefixes (top 8 bases), k-mer weighted, normalised;
  - per-partition statistics of the sample: superkmer share, k-mer share, distinct/instance ratio;
  - GC fraction, mean read length, read-length coefficient of variation, log10 of the k-mer total.
Outputs, per partition p (2 x parts): log((skm_p + 1) / (N/parts)) and log((distinct_p + 1) / (N/parts)),
where N = total k-mers, known from read lengths before any counting.

The network (2 x 128 hidden, ReLU, BF16 weights, FP32 accumulation) is trained on the device by
librgkc's own forward/backward kernels on counter outputs from seeded genomes (GC 30-65%,
repeats 0-50%, errors 0.1-5%). The predicted distinct counts size the table regions and the count
sweeps (rgkc_set_sizing); an overflow falls back to exact sizing inside the library, and the exact
counts measure the predictor's error.

The reference sizes by instance counts (VRAM grouping gkc_main.cpp:324-352, hash table from N
gpu_hash_counter.cu:407-419) and has no learned component.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import native, synth

HIST_BASES = 8


def sample_csr(chars, offs, frac: float, seed: int):
    """A seeded uniform sample of reads (CSR)."""
    n = len(offs) - 1
    m = max(1, int(round(n * frac)))
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(n, m, replace=False))
    lens = (offs[idx + 1] - offs[idx]).astype(np.int64)
    o = np.zeros(m + 1, dtype=np.uint64)
    np.cumsum(lens, out=o[1:])
    gather = np.repeat(offs[idx].astype(np.int64) - o[:-1].astype(np.int64), lens) + np.arange(int(o[-1]))
    return chars[gather], o


def total_kmers(offs, k):
    lens = np.diff(offs.astype(np.int64))
    return int(np.maximum(lens - k + 1, 0).sum())


def features(counter: native.Counter, chars, offs, k, p, parts, frac=0.01, seed=0):
    sc_, so = sample_csr(chars, offs, frac, seed)
    counter.set_reads(sc_, so)
    hist, skm, kmr, _ = counter.sample_features(k, p, parts)
    srows, _ = counter.count(1, 0xFFFFFFFF)
    ratio = srows.astype(np.float64) / np.maximum(kmr.astype(np.float64), 1.0)
    h = hist.astype(np.float64)
    h = h / max(h.sum(), 1.0) * len(h)
    s = skm.astype(np.float64); s = s / max(s.sum(), 1.0) * parts
    q = kmr.astype(np.float64); q = q / max(q.sum(), 1.0) * parts
    bases = sc_
    gc = float(np.isin(bases, np.frombuffer(b"CGcg", dtype=np.uint8)).mean()) if len(bases) else 0.0
    lens = np.diff(so.astype(np.float64))
    N = total_kmers(offs, k)
    stats = np.array([gc, lens.mean() / 1000.0, lens.std() / max(lens.mean(), 1.0), math.log10(max(N, 1)) / 10.0])
    x = np.concatenate([np.log1p(h), s, q, ratio, stats]).astype(np.float32)
    return x, (skm.astype(np.float64), float(kmr.sum()), q / parts), N


def targets(counter: native.Counter, chars, offs, k, p, parts):
    """Exact per-partition superkmer and distinct counts from the counter itself."""
    counter.set_reads(chars, offs)
    _, skm, _ = counter.superkmers(k, p, parts)
    rows, _ = counter.count(1, 0xFFFFFFFF)
    N = total_kmers(offs, k)
    scale = N / parts
    t = np.concatenate([np.log((skm + 1.0) / scale), np.log((rows + 1.0) / scale)]).astype(np.float32)
    return t, skm.astype(np.float64), rows.astype(np.float64), N


def genome_params(i: int, coverage=None):
    """Seeded training genome i: GC 30-65%, interspersed repeats 0-50%, errors 0.1-5% (log-uniform),
    coverage 10-40x unless fixed."""
    r = np.random.default_rng(1000 + i)
    gp = dict(gc=float(r.uniform(0.30, 0.65)), rep=float(r.uniform(0.0, 0.5)),
              err=float(math.exp(r.uniform(math.log(0.001), math.log(0.05)))))
    cov = float(r.uniform(10.0, 40.0))
    gp["coverage"] = float(coverage) if coverage else cov
    return gp


def example_reads(i, genome_len, read_len=150, coverage=None):
    gp = genome_params(i, coverage)
    g = synth.genome(1000 + i, genome_len, gc=gp["gc"], inter_frac=gp["rep"], rep_div=0.02)
    n_reads = int(gp["coverage"] * genome_len / read_len)
    chars, offs = synth.short_reads(g, 5000 + i, n_reads, read_len, gp["err"])
    return chars, offs, gp


def make_example(counter, i, k=31, p=11, parts=512, genome_len=2_000_000, coverage=None, read_len=150, frac=0.01):
    chars, offs, gp = example_reads(i, genome_len, read_len, coverage)
    x, (sample_skm, sample_kmers, kshare), N = features(counter, chars, offs, k, p, parts, frac, seed=i)
    t, skm, dist, _ = targets(counter, chars, offs, k, p, parts)
    return x, t, dict(N=N, skm=skm, dist=dist, sample_skm=sample_skm, sample_kmers=sample_kmers, kshare=kshare, **gp)


def make_dataset(counter, n_genomes=200, first=0, log=None, **kw):
    X, T, meta = [], [], []
    t0 = time.time()
    for i in range(first, first + n_genomes):
        x, t, m = make_example(counter, i, **kw)
        X.append(x); T.append(t); meta.append(m)
        if log and (i - first) % 20 == 0:
            log(f"[predictor] genome {i - first}/{n_genomes} ({time.time() - t0:.1f}s)")
    return np.stack(X), np.stack(T), meta


class Predictor:
    """The trained network plus its input/output conventions."""

    def __init__(self, counter: native.Counter, n_in: int, parts: int, hidden=(128, 128), seed=0):
        self.parts = parts
        self.net = native.Mlp(counter, n_in, hidden[0], hidden[1], 2 * parts)
        rng = np.random.default_rng(seed)
        dims = (n_in, hidden[0], hidden[1], 2 * parts)
        w = []
        for a, b in [(dims[0], dims[1]), (dims[1], dims[2]), (dims[2], dims[3])]:
            w.append((rng.standard_normal(a * b) * math.sqrt(2.0 / a)).astype(np.float32))
            w.append(np.zeros(b, dtype=np.float32))
        self.net.set_params(np.concatenate(w))
        self.t_mean = None

    def fit(self, X, T, epochs=150, batch=32, lr=1e-3, wd=1e-5, seed=0, log=None):
        self.t_mean = T.mean(0)
        Tn = T - self.t_mean  # learn the deviation from the training mean
        rng = np.random.default_rng(seed)
        n = len(X)
        losses = []
        for ep in range(epochs):
            order = rng.permutation(n)
            el = 0.0
            for b0 in range(0, n, batch):
                idx = order[b0:b0 + batch]
                el += self.net.train_step(X[idx], Tn[idx], lr=lr, wd=wd) * len(idx)
            losses.append(el / n)
            if log and ep % 25 == 0:
                log(f"[predictor] epoch {ep} loss {losses[-1]:.4f}")
        return losses

    def predict_t(self, X):
        return self.net.forward(X) + self.t_mean

    def predict_counts(self, x, N):
        t = self.predict_t(x[None, :])[0]
        scale = N / self.parts
        skm = np.exp(t[: self.parts]) * scale - 1.0
        dist = np.exp(t[self.parts:]) * scale - 1.0
        return np.maximum(skm, 0.0), np.maximum(dist, 0.0)


def naive_skm(meta_one):
    """The method without the network for superkmers: the sample's per-partition counts scaled by
    (total k-mers / sample k-mers). For distinct k-mers the method without the network is the
    exact bound, the partition's instance count."""
    return meta_one["sample_skm"] * (meta_one["N"] / max(meta_one["sample_kmers"], 1.0))


def relative_error(pred, true):
    return float(np.abs(pred - true).sum() / max(true.sum(), 1.0))


def capacities(skm_hat, kshare, N, k, margin=0.15, slack_recs=64):
    """Phase-1 region capacities from predicted superkmers per partition and the sample's k-mer
    shares: records = skm*(1+m)+slack; bytes = (k-mers + records*(k-1))/3 + 1.5*records (a record
    of L bases takes at most L/3 + 1.67 bytes), with the same margin."""
    recs = np.ceil(skm_hat * (1.0 + margin)) + slack_recs
    kmers = np.ceil(kshare * N * (1.0 + margin)) + 4 * slack_recs
    byts = np.ceil((kmers + recs * (k - 1)) / 3.0 + 1.7 * recs) + 256
    return byts.astype(np.uint64), recs.astype(np.uint64)
