"""Learned load predictor: sizes the count stage from a 1% read sample.

Inputs (from the sample, computed on the device by phase 1 with the histogram hook and by
counting the sample), standardised with the training set's per-feature mean and spread:
  - 4^5-bin histogram of minimizer prefixes (top 5 bases), k-mer weighted, normalised;
  - per-partition statistics of the sample: superkmer share, k-mer share, distinct/instance ratio;
  - GC fraction, mean read length, read-length coefficient of variation (nothing that scales with
    the batch size, so genomes of 1 Mbp train a model for batches of 100 Mbp).
Outputs, per partition p (2 x parts): log((skm_p + 1) / (skm_base_p + 1)) and
log((distinct_p + 1) / (inst_base_p + 1)): corrections to the sample-scaled superkmer count and to the
instance count implied by the sample's k-mer share (N = total k-mers, known from read lengths before
any counting). A network that outputs its training mean reproduces those baselines.

The network (2 x 128 hidden, ReLU, BF16 weights, FP32 accumulation) is trained on the device by
librgkc's own forward/backward kernels on counter outputs from seeded genomes (GC 30-65%,
repeats 0-50%, errors 0.1-5%). The predicted distinct counts size the table regions
(rgkc_set_sizing); a partition whose region overflows is counted again with exact sizing inside the
library, and the exact counts measure the predictor's error.

The reference sizes by instance counts (VRAM grouping gkc_main.cpp:324-352, hash table from N
gpu_hash_counter.cu:407-419) and has no learned component.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import native, synth

HIST_BASES = 8
HIST_BINS = 4 ** 5  # network input: the histogram coarsened to its top 5 bases


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
    h = h.reshape(HIST_BINS, -1).sum(1) if len(h) % HIST_BINS == 0 else h  # coarsen to the top 5 bases
    h = h / max(h.sum(), 1.0) * len(h)
    s = skm.astype(np.float64); s = s / max(s.sum(), 1.0) * parts
    q = kmr.astype(np.float64); q = q / max(q.sum(), 1.0) * parts
    bases = sc_
    gc = float(np.isin(bases, np.frombuffer(b"CGcg", dtype=np.uint8)).mean()) if len(bases) else 0.0
    lens = np.diff(so.astype(np.float64))
    N = total_kmers(offs, k)
    # batch-size-free statistics only: the network is trained on small genomes and applied to large batches
    stats = np.array([gc, lens.mean() / 1000.0, lens.std() / max(lens.mean(), 1.0)])
    x = np.concatenate([np.log1p(h), s, q, ratio, stats]).astype(np.float32)
    return x, (skm.astype(np.float64), float(kmr.sum()), q / parts), N


def baselines(N, sample_skm, sample_kmers, kshare):
    """The estimates without the network: superkmers per partition scaled from the sample (its
    counts times total k-mers / sample k-mers), and k-mer instances per partition from the sample's
    k-mer shares (the exact-sizing bound on distinct k-mers)."""
    return np.asarray(sample_skm, np.float64) * (N / max(sample_kmers, 1.0)), np.asarray(kshare, np.float64) * N


def targets(counter: native.Counter, chars, offs, k, p, parts, base):
    """Exact per-partition superkmer and distinct counts from the counter itself, as log ratios to
    the baselines: log((skm + 1) / (skm_base + 1)) and log((distinct + 1) / (instances_base + 1))."""
    counter.set_reads(chars, offs)
    _, skm, _ = counter.superkmers(k, p, parts)
    rows, _ = counter.count(1, 0xFFFFFFFF)
    N = total_kmers(offs, k)
    b_skm, b_inst = base
    t = np.concatenate([np.log((skm + 1.0) / (b_skm + 1.0)), np.log((rows + 1.0) / (b_inst + 1.0))]).astype(np.float32)
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
    t, skm, dist, _ = targets(counter, chars, offs, k, p, parts, baselines(N, sample_skm, sample_kmers, kshare))
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
        self.x_mean = None
        self.x_scale = None
        self.margin_skm = 0.25
        self.margin_dist = 0.25
        self.use_net = [True, True]      # per head: network (True) or baseline (False)
        self.val_err = [None, None]      # per head: (network, baseline) mean |log error| on held-out data

    def _xn(self, X):
        # standardised features, clipped: a feature that barely varies over the training genomes
        # must not blow up on an input where it does
        return np.clip((X - self.x_mean) / self.x_scale, -8.0, 8.0).astype(np.float32)

    def fit(self, X, T, epochs=150, batch=32, lr=3e-4, wd=1e-5, seed=0, log=None):
        self.t_mean = T.mean(0)
        self.x_mean = X.mean(0)
        self.x_scale = np.maximum(X.std(0), 0.05 * np.abs(self.x_mean) + 1e-2)
        Xn = self._xn(X)
        Tn = T - self.t_mean  # learn the deviation from the training mean
        rng = np.random.default_rng(seed)
        n = len(X)
        losses = []
        for ep in range(epochs):
            order = rng.permutation(n)
            el = 0.0
            for b0 in range(0, n, batch):
                idx = order[b0:b0 + batch]
                el += self.net.train_step(Xn[idx], Tn[idx], lr=lr, wd=wd) * len(idx)
            losses.append(el / n)
            if log and ep % 25 == 0:
                log(f"[predictor] epoch {ep} loss {losses[-1]:.4f}")
        return losses

    def calibrate(self, X, T, q=0.9, q_part=0.98):
        """Model selection and sizing margins on held-out genomes. For each head (superkmers,
        distinct k-mers) the network is kept only if its mean absolute log error beats the baseline
        (a zero correction). Superkmers size phase 1, where one partition over its region sends the
        whole batch to exact sizing, and errors are correlated within a genome: that margin is the
        q-quantile, over held-out genomes, of the genome's largest log(true / predicted). A table
        region that is too small costs only its own partition's recount, so the distinct margin is
        the q_part-quantile of log(true / predicted) over all held-out (genome, partition) pairs."""
        P = self.net.forward(self._xn(X)) + self.t_mean
        p = self.parts
        out = []
        for h, sl in enumerate((slice(0, p), slice(p, 2 * p))):
            r_net, r_base = T[:, sl] - P[:, sl], T[:, sl]
            use_net = float(np.abs(r_net).mean()) < float(np.abs(r_base).mean())
            self.use_net[h] = use_net
            r = r_net if use_net else r_base
            worst = np.quantile(r.max(axis=1), q) if h == 0 else np.quantile(r.ravel(), q_part)
            out.append(float(min(max(np.exp(worst) - 1.0, 0.05), 3.0)))
            self.val_err[h] = (float(np.abs(r_net).mean()), float(np.abs(r_base).mean()))
        self.margin_skm, self.margin_dist = out
        return self.margin_skm, self.margin_dist

    def predict_t(self, X):
        return self.net.forward(self._xn(X)) + self.t_mean

    def predict_counts(self, x, N, sample_skm, sample_kmers, kshare):
        """Per-partition superkmers and distinct k-mers for a batch of N k-mers whose 1% sample gave
        features x and the sample counts (see features())."""
        t = self.predict_t(x[None, :])[0]
        b_skm, b_inst = baselines(N, sample_skm, sample_kmers, kshare)
        ts = t[: self.parts] if self.use_net[0] else 0.0
        td = t[self.parts:] if self.use_net[1] else 0.0
        skm = np.exp(np.clip(ts, -3.0, 3.0)) * (b_skm + 1.0) - 1.0
        dist = np.exp(np.clip(td, -6.0, 0.5)) * (b_inst + 1.0) - 1.0
        # a partition has at most as many distinct k-mers as k-mer instances
        return np.maximum(skm, 0.0), np.clip(dist, 0.0, b_inst)

    def predict_meta(self, x, m):
        return self.predict_counts(x, m["N"], m["sample_skm"], m["sample_kmers"], m["kshare"])


def naive_skm(meta_one):
    """The method without the network for superkmers: the sample's per-partition counts scaled by
    (total k-mers / sample k-mers). For distinct k-mers the method without the network is the
    exact bound, the partition's instance count."""
    return baselines(meta_one["N"], meta_one["sample_skm"], meta_one["sample_kmers"], meta_one["kshare"])[0]


def relative_error(pred, true):
    return float(np.abs(pred - true).sum() / max(true.sum(), 1.0))


def sample_margin(sample_skm, z=6.0, floor=0.05):
    """Per-partition margin for superkmer counts scaled from a read sample: a partition's count in
    the sample is close to Poisson, so its relative standard deviation is 1/sqrt(count); the margin
    is z of them plus a floor (a 1% sample of 20 M reads leaves ~700 superkmers per partition of
    512: 28% margin; of 133 M reads ~4500: 14%)."""
    return floor + z / np.sqrt(np.asarray(sample_skm, dtype=np.float64) + 1.0)


def capacities(skm_hat, kshare, N, k, margin=0.25, slack_recs=64):
    """Phase-1 region capacities from predicted superkmers per partition and the sample's k-mer
    shares: records = skm*(1+m)+slack; bytes = (k-mers + records*(k-1))/3 + 1.5*records (a record
    of L bases takes at most L/3 + 1.67 bytes), with the same margin."""
    recs = np.ceil(skm_hat * (1.0 + margin)) + slack_recs
    kmers = np.ceil(kshare * N * (1.0 + margin)) + 4 * slack_recs
    byts = np.ceil((kmers + recs * (k - 1)) / 3.0 + 1.7 * recs) + 256
    return byts.astype(np.uint64), recs.astype(np.uint64)
