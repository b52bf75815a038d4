"""The learned load predictor: its network's CPU oracle is pinned by finite differences, and the
device forward/backward kernels are checked against it (FP32 accumulation vs float64)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402


def tiny_problem(seed=0, dims=(40, 16, 12, 7), B=9):
    rng = np.random.default_rng(seed)
    n = sum(a * b for a, b in [(dims[0], dims[1]), (dims[1], dims[2]), (dims[2], dims[3])]) + dims[1] + dims[2] + dims[3]
    w = (rng.standard_normal(n) * 0.3).astype(np.float32)
    x = rng.standard_normal((B, dims[0])).astype(np.float32)
    t = rng.standard_normal((B, dims[3])).astype(np.float32)
    m = (rng.random((B, dims[3])) > 0.2).astype(np.float32)
    return w, x, t, m


def test_oracle_gradient_finite_differences():
    dims = (40, 16, 12, 7)
    w, x, t, m = tiny_problem(1, dims)
    w = w.astype(np.float64)
    _, g, _ = mo.loss_and_grad(w, dims, x, t, m, round_weights=False)
    rng = np.random.default_rng(2)
    idx = rng.choice(len(w), 60, replace=False)
    eps = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps; wm[i] -= eps
        lp, _, _ = mo.loss_and_grad(wp, dims, x, t, m, round_weights=False)
        lm, _, _ = mo.loss_and_grad(wm, dims, x, t, m, round_weights=False)
        num = (lp - lm) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.14159265], dtype=np.float32)
    r = mo.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie -> even
    assert r[2] == 1.0078125 and r[3] == -2.5 and abs(r[4] - 3.140625) < 1e-9


@pytest.mark.gpu
def test_device_gradients_match_oracle():
    from rapidgkc_amd import native
    print("[gpu] case mlp gradients", flush=True)
    dims = (300, 128, 128, 64)
    w, x, t, m = tiny_problem(3, dims, B=33)
    c = native.Counter(0)
    net = native.Mlp(c, *dims)
    assert net.n_params == len(w)
    net.set_params(w)
    assert np.array_equal(net.get_params(), w)
    y = net.forward(x)
    g, loss = net.grad(x, t, m)
    ol, og, oy = mo.loss_and_grad(w, dims, x, t, m)
    # FP32 accumulation of BF16-weight products vs float64: relative 1e-4 on outputs, 1e-3 on
    # gradients (scaled by the gradient's magnitude)
    assert np.allclose(y, oy, rtol=1e-4, atol=1e-4 * np.abs(oy).max())
    assert abs(loss - ol) <= 1e-4 * abs(ol)
    assert np.abs(g - og).max() <= 1e-3 * np.abs(og).max()
    # a few Adam steps reduce the loss
    l0 = net.train_step(x, t, m, lr=1e-3)
    for _ in range(30):
        l1 = net.train_step(x, t, m, lr=1e-3)
    assert l1 < 0.5 * l0
    net.close()
    c.close()


def oa_sort(keys, cnts):
    o = np.argsort(keys, kind="stable")
    return keys[o], cnts[o]


@pytest.mark.gpu
def test_predictor_trains_and_sizes_the_count():
    """Train on counter outputs of seeded genomes (GC 30-65%, repeats 0-50%, errors 0.1-5%), then on
    held-out genomes: (1) counting with predicted sizing gives the same table as exact sizing;
    (2) the table allocation shrinks (the gain over sizing by instance counts); (3) forcing an
    under-prediction triggers the exact fallback and still gives the same table."""
    from rapidgkc_amd import native, predictor as pr
    print("[gpu] case predictor", flush=True)
    k, p, parts = 31, 11, 64
    c = native.Counter(0)
    X, T, meta = pr.make_dataset(c, n_genomes=48, k=k, p=p, parts=parts, genome_len=600_000)
    tr, cal = slice(0, 32), slice(32, 40)
    model = pr.Predictor(c, X.shape[1], parts, seed=1)
    losses = model.fit(X[tr], T[tr], epochs=120, batch=16, lr=1e-3)
    assert losses[-1] < 0.5 * losses[0]
    m_skm, m_dist = model.calibrate(X[cal], T[cal])
    # the distinct-k-mer head must beat the instance bound on held-out genomes (that is its gain)
    assert model.use_net[1], model.val_err
    gains, errs_net, errs_naive = [], [], []
    for j in range(40, 48):
        m = meta[j]
        skm_hat, dist_hat = model.predict_meta(X[j], m)
        errs_net.append(pr.relative_error(skm_hat, m["skm"]))
        errs_naive.append(pr.relative_error(pr.naive_skm(m), m["skm"]))
        chars, offs, _ = pr.example_reads(j, 600_000)
        c.set_reads(chars, offs)
        c.superkmers(k, p, parts)
        c.set_sizing(None)
        rows_e, d_e = c.count(1, 0xFFFFFFFF)
        ke, ce, _ = c.table(rows_e)
        alloc_exact = c.stats()["table_rows_alloc"]
        c.set_sizing(dist_hat, margin=m_dist)
        rows_p, d_p = c.count(1, 0xFFFFFFFF)
        st = c.stats()
        kp, cp, _ = c.table(rows_p)
        assert np.array_equal(rows_e, rows_p) and d_e == d_p
        assert np.array_equal(np.sort(ke), np.sort(kp))
        gains.append(alloc_exact / st["table_rows_alloc"])
        # forced under-prediction: the partitions whose regions overflow are counted again exactly
        # (the others keep their rows), and the table is the same
        c.set_sizing(dist_hat * 0.3, margin=0.0)
        rows_f, d_f = c.count(1, 0xFFFFFFFF)
        n_fb = c.stats()["fallbacks"]
        assert 1 <= n_fb <= parts and np.array_equal(rows_f, rows_e) and d_f == d_e
        kf, cf, rof = c.table(rows_f)
        for i in range(parts):
            a_ = oa_sort(kf[rof[i]:rof[i + 1]], cf[rof[i]:rof[i + 1]])
            b_ = oa_sort(ke[rof[i]:rof[i + 1]], ce[rof[i]:rof[i + 1]])
            assert np.array_equal(a_[0], b_[0]) and np.array_equal(a_[1], b_[1]), f"partition {i} differs after fallback"
        c.set_sizing(None)
    print(f"[predictor] table allocation gain x{np.mean(gains):.2f} (margin {m_dist:.2f}); superkmer error "
          f"{'net' if model.use_net[0] else 'sample scaling'} {np.mean(errs_net):.3f} vs sample scaling "
          f"{np.mean(errs_naive):.3f}; held-out |log err| net/baseline {model.val_err}", flush=True)
    assert np.mean(gains) > 1.5
    model.net.close()
    c.close()
