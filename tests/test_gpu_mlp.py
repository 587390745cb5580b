"""Cost model on the GPU: forward and gradients against the FP64 CPU oracle,
SGD training, and the learned scheduler's regret against the fixed C-SAW
heuristic on held-out chunks (what the trained network gains)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "c-saw_amd"))
import mlp_oracle as MO  # noqa: E402


def data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 1, (n, 16)).astype(np.float32)
    X[:, 0] = 1.0
    T = (X[:, 1:6] * np.array([1.0, -2.0, 0.5, 3.0, -1.0], np.float32) + X[:, 6:7]).astype(np.float32)
    M = (rng.uniform(0, 1, (n, 5)) < 0.8).astype(np.float32)
    return X, T, M


def test_forward_matches_oracle(C):
    m = C.Mlp(3)
    X, _, _ = data(500, 1)
    w = MO.bf16_round(m.params())
    want, _ = MO.forward(w, X)
    got = m.forward(X)
    assert np.allclose(got, want, rtol=1e-4, atol=1e-4 * np.abs(want).max())


def test_gradient_matches_oracle(C):
    m = C.Mlp(5)
    X, T, M = data(700, 2)
    loss, g = m.grad(X, T, M)
    wl, wg = MO.loss_grad(MO.bf16_round(m.params()), X, T, M)
    assert abs(loss - wl) <= 1e-4 * abs(wl) + 1e-6
    assert np.allclose(g, wg, rtol=2e-3, atol=2e-4 * np.abs(wg).max())


def test_training_reduces_loss(C):
    m = C.Mlp(7)
    X, T, M = data(4000, 3)
    l0, _ = m.grad(X, T, M)
    l1 = m.train(X, T, M, epochs=30, batch=256, lr=0.05)
    assert l1 < 0.25 * l0, (l0, l1)
    # master weights stay in FP32; the working copy is their BF16 rounding
    p = m.params()
    assert p.dtype == np.float32 and np.isfinite(p).all()


def test_learned_scheduler_beats_heuristic(C):
    import cost_model as CM
    g = C.Graph.rmat(18, 16, weights=True)

    def specs(key, n):
        out = []
        for i, r in enumerate(C.request_mix(n, 64, 8192, key)):
            r = dict(r)
            r["seeds"] = g.seeds(r["samples"], request=i)
            out.append(r)
        return out

    train = CM.collect(C, g, specs(7, 48))
    test = CM.collect(C, g, specs(42, 48))
    mlp, loss = CM.train(C, train)
    r = CM.regret(test, mlp)
    print("held-out regret", r)
    assert r["chunks"] > 100
    assert r["mlp"] < r["heuristic"], r
    # the best variant forced on every chunk (the best fixed policy) is
    # reported beside it (held-out cost of each in "single_variant")
    print("learned / best fixed variant:", r["mlp"] / min(r["single_variant"]))


def test_learned_scheduler_without_cdf_cache(C):
    # the kernels the streaming sampler runs: integer CDFs built per entry
    import cost_model as CM
    g = C.Graph.rmat(18, 16)
    train = CM.collect(C, g, CM.stream_specs(C, g, 48, 7, hi=8192), cdf_cache=False)
    test = CM.collect(C, g, CM.stream_specs(C, g, 48, 42, hi=8192), cdf_cache=False)
    mlp, loss = CM.train(C, train)
    r = CM.regret(test, mlp)
    print("held-out regret (no CDF cache)", r)
    assert r["chunks"] > 100
    assert r["mlp"] < r["heuristic"], r
