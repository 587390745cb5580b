"""Pins the FP64 MLP oracle: analytic gradients against central finite
differences, and the BF16 rounding against torch's bfloat16 conversion."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as MO  # noqa: E402


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.3, MO.N_PARAMS)
    X = rng.uniform(0, 1, (40, MO.N_IN))
    T = rng.normal(0, 1, (40, MO.N_OUT))
    M = (rng.uniform(0, 1, (40, MO.N_OUT)) < 0.7).astype(float)
    loss, g = MO.loss_grad(w, X, T, M)
    assert np.isclose(loss, MO.loss_only(w, X, T, M))
    idx = rng.choice(MO.N_PARAMS, 80, replace=False)
    eps = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        fd = (MO.loss_only(wp, X, T, M) - MO.loss_only(wm, X, T, M)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_matches_torch():
    import torch
    x = np.random.default_rng(1).normal(0, 3, 10000).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(MO.bf16_round(x), want)
