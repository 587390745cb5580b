"""FP64 CPU oracle of the cost-model MLP (test infrastructure only).

Same layout as c-saw_amd/csrc/mlp.hpp: parameters [W1 (64x16), b1 (64),
W2 (64x64), b2 (64), W3 (5x64), b3 (5)]; the network evaluates the
BF16-rounded parameters; loss = sum M * (Y - T)^2 / 2 / sum M.
Pinned by central finite differences in tests/test_oracle_mlp.py.
"""
import numpy as np

N_IN, N_HID, N_OUT = 16, 64, 5
SHAPES = [("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
          ("W3", (N_OUT, N_HID)), ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16_round(p):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(p, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def unpack(w):
    out, at = {}, 0
    for name, s in SHAPES:
        n = int(np.prod(s))
        out[name] = w[at:at + n].reshape(s)
        at += n
    return out


def forward(w, X):
    P = unpack(np.asarray(w, np.float64))
    X = np.asarray(X, np.float64)
    a1 = X @ P["W1"].T + P["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ P["W2"].T + P["b2"]
    h2 = np.maximum(a2, 0)
    y = h2 @ P["W3"].T + P["b3"]
    return y, (X, a1, h1, a2, h2)


def loss_grad(w, X, T, M):
    """Loss and gradient with respect to the (already rounded) values w."""
    y, (X, a1, h1, a2, h2) = forward(w, X)
    M = np.asarray(M, np.float64)
    cnt = M.sum()
    e = (y - np.asarray(T, np.float64)) * M
    loss = 0.5 * (e * e).sum() / cnt
    dy = e / cnt
    P = unpack(np.asarray(w, np.float64))
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    d2 = (dy @ P["W3"]) * (a2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ P["W2"]) * (a1 > 0)
    g["W1"] = d1.T @ X
    g["b1"] = d1.sum(0)
    return loss, np.concatenate([g[name].ravel() for name, _ in SHAPES])


def loss_only(w, X, T, M):
    y, _ = forward(w, X)
    M = np.asarray(M, np.float64)
    e = (y - np.asarray(T, np.float64)) * M
    return 0.5 * (e * e).sum() / M.sum()
