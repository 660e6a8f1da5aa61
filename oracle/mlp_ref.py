"""CPU restatement of the warm-start network (test infrastructure only).

Forward and backward of the 8 -> 512 -> 512 -> 1 ReLU network with the
weighted binary cross-entropy, in float64, with the device's BF16 roundings
applied where its kernels round (layer-1 activations, W2 and the layer-2
gradient dz2 as MFMA operands) when bf16=True.  The feature map of the
device (xcs_mlp.h, mlp_features) is restated in `features`.  The network has
no counterpart in the reference; its backward is pinned by finite
differences (tests/test_mlp_oracle_cpu.py).
"""
import numpy as np

F, H = 8, 512
OW1, OB1 = 0, H * F
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + H
NP = OB3 + 1


def _bf16(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OW1:OB1].reshape(H, F), p[OB1:OW2], p[OW2:OB2].reshape(H, H), p[OB2:OW3],
            p[OW3:OB3], p[OB3])


def forward_backward(p, X, y, pos_w=1.0, bf16=True):
    """Returns (mean loss, gradient vector (NP,), logits)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    B = X.shape[0]
    rnd = _bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    z1 = X @ W1.T + b1
    h1 = rnd(np.maximum(z1, 0.0).astype(np.float32) if bf16 else np.maximum(z1, 0.0))
    W2q = rnd(W2.astype(np.float32) if bf16 else W2)
    z2 = h1 @ W2q.T
    a2 = np.maximum(z2 + b2, 0.0)
    logit = a2 @ W3 + b3
    w = np.where(y > 0.5, pos_w, 1.0)
    loss = w * (np.maximum(logit, 0) - logit * y + np.log1p(np.exp(-np.abs(logit))))
    sig = 1.0 / (1.0 + np.exp(-logit))
    dl = w * (sig - y) / B
    g = np.zeros(NP)
    g[OW3:OB3] = dl @ a2
    g[OB3] = dl.sum()
    dz2 = np.where(a2 > 0, dl[:, None] * W3[None, :], 0.0)
    g[OB2:OW3] = dz2.sum(axis=0)
    g[OW2:OB2] = (dz2.T @ h1).ravel()
    dz2q = rnd(dz2.astype(np.float32) if bf16 else dz2)
    dh1 = dz2q @ W2q
    dz1 = np.where(h1 > 0, dh1, 0.0)
    g[OW1:OB1] = (dz1.T @ X).ravel()
    g[OB1:OW2] = dz1.sum(axis=0)
    return float(loss.mean()), g, logit


def kth_largest_abs(g, k):
    a = np.sort(np.abs(np.asarray(g, np.float32)))[::-1]
    return a[min(max(k, 1), len(a)) - 1]


def features(g, m, k):
    """Per-coordinate features of g = A^T y (float32, the device's formulas)."""
    g = np.asarray(g, np.float32)
    n = g.size
    kk = max(1, min(k, n))
    tk = kth_largest_abs(g, kk)
    t2k = kth_largest_abs(g, min(n, 2 * kk))
    th = kth_largest_abs(g, max(1, (kk + 1) // 2))
    a = np.abs(g)
    mx = np.float32(np.abs(g.astype(np.float64)).max())
    rms = np.float32(np.sqrt((g.astype(np.float64) ** 2).sum() / n))
    inv = lambda d: np.float32(1.0) / d if d > 0 else np.float32(0.0)
    X = np.empty((n, F), np.float32)
    X[:, 0] = a * inv(tk)
    X[:, 1] = a * inv(t2k)
    X[:, 2] = a * inv(th)
    X[:, 3] = a * inv(mx)
    X[:, 4] = a * inv(rms)
    X[:, 5] = np.float32(m) / np.float32(n)
    X[:, 6] = np.float32(k) / np.float32(m)
    X[:, 7] = np.minimum(np.log2(np.float32(1) + a * inv(rms)), np.float32(8)) * np.float32(0.125)
    return X
