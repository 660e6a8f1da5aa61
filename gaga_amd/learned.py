"""Learned warm start for CGIHT: training data, training and evaluation.

Training data are CGIHT recoveries on the (delta, rho) grid of the phase-
transition sweep (delta = linspace(0.05, 0.95, 16), rho = linspace(0.05, 0.5,
16)) at n = 4096: for each grid point one gen A shared by many signals,
solved by the many-right-hand-side CGIHT (gaga_amd.ManyRHS); every
successful recovery (relative l_inf error < 1e-3) contributes per-coordinate
samples labelled by its recovered support.  The network is trained on the
GPU by the library's own kernels (xcs_mlp_train).  At solve time
(Plan.solve_warm) the k best-scored coordinates start CGIHT; plain CGIHT from
x = 0 stays the fallback and the success test stays the reference's.

No counterpart in the reference; its sweep driver is
GAGA/CGIHTcode/shared/tests_delta_fixed.m (the grid) and its CGIHT start is
GAGA/gaga/algorithms.cu:3752-3790.
"""
import math
import os

import numpy as np

from ._capi import MLP, ManyRHS, Plan

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warm_start_mlp.npz")
SUCCESS_TOL = 1e-3


def grid_points(n=4096, deltas=None, rhos=None):
    """(m, k, delta, rho) per grid point; m rounded up to a multiple of 128 (the
    many-RHS product tiles), k = ceil(rho m)."""
    deltas = np.linspace(0.05, 0.95, 16) if deltas is None else deltas
    rhos = np.linspace(0.05, 0.5, 16) if rhos is None else rhos
    pts = []
    for d in deltas:
        m = int(math.ceil(d * n / 128.0) * 128)
        for r in rhos:
            k = max(1, int(math.ceil(r * m - 1e-9)))
            pts.append((m, k, float(d), float(r)))
    return pts


def make_training_set(total=200_000, n=4096, per=256, seed=20260101, first_instance=1 << 24,
                      maxiter=300, log=None):
    """Recoveries over the grid: returns (X, labels, stats).  total = number of
    CGIHT recoveries (signals) over all grid points."""
    import gaga_amd as ga
    pts = grid_points(n)
    count = int(math.ceil(total / len(pts)))
    Xs, ys = [], []
    solved = ok_total = 0
    for g, (m, k, d, r) in enumerate(pts):
        opts = ga.options(tol=1e-5, maxiter=maxiter, seed=seed, vecDistribution="gaussian",
                          instance=first_instance + g * (count + 1))
        R = ManyRHS(m, n, k, count, "bf16", opts)
        R.solve()
        ok = R.results()["norms"][:, 2] < SUCCESS_TOL
        X, y = R.samples(per, seed=seed + g)
        R.close()
        solved += count
        ok_total += int(ok.sum())
        keep = np.repeat(ok, per)
        Xs.append(X[keep])
        ys.append(y[keep])
        if log:
            log("grid %3d/%d delta %.2f rho %.2f (m %d k %d): %d/%d recovered"
                % (g + 1, len(pts), d, r, m, k, ok.sum(), count))
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    return X, y, dict(recoveries=solved, successful=ok_total, samples=len(y),
                      positive_fraction=float(y.mean()) if len(y) else 0.0)


def train(net, X, y, epochs=4, batch=16384, lr=1e-3, seed=1, log=None):
    """Adam over shuffled data resident on the GPU; returns per-epoch losses."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(y))
    pos = float(y.mean())
    pos_w = (1.0 - pos) / max(pos, 1e-6)  # balance the two classes
    net.set_data(X[perm], y[perm])
    steps = max(1, len(y) // batch)
    losses = []
    for e in range(epochs):
        loss, ms = net.train(steps, batch=batch, lr=lr * (0.5 ** e), pos_w=pos_w)
        losses.append(loss)
        if log:
            log("epoch %d: %d steps, last-batch loss %.4f, %.1f ms on the GPU" % (e + 1, steps, loss, ms))
    return losses, pos_w


def save(net, path=WEIGHTS, **meta):
    np.savez(path, params=net.params(), **{k: np.asarray(v) for k, v in meta.items()})


def load_net(path=WEIGHTS):
    """A network with the saved weights (plain data file, no pickles)."""
    d = np.load(path, allow_pickle=False)
    net = MLP(seed=1)
    net.set_params(d["params"])
    return net


def evaluate(net, ens, m, n, k, instances, seed=20260101, first_instance=1 << 26, p=7,
             cg_steps=3):
    """Plain CGIHT against the learned warm start on fresh instances; returns
    per-instance (ms, iterations, success) for both and the fallback flags."""
    import gaga_amd as ga
    rows = []
    for i in range(instances):
        opts = ga.options(tol=1e-5, seed=seed, vecDistribution="gaussian",
                          instance=first_instance + i)
        pl = Plan("CGIHT", ens, k, m, n, p if ens == "smv" else 0, opts)
        try:
            pl.solve()  # warm the plan
            ms_plain = pl.solve()
            r = pl.results()
            plain = (ms_plain, int(r["iters"]), bool(r["norms"][2] < SUCCESS_TOL))
            ms_w, ms_pre, fb = pl.solve_warm(net, cg_steps)
            r = pl.results()
            warm = (ms_w, int(r["iters"]), bool(r["norms"][2] < SUCCESS_TOL))
            ms_m, _, fbm = pl.solve_warm(None, cg_steps)
            r = pl.results()
            mag = (ms_m, int(r["iters"]), bool(r["norms"][2] < SUCCESS_TOL))
            rows.append(dict(plain=plain, warm=warm, magnitude=mag, fallback=fb,
                             magnitude_fallback=fbm, warm_prep_ms=ms_pre))
        finally:
            pl.close()
    return rows
