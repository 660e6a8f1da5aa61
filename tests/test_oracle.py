"""Pins the CPU oracle (runs without a GPU).

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails it: library-computed
fixtures (tests/golden), known-answer vectors, closed forms, invariants and
special cases that reduce to a textbook routine.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SEED = 20260101


# ------------------------------------------------------------------ philox --
def test_philox_known_answers(oracle):
    # Random123 kat_vectors, philox4x32 with 10 rounds
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, out in kat:
        assert oracle.philox(ctr, key).tolist() == out


# --------------------------------------------------------------- operators --
def test_gen_ops_fixture(oracle):
    f = np.load(os.path.join(GOLDEN, "gen_ops.npz"))
    P = oracle.Problem.gen(f["A_cols"])
    np.testing.assert_allclose(P.apply_A(f["x"]), f["Ax"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(P.apply_AT(f["r"]), f["ATr"], rtol=1e-12, atol=1e-12)


def test_smv_ops_fixture(oracle):
    f = np.load(os.path.join(GOLDEN, "smv_ops.npz"))
    P = oracle.Problem("smv", int(f["m"]), int(f["n"]), col_ptr=f["col_ptr"], rows=f["rows"],
                       vals=f["vals"])
    np.testing.assert_allclose(P.apply_A(f["x"]), f["Ax"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(P.apply_AT(f["r"]), f["ATr"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n", [2, 8, 64, 1024])
def test_dct_ops_fixture(n, oracle):
    f = np.load(os.path.join(GOLDEN, "dct_ops.npz"))
    rows = f["rows_%d" % n]
    P = oracle.Problem("dct", rows.size, n, dct_rows=rows)
    np.testing.assert_allclose(P.apply_A(f["x_%d" % n]), f["Ax_%d" % n], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(P.apply_AT(f["r_%d" % n]), f["ATr_%d" % n], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("ens", ["gen", "smv", "dct"])
def test_adjoint_and_dct_rows_orthonormal(ens, oracle):
    m, n = 64, 256
    P = oracle.make_problem(ens, m, n, SEED, 3, p=5)
    r = np.random.default_rng(1)
    x, y = r.standard_normal(n), r.standard_normal(m)
    # <A x, y> == <x, A^T y>
    assert abs(P.apply_A(x) @ y - x @ P.apply_AT(y)) < 1e-10 * max(1.0, abs(x @ P.apply_AT(y)))
    if ens == "dct":
        np.testing.assert_allclose(P.apply_A(P.apply_AT(y)), y, atol=1e-12)  # A A^T = I_m


def test_dct_known_answers(oracle):
    n = 16
    # DCT-II of a delta at 0 is the first column of the transform: w_q cos(0) = w_q
    d = np.zeros(n)
    d[0] = 1.0
    X = np.zeros(n)
    import ctypes
    oracle.lib().orc_dct2_ortho(n, d.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p))
    w = np.full(n, np.sqrt(2.0 / n))
    w[0] = np.sqrt(1.0 / n)
    q = np.arange(n)
    np.testing.assert_allclose(X, w * np.cos(np.pi * q * (2 * 0 + 1) / (2 * n)), atol=1e-14)
    # a constant maps to sqrt(n) * c on the DC coefficient only
    c = np.full(n, 3.0)
    oracle.lib().orc_dct2_ortho(n, c.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p))
    np.testing.assert_allclose(X, np.r_[3.0 * np.sqrt(n), np.zeros(n - 1)], atol=1e-12)


# ----------------------------------------------------------------- binning --
def _state(v, num_bins, alpha=0.25):
    mx = np.float32(np.max(np.abs(v)))
    slope = np.float32(num_bins - 1) / mx
    max_bin = int(np.float32(num_bins) * (np.float32(1) - np.float32(alpha)))
    return float(slope), float(mx), max_bin


def test_binning_topk_special_case(oracle):
    # fine bins and well separated magnitudes: at most one entry per bin, so the
    # kept set {bin <= k_bin} is exactly the k largest magnitudes
    r = np.random.default_rng(2)
    n, k, num_bins = 200, 17, 100000
    v = (r.permutation(n) + 1).astype(np.float32) * np.where(r.random(n) < 0.5, -1, 1)
    slope, mx, max_bin = _state(v, num_bins)
    s, bins, st = oracle.find_support_set(v, k, num_bins, slope, mx, 0.0, 0.0, 0.25, max_bin, 0)
    kept = set(np.nonzero(bins <= st["k_bin"])[0])
    top = set(np.argpartition(-np.abs(v), k)[:k])
    assert s == k and kept == top


@pytest.mark.parametrize("seed", range(6))
def test_binning_invariants(seed, oracle):
    r = np.random.default_rng(seed)
    n = int(r.integers(5, 3000))
    k = int(r.integers(1, n))
    v = (r.standard_normal(n) ** 3).astype(np.float32)
    v[r.random(n) < 0.2] = 0.0
    num_bins = max(n // 20, 1000)
    slope, mx, max_bin = _state(v, num_bins)
    s, bins, st = oracle.find_support_set(v, k, num_bins, slope, mx, 0.0, 0.0, 0.25, max_bin, 0)
    a = np.abs(v)
    order = np.argsort(-a, kind="stable")
    # bins are non-increasing in magnitude
    assert np.all(np.diff(bins[order]) >= 0)
    # counters are the histogram of the counted entries
    counted = (a.astype(np.float64) > np.float64(mx) * 1e-6) & (bins < max_bin)
    np.testing.assert_array_equal(st["counters"][:max_bin], np.bincount(bins[counted], minlength=num_bins)[:max_bin])
    kb = st["k_bin"]
    cum = np.cumsum(st["counters"][:max_bin])
    if s >= k:  # k_bin is the first bin at which the running count reaches k
        assert cum[kb] == s and (kb == 0 or cum[kb - 1] < k)
        assert np.count_nonzero(bins <= kb) == s  # the kept set has `sum` entries
        assert st["alpha"] == 0.25
    else:       # range exhausted: alpha halves, MaxBin widens
        assert kb == max_bin - 1 and s == cum[-1]
        assert st["alpha"] == np.float32(0.125)
        assert st["max_bin"] == int(np.float32(num_bins) * (np.float32(1) - np.float32(0.125)))
    # kept set is upward closed in magnitude
    kept = bins <= kb
    if kept.any() and (~kept).any():
        assert a[kept].min() >= a[~kept].max()
    # min_val is the magnitude at the k-bin edge
    assert st["min_val"] == np.float32(mx) - np.float32(kb + 1) / np.float32(slope)


def test_binning_skip_and_zero_k(oracle):
    v = np.arange(1, 11, dtype=np.float32)
    slope, mx, max_bin = _state(v, 1000)
    s, _, st = oracle.find_support_set(v, 3, 1000, slope, mx, 0.5, 2.0, 0.25, max_bin, 7)
    assert s == -1 and st["k_bin"] == 7 and st["min_val"] == np.float32(1.5)
    s, bins, st = oracle.find_support_set(v, 0, 1000, slope, mx, 0.0, 0.0, 0.25, max_bin, 0)
    assert s == 0 and st["k_bin"] == -1 and not np.any(bins <= -1)


# ---------------------------------------------------------- CG and stops ----
def test_cg_fixed_support_is_least_squares(oracle):
    f = np.load(os.path.join(GOLDEN, "lstsq.npz"))
    P = oracle.Problem.gen(f["A_cols"])
    mask = np.zeros(P.n, np.int32)
    mask[f["supp"]] = 1
    x, steps = oracle.cg_fixed(P, f["y"], mask, tol=1e-12)
    assert steps <= 15
    np.testing.assert_allclose(x[f["supp"]], f["sol"], rtol=1e-6, atol=1e-9)
    assert np.all(x[mask == 0] == 0)


def test_stop_rules_synthetic_histories(oracle):
    evo0 = np.ones(16, np.float32)
    # decreasing residuals: diff = largest decrease, no cycle, no fail
    prev = np.float32(0.5) ** np.arange(16, dtype=np.float32)
    evo, diff, cyc, fail = oracle.stop_update(prev, evo0, 10, 749)
    assert evo[15] == prev[14] - prev[15] and diff == 1.0 and cyc == 1 and fail == 0
    # exact repeat of an earlier norm triggers the cycle test, but only after iter 4
    prev2 = prev.copy()
    prev2[15] = prev2[3]
    assert oracle.stop_update(prev2, evo0, 5, 749)[2] == 0
    assert oracle.stop_update(prev2, evo0, 4, 749)[2] == 1
    # slow geometric decay: rate^(1/15) = 0.9995 > 0.999 fails only after the start iteration
    slow = (np.float32(0.9995) ** np.arange(16)).astype(np.float32)
    assert oracle.stop_update(slow, evo0, 750, 749)[3] == 1
    assert oracle.stop_update(slow, evo0, 749, 749)[3] == 0
    assert oracle.stop_update(slow, evo0, 126, 125)[3] == 1
    fast = (np.float32(0.99) ** np.arange(16)).astype(np.float32)
    assert oracle.stop_update(fast, evo0, 800, 749)[3] == 0


# -------------------------------------------------------------- generators --
def test_generators(oracle):
    n, k = 5000, 123
    idx = oracle.support(n, k, SEED, 9)
    assert idx.size == k and np.all(np.diff(idx) > 0) and idx.min() >= 0 and idx.max() < n
    x = oracle.sparse_vector(n, k, "binary", SEED, 9)
    assert set(np.unique(x[x != 0])) <= {-1.0, 1.0} and np.count_nonzero(x) == k
    m = 400
    A = oracle.gen_matrix(m, 300, SEED, 4)
    assert abs(A.mean()) < 5e-3 and abs(A.var() * m - 1.0) < 0.02
    rows, vals, empty = oracle.smv_matrix(256, 2048, 7, 2, SEED, 1)
    R = rows.reshape(2048, 7)
    assert np.all(np.diff(R, axis=1) > 0) and empty == 0
    np.testing.assert_allclose(np.abs(vals), 1 / np.sqrt(np.float32(7)), rtol=1e-7)
    d = oracle.dct_rows(300, 4096, SEED, 2)
    assert d.size == 300 and np.all(np.diff(d) > 0) and d.max() < 4096


def test_results_closed_forms(oracle):
    xt = np.array([1, 0, -2, 0, 0, 4], np.float32)
    xo = np.array([1, 0.5, -1, 0, 0, 4], np.float32)
    norms, sup = oracle.results(xt, xo, 3)
    np.testing.assert_allclose(norms, [1.5 / 7, np.sqrt(1.25) / np.sqrt(21), 1.0 / 4], rtol=1e-12)
    assert sup.tolist() == [3, 0, 2, 1]


# -------------------------------------------------------------- algorithms --
@pytest.mark.parametrize("alg", ["NIHT", "HTP", "CSMPSP", "CGIHT"])
@pytest.mark.parametrize("ens,m,n,k", [("gen", 256, 1024, 20), ("smv", 256, 2048, 20),
                                       ("dct", 256, 1024, 20)])
def test_recovery_both_precisions(alg, ens, m, n, k, oracle):
    for inst in range(2):
        P = oracle.make_problem(ens, m, n, SEED, inst, p=7)
        x = oracle.sparse_vector(n, k, "binary", SEED, inst)
        y = P.apply_A(x).astype(np.float32)
        x64, r64, tk64, _ = oracle.solve(alg, P, y, k, tol=1e-5, precision=64, trace=True)
        x32, r32, tk32, _ = oracle.solve(alg, P, y, k, tol=1e-5, precision=32, trace=True)
        assert np.max(np.abs(x64 - x)) < 1e-3 and np.max(np.abs(x32 - x)) < 1e-3
        assert r64.stop_reason == 0 and np.all(tk64 >= 0)
        np.testing.assert_array_equal(x32 != 0, x != 0)


def test_cgiht_restart_and_maxiter(oracle):
    P = oracle.make_problem("gen", 128, 512, SEED, 5)
    x = oracle.sparse_vector(512, 10, "gaussian", SEED, 5)
    y = P.apply_A(x).astype(np.float32)
    xr, rr, _, _ = oracle.solve("CGIHT", P, y, 10, tol=1e-6, restart=1)
    assert np.max(np.abs(xr - x)) / np.max(np.abs(x)) < 1e-3
    _, rm, _, _ = oracle.solve("NIHT", P, y, 10, tol=1e-6, maxiter=3)
    assert rm.iters == 3 and rm.stop_reason == 1
