"""The oracle's CGIHTprojected (CGIHTprojected_S_gen, algorithms.cu:4486-4789) pinned
by what the mathematics fixes, not by restating it:

- with projFracTol huge, every iteration after the first is a CG step on the first
  iteration's support, so the k-th iterate minimises ||y - A z|| over z in
  x_1 + K_{k-1}(A_S^T A_S, (A^T r_1)_S) (the Krylov optimality of CG, checked with an
  explicit orthonormal Krylov basis and numpy.linalg.lstsq);
- with projFracTol tiny, every iteration is a steepest-descent step with the exact
  line search along the gradient restricted to the current support, followed by the
  hard threshold: x_j equals x_{j-1} + mu g on its support, which holds the largest
  entries of that point;
- at the default projFracTol = 3 (gaga_gen.cu:190) it recovers the planted signals
  of the parity problems in both precisions.
"""
import numpy as np
import pytest

SEED = 20260101


def _problem(O, ens, m, n, k, inst):
    P = O.make_problem(ens, m, n, SEED, inst)
    x = O.sparse_vector(n, k, "binary", SEED, inst)
    y = P.apply_A(x).astype(np.float32)
    return P, x, y


def _columns(P, S, n):
    return np.stack([P.apply_A(np.eye(n)[j]) for j in S], axis=1)


@pytest.mark.parametrize("ens,m,n,k", [("gen", 128, 512, 40), ("smv", 128, 1024, 30),
                                       ("dct", 128, 512, 35)])
def test_cg_phase_is_krylov_optimal(oracle, ens, m, n, k):
    P, _, y = _problem(oracle, ens, m, n, k, 0)
    y64 = y.astype(np.float64)
    x1, _, _, _ = oracle.solve("CGIHTprojected", P, y, k, tol=1e-5, maxiter=1, precision=64,
                               proj_frac_tol=1e30)
    S = np.flatnonzero(x1)
    AS = _columns(P, S, n)
    r1 = y64 - AS @ x1[S]
    g = AS.T @ r1
    H = AS.T @ AS
    checked = 0
    for j in range(2, 9):
        xj, rj, _, _ = oracle.solve("CGIHTprojected", P, y, k, tol=1e-5, maxiter=j, precision=64,
                                    proj_frac_tol=1e30)
        if rj.iters < j or not np.array_equal(np.flatnonzero(xj), S):
            break  # converged, or the projection guard forced a fresh threshold
        V = [g / np.linalg.norm(g)]
        for _ in range(j - 2):
            w = H @ V[-1]
            for _ in range(2):  # twice-orthogonalised Gram-Schmidt
                for v in V:
                    w = w - (v @ w) * v
            V.append(w / np.linalg.norm(w))
        V = np.stack(V, 1)
        c = np.linalg.lstsq(AS @ V, r1, rcond=None)[0]
        best = np.linalg.norm(r1 - AS @ (V @ c))
        got = np.linalg.norm(y64 - P.apply_A(xj))
        assert abs(got - best) <= 1e-10 * best, (j, got, best)
        checked += 1
    assert checked >= 5


@pytest.mark.parametrize("ens,m,n,k", [("gen", 256, 1024, 20), ("dct", 256, 1024, 20),
                                       ("gen", 128, 512, 40)])
def test_sd_phase_is_exact_line_search_then_threshold(oracle, ens, m, n, k):
    P, _, y = _problem(oracle, ens, m, n, k, 1)
    y64 = y.astype(np.float64)
    x_prev, _, _, _ = oracle.solve("CGIHTprojected", P, y, k, tol=1e-5, maxiter=0, precision=64,
                                   proj_frac_tol=1e-30)
    # x_0 is the initial hard threshold of A^T y: it keeps the largest entries
    g0 = P.apply_AT(y64)
    S0 = np.flatnonzero(x_prev)
    np.testing.assert_allclose(x_prev[S0], g0[S0], rtol=1e-12)
    assert np.min(np.abs(g0[S0])) >= np.max(np.abs(np.delete(g0, S0)))
    for j in range(1, 6):
        xj, rj, _, _ = oracle.solve("CGIHTprojected", P, y, k, tol=1e-5, maxiter=j, precision=64,
                                    proj_frac_tol=1e-30)
        if rj.iters < j:
            break
        S = np.flatnonzero(x_prev)
        r = y64 - P.apply_A(x_prev)
        g = P.apply_AT(r)
        gS = np.zeros(n)
        gS[S] = g[S]
        d = P.apply_A(gS)
        mu = np.linalg.lstsq(d[:, None], r, rcond=None)[0][0]  # argmin ||r - mu A g_S||
        z = x_prev + mu * g
        Sj = np.flatnonzero(xj)
        # a top set of z (fewer than k entries when bins past MaxBin were not counted)
        assert len(Sj) > 0
        np.testing.assert_allclose(xj[Sj], z[Sj], rtol=1e-9, atol=1e-12 * np.max(np.abs(z)))
        assert np.min(np.abs(z[Sj])) >= np.max(np.abs(np.delete(z, Sj))) * (1 - 1e-9)
        x_prev = xj


@pytest.mark.parametrize("ens,m,n", [("gen", 256, 1024), ("smv", 256, 2048), ("dct", 256, 1024)])
def test_cgihtp_recovers(oracle, ens, m, n):
    for inst in range(3):
        P, x, y = _problem(oracle, ens, m, n, 20, inst)
        for prec in (32, 64):
            xo, res, _, _ = oracle.solve("CGIHTprojected", P, y, 20, tol=1e-5, precision=prec)
            assert res.stop_reason == 0
            assert np.max(np.abs(xo - x)) < 1e-4


def test_proj_frac_tol_selects_the_step(oracle):
    """A huge projFracTol and a tiny one give different iterates after the first
    (CG against steepest descent) and the same first iterate."""
    P, _, y = _problem(oracle, "gen", 128, 512, 40, 0)
    xs = []
    for pft in (1e30, 1e-30):
        xj, _, _, _ = oracle.solve("CGIHTprojected", P, y, 40, tol=1e-5, maxiter=3, precision=64,
                                   proj_frac_tol=pft)
        xs.append(xj)
    assert np.max(np.abs(xs[0] - xs[1])) > 1e-6
    x1a, _, _, _ = oracle.solve("CGIHTprojected", P, y, 40, tol=1e-5, maxiter=1, precision=64,
                                proj_frac_tol=1e30)
    x1b, _, _, _ = oracle.solve("CGIHTprojected", P, y, 40, tol=1e-5, maxiter=1, precision=64,
                                proj_frac_tol=1e-30)
    np.testing.assert_array_equal(x1a, x1b)  # the first iteration is always a descent step
