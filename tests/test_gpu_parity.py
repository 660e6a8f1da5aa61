"""GPU parity: the HIP path against the CPU oracle on identical inputs.

Bars: bit-exact for support selection (bins, k_bin, sum and the binning
state); FP32 matvecs within a bound of the FP64 oracle; recovered vectors
within 1e-5 relative of the oracle's (and of the planted vector).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20260101


def _rng(s):
    return np.random.default_rng(s)


# ---------------------------------------------------------------- generator --
def test_gen_matrix_matches_oracle(oracle, lib):
    from gaga_amd import _capi
    m, n = 96, 40
    A_dev = _capi.generate_gen(m, n, SEED, 7)
    A_orc = oracle.gen_matrix(m, n, SEED, 7)
    d = np.abs(A_dev.view(np.int32).astype(np.int64) - A_orc.view(np.int32).astype(np.int64))
    assert d.max() <= 1, "device Box-Muller differs from the oracle by more than 1 ulp"
    assert (d == 0).mean() > 0.99


@pytest.mark.parametrize("m,n,p,ens,inst", [(1000, 1024, 2, 2, 3), (60, 64, 1, 2, 1),
                                            (900, 1024, 2, 3, 2), (256, 2048, 7, 2, 1),
                                            (3000, 3000, 7, 3, 9), (64, 64, 64, 2, 1),
                                            (8, 16384, 4, 2, 1), (32768, 262144, 7, 3, 0)])
def test_smv_generator_matches_oracle(m, n, p, ens, inst, oracle, lib):
    """Device smv draws, zero-row repair and CSR copy equal the oracle bit for bit
    (the first five cases have empty rows before the repair; m = 8, n = 16384 has
    rows longer than the row sort stages in LDS)."""
    from gaga_amd import _capi
    print("case smv generator", m, n, p, ens, inst, flush=True)
    rows, vals, rp, cols, rv, empty = _capi.generate_smv(m, n, p, ens, SEED, inst)
    r_o, v_o, e_o = oracle.smv_matrix(m, n, p, ens, SEED, inst)
    assert empty == e_o
    np.testing.assert_array_equal(rows, r_o)
    np.testing.assert_array_equal(vals.view(np.int32), v_o.view(np.int32))
    colidx = np.repeat(np.arange(n, dtype=np.int32), p)
    order = np.lexsort((colidx, r_o))
    np.testing.assert_array_equal(rp, np.concatenate([[0], np.cumsum(np.bincount(r_o, minlength=m))]))
    np.testing.assert_array_equal(cols, colidx[order])
    np.testing.assert_array_equal(rv.view(np.int32), v_o[order].view(np.int32))
    assert np.all(np.bincount(rows, minlength=m) > 0)


# ------------------------------------------------------------------ binning --
def _binning_cases():
    r = _rng(1)
    cases = []
    for n, k in [(1, 1), (7, 3), (1000, 20), (4096, 205), (20000, 700), (100000, 13108)]:
        for kind in ("gauss", "ties", "sparse", "tiny"):
            if kind == "gauss":
                v = r.standard_normal(n).astype(np.float32)
            elif kind == "ties":
                v = r.integers(-5, 6, n).astype(np.float32)
            elif kind == "sparse":
                v = np.zeros(n, np.float32)
                idx = r.choice(n, max(1, n // 10), replace=False)
                v[idx] = r.standard_normal(idx.size).astype(np.float32)
            else:
                v = (r.standard_normal(n) * 1e-7).astype(np.float32)
                v[0] = 1.0
            cases.append((n, k, v))
    return cases


@pytest.mark.parametrize("case", range(24))
def test_find_support_set_bit_exact(case, oracle, lib):
    from gaga_amd import _capi
    n, k, v = _binning_cases()[case]
    num_bins = max(n // 20, 1000)
    mx = float(np.max(np.abs(v)))
    if mx == 0:
        pytest.skip("all-zero vector")
    slope = np.float32(num_bins - 1) / np.float32(mx)
    for alpha in (0.25, 0.125, 1e-9):
        max_bin = int(np.float32(num_bins) * (np.float32(1) - np.float32(alpha)))
        for kk in (k, 0, n):
            ref = oracle.find_support_set(v, kk, num_bins, float(slope), mx, 0.0, 0.0, alpha,
                                          max_bin, 0)
            dev = _capi.find_support_set(v, kk, num_bins, float(slope), mx, 0.0, 0.0, alpha,
                                         max_bin, 0)
            assert dev[0] == ref[0], "sum"
            np.testing.assert_array_equal(dev[1], ref[1])
            for key in ("min_val", "alpha", "max_bin", "k_bin"):
                assert dev[2][key] == ref[2][key], key


def test_find_support_set_skip_branch(oracle, lib):
    from gaga_amd import _capi
    v = _rng(3).standard_normal(100).astype(np.float32)
    s, _, st = _capi.find_support_set(v, 10, 1000, 1.0, 1.0, 0.25, 0.75, 0.25, 750, 17)
    assert s == -1 and st["k_bin"] == 17
    assert st["min_val"] == np.float32(0.75) - np.float32(0.25)


# ---------------------------------------------------------------- operators --
def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("m,n", [(1, 1), (5, 3), (256, 1024), (300, 777), (2048, 8192), (9000, 300)])
def test_gen_apply(m, n, oracle, lib):
    from gaga_amd import _capi
    Acols = oracle.gen_matrix(m, n, SEED, 11)
    P = oracle.Problem.gen(Acols, m=m)
    A = Acols.T  # (m, n)
    r = _rng(m + n)
    x = r.standard_normal(n).astype(np.float32)
    x[r.random(n) < 0.7] = 0
    y = _capi.apply("gen", 0, m, n, x, A=A)
    assert _rel(y, P.apply_A(x)) < 1e-5
    res = r.standard_normal(m).astype(np.float32)
    g = _capi.apply("gen", 1, m, n, res, A=A)
    assert _rel(g, P.apply_AT(res)) < 1e-5


@pytest.mark.parametrize("m,n,p", [(16, 64, 3), (1024, 8192, 7), (32768, 262144, 7)])
def test_smv_apply(m, n, p, oracle, lib):
    from gaga_amd import _capi
    rows, vals, _ = oracle.smv_matrix(m, n, p, 2, SEED, 5)
    P = oracle.Problem.smv_fixed_p(m, n, p, rows, vals)
    r = _rng(n)
    x = r.standard_normal(n).astype(np.float32)
    x[r.random(n) < 0.9] = 0
    y = _capi.apply("smv", 0, m, n, x, col_ptr=P.col_ptr, rows=rows, vals=vals)
    assert _rel(y, P.apply_A(x)) < 1e-5
    res = r.standard_normal(m).astype(np.float32)
    g = _capi.apply("smv", 1, m, n, res, col_ptr=P.col_ptr, rows=rows, vals=vals)
    assert _rel(g, P.apply_AT(res)) < 1e-5


@pytest.mark.parametrize("m,n", [(2, 2), (3, 8), (100, 1024), (512, 4096), (1000, 8192),
                                 (8192, 65536), (131072, 1048576)])
def test_dct_apply(m, n, oracle, lib):
    from gaga_amd import _capi
    rows = oracle.dct_rows(m, n, SEED, 3)
    P = oracle.Problem("dct", m, n, dct_rows=rows)
    r = _rng(n)
    x = r.standard_normal(n).astype(np.float32)
    y = _capi.apply("dct", 0, m, n, x, dct_rows=rows)
    assert _rel(y, P.apply_A(x)) < 2e-6 * np.log2(n) + 1e-6
    res = r.standard_normal(m).astype(np.float32)
    g = _capi.apply("dct", 1, m, n, res, dct_rows=rows)
    assert _rel(g, P.apply_AT(res)) < 2e-6 * np.log2(n) + 1e-6


# --------------------------------------------------------------- end to end --
def _problem(ens, m, n, k, inst, dist="binary", p=7):
    from oracle import oracle as O
    P = O.make_problem(ens, m, n, SEED, inst, p=p)
    x = O.sparse_vector(n, k, dist, SEED, inst)
    y = P.apply_A(x).astype(np.float32)
    return P, x, y


def _solve_dev(alg, ens, P, y, k, tol, maxiter=None):
    import gaga_amd as ga
    opts = ga.options(tol=tol, **({"maxiter": maxiter} if maxiter else {}))
    if ens == "gen":
        return ga.cs(alg, "gen", y, P.A.T[:P.m], k, opts)
    if ens == "smv":
        nz = P.rows.size
        cols = np.repeat(np.arange(P.n, dtype=np.int32), np.diff(P.col_ptr))
        return ga.cs(alg, "smv", y, P.rows + 1, cols + 1, P.vals, k, opts)
    return ga.cs(alg, "dct", y, P.dct_rows + 1, k, P.m, P.n, opts)


CSMPSP_OFF = ("CSMPSP is switched off on the GPU: this case (gen 256x1024, instances 0-2, one "
              "workgroup) faulted the device once after passing before, a 4-workgroup CSMPSP "
              "case faulted earlier, and no cause has been shown; the oracle tests keep CSMPSP")


@pytest.mark.parametrize("alg", ["NIHT", "HTP", pytest.param("CSMPSP", marks=pytest.mark.skip(
    reason=CSMPSP_OFF)), "CGIHT"])
@pytest.mark.parametrize("ens,m,n,k", [("gen", 256, 1024, 20), ("smv", 256, 2048, 20),
                                       ("dct", 256, 1024, 20)])
def test_solve_matches_oracle(alg, ens, m, n, k, oracle):
    for inst in range(3):
        P, x, y = _problem(ens, m, n, k, inst)
        xd, it, rate = _solve_dev(alg, ens, P, y, k, 1e-5)
        x32, r32, _, _ = oracle.solve(alg, P, y, k, tol=1e-5, precision=32)
        x64, r64, _, _ = oracle.solve(alg, P, y, k, tol=1e-5, precision=64)
        # recovered, and the same answer as the oracle's
        assert np.max(np.abs(xd - x)) / np.max(np.abs(x)) < 1e-3
        np.testing.assert_array_equal(xd != 0, x32 != 0)
        assert _rel(xd, x64) < 1e-5
        assert abs(it - r32.iters) <= max(2, r32.iters // 10)


def test_csmpsp_is_switched_off():
    """The switched-off algorithm fails loudly before any device work."""
    import gaga_amd as ga
    with pytest.raises(ga.GagaError, match="switched off"):
        ga.cs("CSMPSP", "gen", 20, 256, 1024, ga.options(tol=1e-5))
