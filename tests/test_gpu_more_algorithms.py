"""IHT, ThresholdSD, ThresholdCG, FIHT, ALPS and CGIHTprojected on the GPU against the
oracle: same support as the FP32 oracle, the FP64 oracle's solution to 1e-5 (1e-4 on
an instance IHT fails), iteration counts within 10 % (20 % where the stagnation rule
stops), and the device and host-stepped forms bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20260101
ALGS = ["IHT", "ThresholdSD", "ThresholdCG", "FIHT", "ALPS", "CGIHTprojected"]


def _problem(O, ens, m, n, k, inst):
    P = O.make_problem(ens, m, n, SEED, inst)
    x = O.sparse_vector(n, k, "binary", SEED, inst)
    y = P.apply_A(x).astype(np.float32)
    return P, x, y


def _solve_dev(alg, ens, P, y, k, tol, **kw):
    import gaga_amd as ga
    opts = ga.options(tol=tol, **kw)
    if ens == "gen":
        return ga.cs(alg, "gen", y, P.A.T[:P.m], k, opts)
    if ens == "smv":
        cols = np.repeat(np.arange(P.n, dtype=np.int32), np.diff(P.col_ptr))
        return ga.cs(alg, "smv", y, P.rows + 1, cols + 1, P.vals, k, opts)
    return ga.cs(alg, "dct", y, P.dct_rows + 1, k, P.m, P.n, opts)


@pytest.mark.parametrize("alg", ALGS)
@pytest.mark.parametrize("ens,m,n,k", [("gen", 256, 1024, 20), ("smv", 256, 2048, 20),
                                       ("dct", 256, 1024, 20), ("gen", 256, 1024, 5)])
def test_matches_oracle(alg, ens, m, n, k, oracle):
    for inst in range(3):
        print("case", alg, ens, m, n, k, "instance", inst, flush=True)
        P, x, y = _problem(oracle, ens, m, n, k, inst)
        xd, it, rate = _solve_dev(alg, ens, P, y, k, 1e-5)
        x32, r32, _, _ = oracle.solve(alg, P, y, k, tol=1e-5, precision=32)
        x64, r64, _, _ = oracle.solve(alg, P, y, k, tol=1e-5, precision=64)
        if alg == "ALPS" and r64.stop_reason == 5:
            # the cycle test ends ALPS on smv instance 0 at iteration 9 in FP32 and 13 in
            # FP64, unrecovered in both: the device must end it unrecovered as well
            assert np.max(np.abs(xd - x)) / np.max(np.abs(x)) > 1e-3
            continue
        np.testing.assert_array_equal(xd != 0, x32 != 0)
        # where the stop is the stagnation rule (the one-shot thresholds, a failing IHT)
        # its iteration is rounding-sensitive: the FP32 and FP64 oracles differ by up to 5
        # iterations there, and by 9e-6 in x on the failing IHT instance
        stagnated = r64.stop_reason != 0
        assert np.linalg.norm(xd - x64) <= (1e-4 if stagnated and alg == "IHT" else 1e-5) * \
            np.linalg.norm(x64)
        slack = max(5, r32.iters // 5) if stagnated else max(2, r32.iters // 10)
        if stagnated and alg == "ThresholdCG":
            # at the FP32 floor the CG's stop is the 16-step window over squared
            # residuals near 1e-9: rounding noise, so only the count's scale is checked
            slack = max(10, r32.iters // 2)
        assert abs(it - r32.iters) <= slack
        if (alg in ("IHT", "FIHT") and ens != "dct") or alg in ("FIHT", "ALPS", "CGIHTprojected") \
                or k == 5:
            assert np.max(np.abs(xd - x)) / np.max(np.abs(x)) < 1e-3


@pytest.mark.parametrize("alg", ALGS)
def test_plain_and_device_forms_identical(alg):
    import gaga_amd as ga
    res = []
    for mode in ("device", "plain"):
        pl = ga.Plan(alg, "gen", 20, 256, 1024, 0,
                     ga.options(tol=1e-5, seed=SEED, vecDistribution="binary", instance=4,
                                mode=mode))
        pl.solve()
        res.append(pl.results())
        pl.close()
    np.testing.assert_array_equal(res[0]["x"], res[1]["x"])
    assert res[0]["iters"] == res[1]["iters"]


@pytest.mark.parametrize("pft", [1e30, 1e-30])
@pytest.mark.parametrize("ens", ["gen", "dct"])
def test_cgihtp_proj_frac_tol_reaches_the_device(pft, ens, oracle):
    """projFracTol 1e30 keeps CG steps on the first support until the projection guard
    trips; 1e-30 makes every step a steepest-descent step with a fresh threshold.  The
    device follows the oracle in both (support, iterations, FP64 solution)."""
    for inst in range(2):
        print("case CGIHTprojected", ens, "projFracTol", pft, "instance", inst, flush=True)
        P, x, y = _problem(oracle, ens, 256, 1024, 20, inst)
        xd, it, _ = _solve_dev("CGIHTprojected", ens, P, y, 20, 1e-5, projFracTol=pft)
        x32, r32, _, _ = oracle.solve("CGIHTprojected", P, y, 20, tol=1e-5, precision=32,
                                      proj_frac_tol=pft)
        x64, r64, _, _ = oracle.solve("CGIHTprojected", P, y, 20, tol=1e-5, precision=64,
                                      proj_frac_tol=pft)
        np.testing.assert_array_equal(xd != 0, x32 != 0)
        assert np.linalg.norm(xd - x64) <= 1e-5 * np.linalg.norm(x64)
        assert abs(it - r32.iters) <= 2


def test_cgihtp_batch_bit_identical_to_single_instance():
    import gaga_amd as ga
    specs = [dict(ens="gen", k=20, m=256, n=1024, instance=0),
             dict(ens="smv", k=12, m=128, n=1024, p=7, instance=1),
             dict(ens="dct", k=20, m=256, n=1024, instance=2),
             dict(ens="gen", k=9, m=100, n=300, instance=3)]
    base = dict(tol=1e-5, seed=SEED, vecDistribution="binary")
    B = ga.Batch("CGIHTprojected", specs, ga.options(**base))
    B.solve()
    rb = B.results()
    for i, s in enumerate(specs):
        extra = {"matrixEnsemble": "binary"} if s["ens"] == "smv" else {}
        pl = ga.Plan("CGIHTprojected", s["ens"], s["k"], s["m"], s["n"], s.get("p", 0),
                     ga.options(instance=s["instance"], teamSize=1, blockSize=256, **base, **extra))
        pl.solve()
        r = pl.results()
        xb, _ = B.x(i)
        np.testing.assert_array_equal(xb, r["x"])
        assert rb["iters"][i] == r["iters"]
        assert rb["norms"][i, 2] < 1e-3
        pl.close()
    B.close()


@pytest.mark.parametrize("alg,sf", [(a, None) for a in ALGS] + [("NIHT", s) for s in range(4)]
                         + [("HTP", s) for s in range(4)])
def test_edge_cases(alg, sf):
    """gen 50 x 77 (neither a multiple of 64): y = 0 returns x = 0 without iterating,
    k = 0 returns x = 0 (the support stays empty), k = 3 recovers a 3-sparse x; the
    sort-based supportFlags refuse k = 0."""
    import gaga_amd as ga
    m, n = 50, 77
    rng = np.random.default_rng(0)
    A = (rng.standard_normal((m, n)) / np.sqrt(m)).astype(np.float32)
    extra = {} if sf is None else dict(timing="on", supportFlag=sf)
    print("case edge", alg, extra, flush=True)
    x, it, _ = ga.cs(alg, "gen", np.zeros(m, np.float32), A, 3, ga.options(tol=1e-5, **extra))
    assert it == 0 and not x.any()
    xt = np.zeros(n, np.float32)
    xt[:3] = 1
    y = (A @ xt).astype(np.float32)
    if sf is not None and sf >= 2:
        with pytest.raises(ga.GagaError, match="k >= 1"):
            ga.cs(alg, "gen", y, A, 0, ga.options(tol=1e-5, maxiter=20, **extra))
    else:
        x, it, _ = ga.cs(alg, "gen", y, A, 0, ga.options(tol=1e-5, maxiter=20, **extra))
        assert not x.any()
    x, it, _ = ga.cs(alg, "gen", y, A, 3, ga.options(tol=1e-5, maxiter=200, **extra))
    assert np.max(np.abs(x - xt)) < 1e-3
