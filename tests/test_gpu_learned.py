"""Learned warm start on the GPU: network gradients against the CPU oracle,
features against their restatement, training, the warm-start solve and its
fallback, and what the trained network gains over CGIHT without it."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20260101


def test_gradients_match_oracle():
    import gaga_amd as ga
    from oracle import mlp_ref as R
    net = ga.MLP(seed=3)
    rng = np.random.default_rng(5)
    B = 777  # not a multiple of any tile or chunk
    X = np.abs(rng.standard_normal((B, 8))).astype(np.float32) * 2
    y = (rng.random(B) < 0.3).astype(np.float32)
    g, loss = net.grad(X, y, pos_w=2.0)
    p = net.params()
    loss_ref, g_ref, _ = R.forward_backward(p, X, y, pos_w=2.0, bf16=True)
    assert abs(loss - loss_ref) <= 1e-5 * abs(loss_ref)
    blocks = {"W1": (R.OW1, R.OB1), "b1": (R.OB1, R.OW2), "W2": (R.OW2, R.OB2),
              "b2": (R.OB2, R.OW3), "W3": (R.OW3, R.OB3), "b3": (R.OB3, R.NP)}
    for name, (a, b) in blocks.items():
        err = np.linalg.norm(g[a:b] - g_ref[a:b]) / np.linalg.norm(g_ref[a:b])
        assert err < 2e-3, (name, err)
    net.close()


def test_features_match_oracle():
    import gaga_amd as ga
    from oracle import mlp_ref as R
    rng = np.random.default_rng(2)
    for n, m, k in ((1000, 300, 30), (4096, 1024, 200)):
        g = rng.standard_normal(n).astype(np.float32)
        X = ga.mlp_features(g, m, k)
        np.testing.assert_allclose(X, R.features(g, m, k), rtol=2e-6, atol=1e-7)


def test_training_learns_a_rule():
    import gaga_amd as ga
    rng = np.random.default_rng(7)
    N = 65536
    X = (np.abs(rng.standard_normal((N, 8))) * 2).astype(np.float32)
    y = ((X[:, 0] > 1.0) & (X[:, 4] > 0.5)).astype(np.float32)
    net = ga.MLP(seed=4)
    net.set_data(X, y)
    l0, _ = net.train(1, batch=8192, lr=1e-3)
    l1, ms = net.train(300, batch=8192, lr=1e-3)
    assert l1 < 0.5 * l0
    acc = ((net.scores(X[:8192]) > 0) == (y[:8192] > 0.5)).mean()
    assert acc > 0.97
    net.close()


WARM_OFF = ("one run of this test ended in a GPU memory fault inside the warm-start solve whose "
            "cause is not yet shown (a CPU run of the same solve, on the same instance, team "
            "size and masks, under ASan, guard pages and TSan, is clean); the warm-start solve "
            "is switched off and stays off the GPU until the cause is found")


def test_warm_start_reports_switched_off():
    """The switched-off entry point fails loudly and launches nothing."""
    import gaga_amd as ga
    pl = ga.Plan("CGIHT", "gen", 20, 256, 1024, 0,
                 ga.options(tol=1e-5, seed=SEED, vecDistribution="binary", instance=1))
    with pytest.raises(ga.GagaError, match="switched off"):
        pl.solve_warm(None)
    pl.solve()  # the plan still solves without the warm start
    assert pl.results()["norms"][2] < 1e-3
    pl.close()


@pytest.mark.skip(reason=WARM_OFF)
def test_warm_start_solves_and_falls_back():
    import gaga_amd as ga
    pl = ga.Plan("CGIHT", "gen", 40, 512, 2048, 0,
                 ga.options(tol=1e-5, seed=SEED, vecDistribution="gaussian", instance=11))
    # magnitude start (no network)
    ms, pre, fb = pl.solve_warm(None)
    assert pl.results()["norms"][2] < 1e-3
    # an untrained network: success is still reached (by itself or by the fallback)
    net = ga.MLP(seed=9)
    ms, pre, fb = pl.solve_warm(net)
    assert pl.results()["norms"][2] < 1e-3 and pre > 0
    net.close()
    pl.close()


@pytest.mark.skip(reason=WARM_OFF)
def test_trained_network_gain():
    """The committed network (trained on the GPU by train_warm_start on 200k
    recoveries) against CGIHT without it on fresh subsampled-dct instances
    near the transition (n = 2^16, delta = 1/4, rho = 0.2): no fewer
    recoveries, fewer CGIHT iterations."""
    from gaga_amd import learned
    if not os.path.exists(learned.WEIGHTS):
        pytest.skip("no trained network committed (gaga_amd/warm_start_mlp.npz)")
    net = learned.load_net()
    rows = learned.evaluate(net, "dct", 16384, 65536, 3277, instances=6, cg_steps=3)
    plain_ok = sum(r["plain"][2] for r in rows)
    warm_ok = sum(r["warm"][2] for r in rows)
    it_plain = np.mean([r["plain"][1] for r in rows])
    it_warm = np.mean([r["warm"][1] for r in rows])
    print("plain %d/6 ok, %.1f iterations, %.1f ms; warm %d/6 ok, %.1f iterations, %.1f ms"
          % (plain_ok, it_plain, np.mean([r["plain"][0] for r in rows]), warm_ok, it_warm,
             np.mean([r["warm"][0] for r in rows])))
    assert warm_ok >= plain_ok
    assert it_warm < it_plain
    net.close()
