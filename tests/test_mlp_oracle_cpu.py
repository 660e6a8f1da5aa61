"""The CPU restatement of the warm-start network: backward pinned by central
finite differences (float64, no BF16 rounding), and the feature map by
closed-form values."""
import numpy as np

from oracle import mlp_ref as R


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(R.NP)
    p[R.OW1:R.OB1] = rng.standard_normal(R.H * R.F) * 0.5
    p[R.OB1:R.OW2] = rng.standard_normal(R.H) * 0.1
    p[R.OW2:R.OB2] = rng.standard_normal(R.H * R.H) * (2.0 / R.H) ** 0.5
    p[R.OB2:R.OW3] = rng.standard_normal(R.H) * 0.1
    p[R.OW3:R.OB3] = rng.standard_normal(R.H) * (1.0 / R.H) ** 0.5
    p[R.OB3] = 0.1
    return p


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = _params(1)
    X = rng.standard_normal((40, R.F))
    y = (rng.random(40) < 0.3).astype(float)
    _, g, _ = R.forward_backward(p, X, y, pos_w=2.0, bf16=False)
    idx = np.concatenate([rng.integers(R.OW1, R.OB1, 6), rng.integers(R.OB1, R.OW2, 4),
                          rng.integers(R.OW2, R.OB2, 8), rng.integers(R.OB2, R.OW3, 4),
                          rng.integers(R.OW3, R.OB3, 4), [R.OB3]])
    h = 1e-6
    for i in idx:
        e = np.zeros(R.NP)
        e[i] = h
        lp, _, _ = R.forward_backward(p + e, X, y, pos_w=2.0, bf16=False)
        lm, _, _ = R.forward_backward(p - e, X, y, pos_w=2.0, bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_loss_closed_form():
    p = np.zeros(R.NP)
    p[R.OB3] = 0.0  # logit 0 everywhere: loss = log 2
    X = np.ones((5, R.F))
    y = np.array([0, 1, 0, 1, 1.0])
    loss, g, logit = R.forward_backward(p, X, y, bf16=False)
    assert np.allclose(logit, 0.0) and abs(loss - np.log(2)) < 1e-12
    assert abs(g[R.OB3] - np.mean(0.5 - y)) < 1e-12


def test_features_closed_form():
    g = np.array([4.0, -2.0, 1.0, 0.0, -8.0, 0.5], np.float32)
    X = mlp = R.features(g, m=3, k=2)
    # t_2 = 4, t_4 = 1, t_1 = 8, max = 8, rms = sqrt(85.25 / 6)
    np.testing.assert_allclose(X[:, 0], np.abs(g) / 4, rtol=1e-6)
    np.testing.assert_allclose(X[:, 1], np.abs(g) / 1, rtol=1e-6)
    np.testing.assert_allclose(X[:, 2], np.abs(g) / 8, rtol=1e-6)
    np.testing.assert_allclose(X[:, 3], np.abs(g) / 8, rtol=1e-6)
    np.testing.assert_allclose(X[:, 4], np.abs(g) / np.sqrt(85.25 / 6), rtol=1e-6)
    assert (X[:, 5] == np.float32(0.5)).all() and (X[:, 6] == np.float32(2 / 3)).all()
    assert mlp.shape == (6, 8)
