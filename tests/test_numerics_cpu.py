"""Known values pinning the test-side roundings of bf16 and OCP e4m3."""
import numpy as np

from numerics import bf16_rne, e4m3_rne


def test_bf16_known_values():
    f = np.float32
    assert bf16_rne(f(1.0)) == 1.0
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert bf16_rne(f(1.0 + 2 ** -8)) == 1.0
    # 1 + 3*2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even -> 1 + 2^-6
    assert bf16_rne(f(1.0 + 3 * 2 ** -8)) == f(1.0 + 2 ** -6)
    assert bf16_rne(f(1.0 + 2 ** -8 + 2 ** -20)) == f(1.0 + 2 ** -7)
    assert bf16_rne(f(-3.0)) == -3.0
    # the 256 bf16 values of [1, 2) are fixed points
    vals = (1 + np.arange(128) / 128).astype(np.float32)
    np.testing.assert_array_equal(bf16_rne(vals), vals)


def test_e4m3_known_values():
    assert e4m3_rne(448.0) == 448.0
    assert e4m3_rne(-448.0) == -448.0
    assert e4m3_rne(1.0625) == 1.0        # tie between 1 and 1.125 -> even mantissa
    assert e4m3_rne(1.1875) == 1.25       # tie between 1.125 and 1.25 -> even
    assert e4m3_rne(2.0 ** -9) == 2.0 ** -9  # smallest subnormal
    assert e4m3_rne(2.0 ** -10) == 0.0    # tie with zero -> zero
    assert e4m3_rne(3 * 2.0 ** -10) == 2.0 ** -8  # tie between 2^-9 and 2^-8 -> even (2^-8)
    assert e4m3_rne(0.0) == 0.0
    # every e4m3 value is a fixed point, and there are 126 positive finite ones
    pos = [m * 2.0 ** -9 for m in range(1, 8)]
    pos += [(1 + mt / 8) * 2.0 ** e for e in range(-6, 9) for mt in range(8)]
    pos = [p for p in pos if p <= 448.0]
    assert len(pos) == 126
    np.testing.assert_array_equal(e4m3_rne(np.array(pos)), np.array(pos))
    # rounding is monotone and idempotent on a dense sweep
    xs = np.linspace(-448, 448, 100001)
    r = e4m3_rne(xs)
    assert (np.diff(r) >= 0).all()
    np.testing.assert_array_equal(e4m3_rne(r), r)


def test_rel_errors_closed_form():
    from numerics import rel_errors
    t = np.array([1.0, -2.0, 0.0, 3.0], np.float32)
    x = np.array([1.5, -2.0, 0.0, 2.0], np.float32)
    np.testing.assert_allclose(rel_errors(x, t), [1.5 / 6, np.sqrt(1.25 / 14), 1 / 3], rtol=1e-15)
    np.testing.assert_array_equal(rel_errors(t, t), [0.0, 0.0, 0.0])
