"""Reference roundings for the low-precision formats (test helpers)."""
import numpy as np


def bf16_rne(a):
    """float32 -> bfloat16 (round to nearest even) -> float32."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def e4m3_rne(v):
    """Round values with |v| <= 448 to OCP FP8 e4m3 (3 mantissa bits, exponent
    bias 7, subnormal quantum 2^-9), nearest even; returns float64."""
    v = np.asarray(v, np.float64)
    a = np.abs(v)
    e = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    e = np.maximum(e, -6)
    q = 2.0 ** (e - 3)
    r = np.round(a / q) * q  # numpy rounds half to even
    return np.sign(v) * np.minimum(r, 448.0)


def rel_errors(x, x_true):
    """Relative l1, l2 and l_inf errors of x against x_true in float64, the
    closed forms of the reference's results (results.cu:363-378)."""
    t = np.asarray(x_true, np.float64)
    d = t - np.asarray(x, np.float64)
    return np.array([np.abs(d).sum() / np.abs(t).sum(), np.linalg.norm(d) / np.linalg.norm(t),
                     np.abs(d).max() / np.abs(t).max()])
