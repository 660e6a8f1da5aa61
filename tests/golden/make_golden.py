"""Writes the golden fixtures used by tests/test_oracle.py.

Every expected output here comes from numpy / scipy (library routines), never
from this project's code, so the fixtures pin the oracle independently:
  gen_ops.npz   dense A x and A^T r in float64 (numpy matmul)
  smv_ops.npz   CSC products (scipy.sparse.csc_matrix)
  dct_ops.npz   orthonormal DCT-II rows / DCT-III of the zero fill (scipy.fft)
  lstsq.npz     least squares on a fixed support (numpy.linalg.lstsq)
Run: python tests/golden/make_golden.py   (rewrites the .npz files in place)
"""
import os

import numpy as np
import scipy.fft
import scipy.sparse

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20260101)
    # gen: column-major A stored as (n, m) rows = columns
    m, n = 37, 53
    A = rng.standard_normal((m, n)).astype(np.float32)
    x = rng.standard_normal(n)
    x[rng.random(n) < 0.6] = 0.0
    r = rng.standard_normal(m)
    np.savez(os.path.join(HERE, "gen_ops.npz"), A_cols=np.ascontiguousarray(A.T), x=x, r=r,
             Ax=A.astype(np.float64) @ x, ATr=A.astype(np.float64).T @ r)
    # smv: p distinct rows per column, +-1/sqrt(p)
    m, n, p = 40, 97, 5
    rows = np.concatenate([np.sort(rng.choice(m, p, replace=False)) for _ in range(n)]).astype(np.int32)
    vals = (rng.choice([-1.0, 1.0], n * p) / np.sqrt(p)).astype(np.float32)
    col_ptr = np.arange(0, n * p + 1, p, dtype=np.int32)
    S = scipy.sparse.csc_matrix((vals.astype(np.float64), rows, col_ptr), shape=(m, n))
    x = rng.standard_normal(n)
    r = rng.standard_normal(m)
    np.savez(os.path.join(HERE, "smv_ops.npz"), m=m, n=n, p=p, rows=rows, vals=vals, col_ptr=col_ptr,
             x=x, r=r, Ax=S @ x, ATr=S.T @ r)
    # dct: rows of the orthonormal DCT-II, adjoint = DCT-III of the zero fill
    out = {}
    for n, m in [(2, 1), (8, 3), (64, 20), (1024, 300)]:
        rows_d = np.sort(rng.choice(n, m, replace=False)).astype(np.int32)
        x = rng.standard_normal(n)
        r = rng.standard_normal(m)
        z = np.zeros(n)
        z[rows_d] = r
        out["rows_%d" % n] = rows_d
        out["x_%d" % n] = x
        out["r_%d" % n] = r
        out["Ax_%d" % n] = scipy.fft.dct(x, type=2, norm="ortho")[rows_d]
        out["ATr_%d" % n] = scipy.fft.idct(z, type=2, norm="ortho")
    np.savez(os.path.join(HERE, "dct_ops.npz"), **out)
    # least squares on a fixed support
    m, n, s = 60, 120, 9
    A = rng.standard_normal((m, n)).astype(np.float32) / np.sqrt(m)
    supp = np.sort(rng.choice(n, s, replace=False))
    y = rng.standard_normal(m).astype(np.float32)
    sol, *_ = np.linalg.lstsq(A[:, supp].astype(np.float64), y.astype(np.float64), rcond=None)
    np.savez(os.path.join(HERE, "lstsq.npz"), A_cols=np.ascontiguousarray(A.T), y=y, supp=supp, sol=sol)


if __name__ == "__main__":
    main()
