"""Writes c2d_cgiht_dct.json: the FP64 oracle's CGIHT outcome on the full-size dct
configuration (n = 2^20, m = 2^17, k = 13108, N(0,1) nonzeros, tol 1e-5, restartFlag
off), instances 0-3 of seed 20260101.  The oracle takes about 20 s per instance here,
too slow for a GPU test, so the GPU test compares against these numbers.  Run from the
repository root: python tests/golden/make_c2d_cgiht.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

SEED = 20260101
K, M, N = 13108, 2 ** 17, 2 ** 20


def main():
    out = []
    for inst in range(4):
        P = O.make_problem("dct", M, N, SEED, inst)
        x = O.sparse_vector(N, K, "gaussian", SEED, inst)
        y = P.apply_A(x).astype(np.float32)
        xo, res, _, _ = O.solve("CGIHT", P, y, K, tol=1e-5, precision=64)
        out.append(dict(instance=inst, iterations=int(res.iters), stop_reason=int(res.stop_reason),
                        rel_linf=float(np.max(np.abs(xo - x)) / np.max(np.abs(x))),
                        rel_l2=float(np.linalg.norm(xo - x) / np.linalg.norm(x))))
        print(out[-1], flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2d_cgiht_dct.json")
    with open(path, "w") as f:
        json.dump(dict(alg="CGIHT", ens="dct", k=K, m=M, n=N, seed=SEED, tol=1e-5,
                       vecDistribution="gaussian", precision=64, instances=out), f, indent=1)


if __name__ == "__main__":
    main()
