"""Train the CGIHT warm-start network on the GPU and save its weights.

    python -m gaga_amd.train_warm_start [--total 200000] [--epochs 4] [--out PATH]

Generates `total` CGIHT recoveries over the (delta, rho) grid at n = 4096
(many-RHS solves, one A per grid point), trains the network with the
library's kernels, saves the parameters (numpy .npz, plain arrays) and prints
one JSON line with the data and training statistics and a short evaluation
against CGIHT without the network.
"""
import argparse
import json
import sys
import time

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--total", type=int, default=200_000)
    ap.add_argument("--per", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=None)
    ap.add_argument("--eval-instances", type=int, default=8)
    args = ap.parse_args()
    import gaga_amd as ga
    from gaga_amd import learned
    t0 = time.time()
    X, y, stats = learned.make_training_set(total=args.total, per=args.per, log=log)
    t_data = time.time() - t0
    log("training set: %d samples from %d successful of %d recoveries (%.1f s)"
        % (len(y), stats["successful"], stats["recoveries"], t_data))
    net = ga.MLP(seed=1)
    t0 = time.time()
    losses, pos_w = learned.train(net, X, y, epochs=args.epochs, batch=args.batch, lr=args.lr,
                                  log=log)
    t_train = time.time() - t0
    out = args.out or learned.WEIGHTS
    learned.save(net, out, losses=losses, pos_w=pos_w, samples=len(y),
                 recoveries=stats["recoveries"], successful=stats["successful"])
    res = {}
    for ens, m, n, k in (("gen", 1024, 4096, 205), ("dct", 16384, 65536, 3277)):
        rows = learned.evaluate(net, ens, m, n, k, args.eval_instances)
        agg = {}
        for key in ("plain", "magnitude", "warm"):
            agg[key] = dict(ms=float(np.mean([r[key][0] for r in rows])),
                            iterations=float(np.mean([r[key][1] for r in rows])),
                            recovered=int(sum(r[key][2] for r in rows)))
        agg["fallbacks"] = int(sum(r["fallback"] for r in rows))
        res["%s_%dx%d_k%d" % (ens, m, n, k)] = agg
        log(ens, json.dumps(agg))
    print(json.dumps(dict(stats, data_s=t_data, train_s=t_train, losses=losses, weights=out,
                          evaluation=res)))


if __name__ == "__main__":
    main()
