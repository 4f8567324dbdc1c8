"""Writes sgd_cfg2_oracle.json: the FP64 sequential SGD oracle's test-RMSE per
epoch on the 100M-nonzero benchmark tensor (480189 x 17770 x 2182, Zipf(1.0)
modes 0-1, rank-10 CP truth, integer values, 1% test, seed 0x5EED0002), R=32,
lambda 0.05, lr 1e-3 x 0.95 per epoch, scale 1, factors U(-1,1) from
numpy.random.default_rng(1) in mode order, nonzeros visited in mode-0 order.
Takes about 15 s per epoch on one core.  Usage: python make_sgd_cfg2_trajectory.py [epochs]"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import cutc_amd as ct  # noqa: E402
import oracle as O  # noqa: E402


def main():
    epochs = int(sys.argv[1]) if len(sys.argv) > 1 else 20
    dims = (480189, 17770, 2182)
    tr, te = ct.synth(dims, 100_000_000, zipf=(1.0, 1.0, 0.0), true_rank=10, value_map=ct.VALUES_INT_1_5,
                      noise_std=0.3, test_fraction=0.01, seed=0x5EED0002)
    ind, vals = tr.arrays()
    oc = O.Coo(ind, vals, dims)
    del ind, vals
    ind, vals = te.arrays()
    ot = O.Coo(ind, vals, dims)
    rng = np.random.default_rng(1)
    F = [rng.uniform(-1, 1, (d, 32)) for d in dims]
    out = {"rmse_te_initial": O.rmse(ot, F), "rmse_te": []}
    order = np.arange(tr.nnz, dtype=np.uint32)  # the training tensor is in mode-0 order
    lr = 1e-3
    for e in range(epochs):
        O.sgd_epoch(oc, F, lr, 0.05, 1.0, order=order)
        lr *= 0.95
        out["rmse_te"].append(O.rmse(ot, F))
        print(e + 1, out["rmse_te"][-1], flush=True)
    json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "sgd_cfg2_oracle.json"), "w"),
              indent=1)


if __name__ == "__main__":
    main()
