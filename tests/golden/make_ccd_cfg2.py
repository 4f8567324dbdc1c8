"""Writes ccd_cfg2_oracle.npz: the FP64 oracle's CCD++ (oracle.ccd_epoch:
canonical CCD++, SG/ccd.cu:532-554, one inner iteration) on the 100M-nonzero
benchmark tensor (480189 x 17770 x 2182, Zipf(1.0) modes 0-1, rank-10 CP
truth, integer values, 1% test, seed 0x5EED0002), R=32, lambda 0.05, from
factors U(-1,1) drawn by numpy.random.default_rng(1) in mode order and rounded
to float32 (the GPU's starting point).

Stored: the initial and per-epoch test RMSE, and after epochs 1 and the last,
64 seeded sample rows of every factor with their ids.
Usage: python make_ccd_cfg2.py [epochs]"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import cutc_amd as ct  # noqa: E402
import oracle as O  # noqa: E402


def main():
    epochs = int(sys.argv[1]) if len(sys.argv) > 1 else 4
    dims = (480189, 17770, 2182)
    tr, te = ct.synth(dims, 100_000_000, zipf=(1.0, 1.0, 0.0), true_rank=10, value_map=ct.VALUES_INT_1_5,
                      noise_std=0.3, test_fraction=0.01, seed=0x5EED0002)
    ind, vals = tr.arrays()
    oc = O.Coo(ind, vals, dims)
    del ind, vals
    ind, vals = te.arrays()
    ot = O.Coo(ind, vals, dims)
    rng = np.random.default_rng(1)
    F = [rng.uniform(-1, 1, (d, 32)).astype(np.float32).astype(np.float64) for d in dims]
    srng = np.random.default_rng(7)
    rows = [np.sort(srng.choice(d, 64, replace=False)).astype(np.int64) for d in dims]
    nth = os.cpu_count() or 8
    out = {"lam": 0.05, "rmse_te_initial": O.rmse(ot, F, nthreads=nth)}
    for m in range(3):
        out[f"rows{m}"] = rows[m]
    orders = [O.mode_order(oc, m) for m in range(3)]
    res = O.residual(oc, F)
    rm = []
    for e in range(1, epochs + 1):
        t0 = time.time()
        O.ccd_epoch(oc, F, 0.05, res, orders=orders, nthreads=nth)
        rm.append(O.rmse(ot, F, nthreads=nth))
        if e in (1, epochs):
            for m in range(3):
                out[f"f{m}_e{e}"] = F[m][rows[m]].copy()
        print(e, rm[-1], f"{time.time() - t0:.1f}s", flush=True)
    out["rmse_te"] = np.array(rm)
    np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ccd_cfg2_oracle.npz"), **out)


if __name__ == "__main__":
    main()
