"""Writes sgd_cfg2_sync_oracle.npz: the FP64 oracle of libctc's default SGD
schedule (oracle.sgd_sync_epoch: bulk-synchronous walkers of `span` nonzeros,
sub-steps of `sub`, hot rows above `hot` nonzeros and 4x their mode's mean
averaged, cold rows summed) on the 100M-nonzero benchmark tensor
(480189 x 17770 x 2182, Zipf(1.0) modes 0-1, rank-10 CP truth, integer values,
1% test, seed 0x5EED0002), R=32, lambda 0.05, lr 1e-3 x 0.95 per epoch,
scale 1, factors U(-1,1) from numpy.random.default_rng(1) in mode order.

Stored: the initial and per-epoch test RMSE, and after epochs 1, 2, 5 and the
last, 64 seeded sample rows of every factor with their ids.  About 10 s per
epoch on 8 cores.  A fifth argument "mixed" writes sgd_cfg2_sync_mixed_oracle.npz
instead: the oracle of the mixed-precision variant (run deltas of modes 1..
rounded to BF16 before they are summed).
Usage: python make_sgd_cfg2_sync.py [epochs] [span] [sub] [hot] [mixed]"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import cutc_amd as ct  # noqa: E402
import oracle as O  # noqa: E402


def main():
    epochs = int(sys.argv[1]) if len(sys.argv) > 1 else 20
    span = int(sys.argv[2]) if len(sys.argv) > 2 else 2048
    sub = int(sys.argv[3]) if len(sys.argv) > 3 else 32
    hot = int(sys.argv[4]) if len(sys.argv) > 4 else 65536
    mixed = len(sys.argv) > 5 and sys.argv[5] == "mixed"
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
    srng = np.random.default_rng(7)
    rows = [np.sort(srng.choice(d, 64, replace=False)).astype(np.int64) for d in dims]
    out = {"span": span, "sub": sub, "hot": hot, "mixed": int(mixed), "rmse_te_initial": O.rmse(ot, F, nthreads=8)}
    for m in range(3):
        out[f"rows{m}"] = rows[m]
    rm = []
    order = np.arange(tr.nnz, dtype=np.uint32)  # the training tensor is in mode-0 order
    lr = 1e-3
    nth = os.cpu_count() or 8
    for e in range(1, epochs + 1):
        t0 = time.time()
        O.sgd_sync_epoch(oc, F, lr, 0.05, span, sub, hot, order=order, nthreads=nth, mixed=mixed)
        lr *= 0.95
        rm.append(O.rmse(ot, F, nthreads=nth))
        if e in (1, 2, 5, epochs):
            for m in range(3):
                out[f"f{m}_e{e}"] = F[m][rows[m]].copy()
        print(e, rm[-1], f"{time.time() - t0:.1f}s", flush=True)
    out["rmse_te"] = np.array(rm)
    name = "sgd_cfg2_sync_mixed_oracle.npz" if mixed else "sgd_cfg2_sync_oracle.npz"
    np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), name), **out)


if __name__ == "__main__":
    main()
