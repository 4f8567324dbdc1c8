"""Full-size parity of the loss adjoint on the 4-way benchmark tensor
(1M x 500K x 1000 x 24, 140M nonzeros, modes 0-1 Zipf(1.0), values from a
rank-16 CP model plus N(0, 0.2), 1% test; R = 32, lambda 0.05; factors
U(0, 0.6)): the gradient of every mode from one ctc_grad call against the FP64
oracle's rows of the same gradient from the same FP32 factors
(orc_grad_rows_f64), on sampled rows of each mode (rows of at most 200K
nonzeros in modes 0-2, two of the 24 rows of mode 3, whose rows hold about
5.8M nonzeros each)."""
import sys

import numpy as np
import pytest

import cutc_amd as ct
import oracle as O

pytestmark = pytest.mark.gpu
DIMS = (1_000_000, 500_000, 1000, 24)
R, LAM = 32, 0.05


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("per_occ", [False])
def test_adjoint_full_size_sampled_rows_match_oracle(per_occ):
    import torch
    say("[adjoint_full] generating the 140M-nonzero 4-way tensor")
    tr, _ = ct.synth(DIMS, 140_000_000, zipf=(1.0, 1.0, 0.0, 0.0), true_rank=16, value_map=ct.VALUES_RAW,
                     noise_std=0.2, test_fraction=0.01, seed=0x5EED0003)
    m = ct.Model(ct.Context(0), tr, None, None, rank=R)
    rng = np.random.default_rng(3)
    F = [(rng.random((d, R)) * 0.6).astype(np.float32) for d in DIMS]
    dev = torch.device("cuda:0")
    Ft = [torch.from_numpy(f).to(dev) for f in F]
    Gt = [torch.empty_like(f) for f in Ft]
    torch.cuda.synchronize()
    m.grad_device([f.data_ptr() for f in Ft], [g.data_ptr() for g in Gt], LAM, per_occ)
    torch.cuda.synchronize()
    G = [g.cpu().numpy() for g in Gt]
    del Ft, Gt
    ind, vals = tr.arrays()
    oc = O.Coo(ind, vals, tr.dims)
    del ind, vals, tr
    worst = {}
    for mode in range(4):
        order = O.mode_order(oc, mode)
        cnt = np.diff(order[1].astype(np.int64))
        if mode < 3:
            cand = np.flatnonzero((cnt > 0) & (cnt <= 200_000))
            rows = np.sort(rng.choice(cand, size=min(256, len(cand)), replace=False)).astype(np.uint32)
        else:
            rows = np.array([0, 23], np.uint32)
        ref = O.grad_rows(oc, mode, F, LAM, rows, per_occ, order=order, nthreads=16)
        got = G[mode][rows].astype(np.float64)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        worst[mode] = err
        say(f"  mode {mode}: {len(rows)} rows ({int(cnt[rows].sum())} nonzeros), rel err {err:.2e}")
    assert all(e < 1e-5 for e in worst.values()), worst
