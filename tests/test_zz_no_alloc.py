"""ctc_grad_noalloc (the allocation-free gradient) against ctc_grad through
the C ABI: two calls in a row on one GPU, the second on a smaller tensor, each
compared with ctc_grad on the same factors (gradients within 1e-5 relative per
factor: both add float atomics in an order that varies run to run; the loss
within 1e-9 relative: FP64 partial sums)."""
import sys

import numpy as np
import pytest

import cutc_amd as ct

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(ctx, dims, nnz, R, seed, per_occ):
    import torch
    tr, _ = ct.synth(dims, nnz, zipf=tuple([1.0, 0.7] + [0.0] * (len(dims) - 2)), true_rank=4,
                     value_map=ct.VALUES_AFFINE_1_5, noise_std=0.1, test_fraction=0.0, seed=seed)
    m = ct.Model(ctx, tr, None, None, rank=R)
    m.prepare(ct.Options(algorithm=ct.ALS, rank=R, lambda_=0.1))  # builds every mode's CISS
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    F = [torch.from_numpy(rng.uniform(-1, 1, (d, R)).astype(np.float32)).to(dev) for d in tr.dims]
    G1 = [torch.full_like(f, 7.0) for f in F]  # stale values the call must overwrite
    G2 = [torch.empty_like(f) for f in F]
    loss_dev = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    m.grad_device_noalloc([f.data_ptr() for f in F], [g.data_ptr() for g in G1], 0.3, loss_dev.data_ptr(),
                          per_occ)
    torch.cuda.synchronize()
    loss_ref = m.grad_device([f.data_ptr() for f in F], [g.data_ptr() for g in G2], 0.3, per_occ)
    torch.cuda.synchronize()
    lo = float(loss_dev.cpu()[0])
    say(f"  dims={dims} nnz={tr.nnz} R={R} per_occ={per_occ}: loss {lo:.10g} vs {loss_ref:.10g}")
    assert lo == pytest.approx(loss_ref, rel=1e-9)
    for a, b in zip(G1, G2):
        a, b = a.cpu().numpy().astype(np.float64), b.cpu().numpy().astype(np.float64)
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert err < 1e-5, err


@pytest.mark.parametrize("per_occ", [False, True])
def test_grad_noalloc_matches_grad_twice(per_occ):
    ctx = ct.Context(0)
    say(f"[grad_noalloc] per_occ={per_occ}")
    _case(ctx, (3000, 1500, 200, 12), 400_000, 32, 5, per_occ)  # larger first
    _case(ctx, (300, 200, 60), 20_000, 16, 6, per_occ)          # then smaller: stale scratch must not leak


def test_grad_noalloc_requires_prepared_model():
    ctx = ct.Context(0)
    tr, _ = ct.synth((100, 80, 30), 2000, test_fraction=0.0, seed=1)
    m = ct.Model(ctx, tr, None, None, rank=8)
    with pytest.raises(ct.CtcError) as e:
        m.grad_device_noalloc([1, 1, 1], [1, 1, 1], 0.1)  # rejected before any launch
    assert e.value.status == 1
