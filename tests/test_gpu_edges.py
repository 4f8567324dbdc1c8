"""Degenerate and ragged inputs through every kernel path, against the oracle:
a single nonzero, one tile that is partly filled, exactly full or one over,
a mode with a single row (every nonzero in one long row), rows without
nonzeros in every mode, and a small 4-way tensor.  Same bars as the parity
tests at larger sizes: MTTKRP, ALS (span and row kernels) and CCD++ within
1e-5 of the FP64 oracle, the BF16-operand ALS within 1e-4 of its oracle, the
default SGD schedule within 1e-5 (1e-4 on its mixed path), loss and RMSE
within 1e-6 relative."""
import sys

import numpy as np
import pytest

import cutc_amd as ct
import oracle as O

pytestmark = pytest.mark.gpu

CASES = {  # dims, nonzeros
    "one_nonzero": ((4, 3, 2), 1),
    "ragged_tile": ((9, 7, 5), 37),
    "full_tile": ((9, 7, 5), 64),
    "tile_plus_one": ((9, 7, 5), 65),
    "single_row_mode0": ((1, 40, 30), 500),
    "four_way": ((5, 4, 3, 2), 50),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return ct.Context(0)


def tensor(name):
    dims, nnz = CASES[name]
    rng = np.random.default_rng(len(name) * 7919 + nnz)
    flat = rng.choice(int(np.prod(dims)), size=nnz, replace=False)
    ind = [a.astype(np.uint32) for a in np.unravel_index(flat, dims)]
    vals = rng.uniform(1.0, 5.0, nnz)
    t = ct.Coo.from_arrays(ind, vals, dims)
    return t, O.Coo(ind, vals, dims)


def factors(dims, R, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-1, 1, (d, R)).astype(np.float32) for d in dims]


def relerr(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("name", list(CASES))
def test_edges_mttkrp_and_evaluate(ctx, name):
    t, oc = tensor(name)
    say(f"[edges mttkrp] {name} dims={t.dims} nnz={t.nnz}")
    R = 16
    m = ct.Model(ctx, t, t, t, rank=R)
    F = factors(t.dims, R, 1)
    m.set_factors(F)
    F64 = [f.astype(np.float64) for f in F]
    for mode in range(t.nmodes):
        e = relerr(m.mttkrp(mode), O.mttkrp(oc, mode, F64))
        assert e < 1e-5, (mode, e)
    ev = m.evaluate(lam=0.5)
    assert ev["loss"] == pytest.approx(O.loss_sq(oc, F64), rel=1e-6)
    assert ev["rmse_te"] == pytest.approx(O.rmse(oc, F64), rel=1e-6)
    assert ev["frob"] == pytest.approx(O.frob_sq(t.dims, F64, 0.5), rel=1e-6)


ALS_KINDS = [(8, ct.ALS_SPANS, ct.PREC_FP32), (16, ct.ALS_SPANS, ct.PREC_FP32), (32, ct.ALS_ROWS, ct.PREC_FP32),
             (32, ct.ALS_ROWS, ct.PREC_MIXED)]


@pytest.mark.parametrize("R,kernel,prec", ALS_KINDS, ids=["spans8", "spans16", "rows32", "mixed32"])
@pytest.mark.parametrize("name", list(CASES))
def test_edges_als(ctx, name, R, kernel, prec):
    """Every mode once; rows without nonzeros become 0 on both sides."""
    t, oc = tensor(name)
    say(f"[edges als] {name} R={R} kernel={kernel} prec={prec}")
    m = ct.Model(ctx, t, None, None, rank=R)
    F = factors(t.dims, R, 2)
    m.set_factors(F)
    mixed = prec == ct.PREC_MIXED
    opt = ct.Options(algorithm=ct.ALS, rank=R, lambda_=0.05, tile_len=64, als_kernel=kernel, precision=prec)
    for mode in range(t.nmodes):
        m.als_mode_update(mode, opt)
        if mixed:
            assert O.als_mode_bf16(oc, mode, F, 0.05) == 0
            want = F[mode].astype(np.float64)
        else:
            F64 = [f.astype(np.float64) for f in F]
            assert O.als_mode(oc, mode, F64, 0.05) == 0
            want = F64[mode]
            F[mode] = F64[mode].astype(np.float32)
        got = m.factor(mode)
        e = relerr(got, want)
        assert e < (1e-4 if mixed else 1e-5), (mode, e)
        m.set_factors(F)


@pytest.mark.parametrize("name", list(CASES))
def test_edges_ccd(ctx, name):
    t, oc = tensor(name)
    say(f"[edges ccd] {name}")
    R = 8
    m = ct.Model(ctx, t, None, None, rank=R)
    F = factors(t.dims, R, 3)
    m.set_factors(F)
    F64 = [f.astype(np.float64) for f in F]
    res = O.residual(oc, F64)
    orders = [O.mode_order(oc, md) for md in range(t.nmodes)]
    opt = ct.Options(algorithm=ct.CCD, rank=R, lambda_=0.05, tile_len=64)
    for _ in range(2):
        m.ccd_epoch(opt)
        O.ccd_epoch(oc, F64, 0.05, res, orders=orders)
        errs = [relerr(m.factor(md), F64[md]) for md in range(t.nmodes)]
        assert max(errs) < 1e-5, errs


@pytest.mark.parametrize("mixed", [False, True], ids=["fp32", "mixed"])
@pytest.mark.parametrize("name", list(CASES))
def test_edges_sgd_sync(ctx, name, mixed):
    """Walkers of 16 nonzeros in sub-steps of 4: fewer walkers than one wave
    holds, a last walker shorter than its span, and (single_row_mode0) a
    mode-0 row shared by every walker, summed in fixed point."""
    t, oc = tensor(name)
    say(f"[edges sgd] {name} mixed={mixed}")
    R = 16
    m = ct.Model(ctx, t, None, None, rank=R)
    F = factors(t.dims, R, 4)
    m.set_factors(F)
    F64 = [f.astype(np.float64) for f in F]
    order, _ = O.mode_order(oc, 0)
    opt = ct.Options(algorithm=ct.SGD, rank=R, lambda_=0.05, sgd_scale=1.0, tile_len=16, sgd_substep=4,
                     sgd_hot_rows=8, precision=ct.PREC_MIXED if mixed else ct.PREC_FP32)
    for _ in range(2):
        m.sgd_epoch(opt, lr=1e-2)
        O.sgd_sync_epoch(oc, F64, 1e-2, 0.05, 16, 4, 8, order=order, mixed=mixed)
        errs = [relerr(m.factor(k), F64[k]) for k in range(t.nmodes)]
        assert max(errs) < (1e-4 if mixed else 1e-5), errs
