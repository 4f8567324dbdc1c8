"""Full-size parity on the 100M-nonzero benchmark tensor (SGD, CCD++, and
sampled rows of an ALS sweep, each test's docstring).

SGD: the default bulk-synchronous schedule on the
100M-nonzero benchmark tensor (480189 x 17770 x 2182, Zipf(1.0) modes 0-1,
integer values, 1% test; R = 32, lambda 0.05, lr 1e-3 x 0.95 per epoch,
scale 1, factors U(-1,1) from numpy.random.default_rng(1)) against the FP64
oracle's run of the same schedule, stored by
tests/golden/make_sgd_cfg2_sync.py in sgd_cfg2_sync_oracle.npz.

Bars: per-epoch test RMSE within 1e-3 for all 20 epochs; 64 sampled rows of
every factor within 1e-5 relative (Frobenius) after epochs 1, 2, 5 and 20
(FP32 rounding on the GPU against FP64 in the oracle, compounded over the
sweeps: measured 1e-6 after epoch 20).

The mixed-precision path (run deltas of modes 1.. staged in BF16) runs the
same 20 epochs against the oracle's mixed variant
(sgd_cfg2_sync_mixed_oracle.npz, same script with "mixed"): sampled rows
within 5e-4 after epoch 1, and per-epoch test RMSE within 1e-3 of that oracle
and of the FP64 run of the FP32 schedule (the precision target).  A delta
summed in FP32 and in FP64 can fall on different sides of a BF16 rounding
midpoint, which moves it by a whole BF16 step; a mode-2 row takes ~45000
rounded deltas per epoch, so its spread is the largest: the oracle's own FP32
twin of the mixed variant differs from its FP64 run by 1.4e-4 on mode 2 (and
3e-6 on mode 1) after epoch 1, as the GPU does.  Later epochs carry those
moves on and are reported."""
import os
import sys

import numpy as np
import pytest

import cutc_amd as ct

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "sgd_cfg2_sync_oracle.npz")
GOLD_MIXED = os.path.join(os.path.dirname(__file__), "golden", "sgd_cfg2_sync_mixed_oracle.npz")
GOLD_CCD = os.path.join(os.path.dirname(__file__), "golden", "ccd_cfg2_oracle.npz")
DIMS = (480189, 17770, 2182)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def full_tensor():
    say("[sgd_full] generating the 100M-nonzero tensor")
    tr, te = ct.synth(DIMS, 100_000_000, zipf=(1.0, 1.0, 0.0), true_rank=10, value_map=ct.VALUES_INT_1_5,
                      noise_std=0.3, test_fraction=0.01, seed=0x5EED0002)
    yield tr, te
    del tr, te


@pytest.fixture(scope="module")
def full_model(full_tensor):
    tr, te = full_tensor
    m = ct.Model(ct.Context(0), tr, None, te, rank=32)
    yield m
    del m


def run_20(m, gold, precision, tol_rows):
    """20 epochs from the fixture's init; tol_rows: {epoch: bar} for the
    sampled rows (epochs 1, 2, 5 and 20 are compared and printed)."""
    rng = np.random.default_rng(1)
    m.set_factors([rng.uniform(-1, 1, (d, 32)).astype(np.float32) for d in DIMS])
    opt = ct.Options(algorithm=ct.SGD, rank=32, lambda_=0.05, sgd_scale=1.0, precision=precision)
    c = opt.to_c()
    assert (c.tile_len, c.sgd_substep, c.sgd_hot_rows) == (int(gold["span"]), int(gold["sub"]), int(gold["hot"]))
    assert abs(m.evaluate()["rmse_te"] - float(gold["rmse_te_initial"])) < 1e-6
    lr, rm = 1e-3, []
    for e in range(1, 21):
        m.sgd_epoch(opt, lr=lr)
        lr *= 0.95
        rm.append(m.evaluate()["rmse_te"])
        say(f"  epoch {e}: gpu {rm[-1]:.6f} oracle {gold['rmse_te'][e - 1]:.6f}")
        if e in (1, 2, 5, 20):
            for k in range(3):
                rows = gold[f"rows{k}"]
                got = m.factor(k)[rows].astype(np.float64)
                ref = gold[f"f{k}_e{e}"]
                err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
                say(f"    factor {k} sampled rows rel err {err:.2e}")
                if e in tol_rows:
                    assert err < tol_rows[e], (e, k, err)
    return np.array(rm)


@pytest.mark.timeout(600)
def test_sgd_default_schedule_full_size_matches_oracle(full_model):
    gold = np.load(GOLD, allow_pickle=False)
    rm = run_20(full_model, gold, ct.PREC_FP32, {1: 1e-5, 2: 1e-5, 5: 1e-5, 20: 1e-5})
    gap = np.max(np.abs(rm - gold["rmse_te"]))
    say(f"  max test-RMSE gap {gap:.2e}")
    assert gap < 1e-3


@pytest.mark.timeout(600)
def test_sgd_mixed_full_size_matches_oracles(full_model):
    gold = np.load(GOLD_MIXED, allow_pickle=False)
    fp64 = np.load(GOLD, allow_pickle=False)
    assert int(gold["mixed"]) == 1
    rm = run_20(full_model, gold, ct.PREC_MIXED, {1: 5e-4})
    gap_mixed = np.max(np.abs(rm - gold["rmse_te"]))
    gap_fp64 = np.max(np.abs(rm - fp64["rmse_te"]))
    say(f"  max test-RMSE gap: {gap_mixed:.2e} to the mixed oracle, {gap_fp64:.2e} to the FP64 oracle")
    assert gap_mixed < 1e-3
    assert gap_fp64 < 1e-3


@pytest.mark.timeout(600)
def test_ccd_full_size_matches_oracle(full_model):
    """CCD++ (32 columns x 3 modes per epoch, one inner iteration, lambda 0.05)
    on the same tensor from the same float32-rounded init against the FP64
    oracle (ccd_cfg2_oracle.npz from tests/golden/make_ccd_cfg2.py): 64 sampled
    rows of every factor after epoch 1 and the last, and the per-epoch test
    RMSE.  The GPU keeps FP32 factors and residuals with FP64 row sums."""
    gold = np.load(GOLD_CCD, allow_pickle=False)
    m = full_model
    rng = np.random.default_rng(1)
    m.set_factors([rng.uniform(-1, 1, (d, 32)).astype(np.float32) for d in DIMS])
    assert abs(m.evaluate()["rmse_te"] - float(gold["rmse_te_initial"])) < 1e-6
    opt = ct.Options(algorithm=ct.CCD, rank=32, lambda_=float(gold["lam"]))
    E = len(gold["rmse_te"])
    rm = []
    for e in range(1, E + 1):
        m.ccd_epoch(opt)
        rm.append(m.evaluate()["rmse_te"])
        say(f"  epoch {e}: gpu {rm[-1]:.6f} oracle {gold['rmse_te'][e - 1]:.6f}")
        if e in (1, E):
            for k in range(3):
                got = m.factor(k)[gold[f"rows{k}"]].astype(np.float64)
                ref = gold[f"f{k}_e{e}"]
                err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
                say(f"    factor {k} sampled rows rel err {err:.2e}")
                assert err < 1e-4, (e, k, err)
    gap = np.max(np.abs(np.array(rm) - gold["rmse_te"]))
    say(f"  max test-RMSE gap {gap:.2e}")
    assert gap < 1e-4


def test_als_full_size_sampled_rows_match_oracle(full_model, full_tensor):
    """One ALS sweep (modes 0, 1, 2; the span kernels, R = 32, lambda 0.05)
    on the full 100M-nonzero tensor from U(-1,1) factors: before each mode
    update, the FP64 oracle solves 256 sampled rows of the mode (rows of at
    most 100K nonzeros, so that the CPU finishes in seconds) from the same
    FP32 factors (orc_als_rows_f64); the GPU's rows within 1e-5 relative."""
    import oracle as O
    tr, _ = full_tensor
    ind, vals = tr.arrays()
    oc = O.Coo(ind, vals, tr.dims)
    del ind, vals
    m = full_model
    rng = np.random.default_rng(7)
    m.set_factors([rng.uniform(-1, 1, (d, 32)).astype(np.float32) for d in DIMS])
    opt = ct.Options(algorithm=ct.ALS, rank=32, lambda_=0.05)
    for mode in range(3):
        order = O.mode_order(oc, mode)
        cnt = np.diff(order[1].astype(np.int64))
        cand = np.flatnonzero((cnt > 0) & (cnt <= 100_000))
        rows = np.sort(rng.choice(cand, size=min(256, len(cand)), replace=False)).astype(np.uint32)
        F = m.factors()
        ref = O.als_rows(oc, mode, F, 0.05, rows, order=order, nthreads=16)
        m.als_mode_update(mode, opt)
        got = m.factor(mode)[rows].astype(np.float64)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        say(f"  ALS mode {mode}: {len(rows)} rows ({int(cnt[rows].sum())} nonzeros), rel err {err:.2e}")
        assert err < 1e-5, (mode, err)
