"""BF16-operand ALS (mixed precision) against the oracle.

The oracle's `als_mode_bf16` restates the path's arithmetic: BF16-rounded
factor rows, BF16-rounded h in the Gram and in m, then an FP64 solve; it accumulates
G and m in FP64 where the GPU accumulates in FP32 (MFMA), so the bar is 1e-4
relative per factor (FP32 accumulation of up to thousands of terms, amplified
by the condition number of G + lambda I).  The distance to the plain FP64
update (no BF16 rounding) is printed as the precision cost.
"""
import sys

import numpy as np
import pytest

import cutc_amd as ct
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return ct.Context(0)


def relerr(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))


def problem(dims, nnz, seed, nmodes_zipf=True):
    zipf = tuple([1.0, 0.8] + [0.0] * (len(dims) - 2)) if nmodes_zipf else None
    return ct.synth(dims, nnz, zipf=zipf, true_rank=6, value_map=ct.VALUES_AFFINE_1_5, noise_std=0.1,
                    test_fraction=0.1, seed=seed)


@pytest.mark.parametrize("R,tile_len", [(32, 0), (64, 0), (128, 0), (32, 64), (128, 96)])
def test_mixed_mode_update_matches_bf16_oracle(ctx, R, tile_len):
    tr, te = problem((400, 300, 50), 40000, seed=50 + R + tile_len)
    say(f"[mixed] R={R} tile_len={tile_len} (rows above it are cut into chunks)")
    m = ct.Model(ctx, tr, None, te, rank=R)
    rng = np.random.default_rng(R)
    F = [rng.uniform(-1, 1, (d, R)).astype(np.float32) for d in tr.dims]
    m.set_factors(F)
    ind, vals = tr.arrays()
    oc = O.Coo(ind, vals, tr.dims)
    opt = ct.Options(algorithm=ct.ALS, rank=R, lambda_=0.05, precision=ct.PREC_MIXED,
                     tile_len=tile_len or None)
    Fo = [f.copy() for f in F]
    for mode in range(3):
        F64 = [f.astype(np.float64) for f in Fo]
        m.als_mode_update(mode, opt)
        assert O.als_mode_bf16(oc, mode, Fo, 0.05) == 0
        assert O.als_mode(oc, mode, F64, 0.05) == 0
        got = m.factor(mode)
        e = relerr(got, Fo[mode].astype(np.float64))
        e64 = relerr(got, F64[mode])
        say(f"  mode {mode}: rel err vs BF16-operand oracle {e:.2e}, vs FP64 oracle {e64:.2e}")
        assert e < 1e-4, (mode, e)
        # continue from identical inputs
        m.set_factors(Fo)


@pytest.mark.parametrize("R,prec", [(32, "mixed"), (128, "mixed"), (32, "fp32"), (128, "fp32")])
def test_long_rows_reduced_in_groups(ctx, R, prec):
    """Rows cut into more than 64 chunks have their chunk partials summed in
    groups of 32 before the solve; mode 2 here has 4 rows of ~9000 nonzeros,
    ~280 chunks of 32 each."""
    tr, te = problem((2000, 1500, 4), 40000, seed=70 + R)
    say(f"[long rows] R={R} {prec}")
    m = ct.Model(ctx, tr, None, te, rank=R)
    rng = np.random.default_rng(R + 1)
    F = [rng.uniform(-1, 1, (d, R)).astype(np.float32) for d in tr.dims]
    m.set_factors(F)
    ind, vals = tr.arrays()
    oc = O.Coo(ind, vals, tr.dims)
    if prec == "mixed":
        opt = ct.Options(algorithm=ct.ALS, rank=R, lambda_=0.05, precision=ct.PREC_MIXED, tile_len=32)
        Fo = [f.copy() for f in F]
        assert O.als_mode_bf16(oc, 2, Fo, 0.05) == 0
        tol = 1e-4
    else:
        opt = ct.Options(algorithm=ct.ALS, rank=R, lambda_=0.05, als_kernel=ct.ALS_ROWS, tile_len=32)
        Fo = [f.astype(np.float64) for f in F]
        assert O.als_mode(oc, 2, Fo, 0.05) == 0
        tol = 1e-5
    m.als_mode_update(2, opt)
    e = relerr(m.factor(2), Fo[2].astype(np.float64))
    say(f"  mode 2 rel err {e:.2e}")
    assert e < tol, e


def test_mixed_four_modes(ctx):
    tr, te = problem((120, 90, 30, 12), 30000, seed=61)
    say("[mixed 4-way] R=32")
    R = 32
    m = ct.Model(ctx, tr, None, te, rank=R)
    rng = np.random.default_rng(3)
    F = [rng.uniform(-1, 1, (d, R)).astype(np.float32) for d in tr.dims]
    m.set_factors(F)
    ind, vals = tr.arrays()
    oc = O.Coo(ind, vals, tr.dims)
    opt = ct.Options(algorithm=ct.ALS, rank=R, lambda_=0.05, precision=ct.PREC_MIXED, tile_len=128)
    for mode in range(4):
        m.als_mode_update(mode, opt)
        assert O.als_mode_bf16(oc, mode, F, 0.05) == 0
        e = relerr(m.factor(mode), F[mode].astype(np.float64))
        say(f"  mode {mode}: {e:.2e}")
        assert e < 1e-4
        m.set_factors(F)


def test_mixed_als_converges_like_fp64(ctx):
    """Six full sweeps at R=32.  Every sweep's test RMSE within 2e-4 of the
    BF16-operand oracle's; the last within 1e-3 of the FP64 oracle's.  (The
    BF16 rounding itself moves sweep 1 by 2.3e-3 on this tensor while the RMSE
    is still near 0.9; from sweep 2 on the gap stays below 2e-4.)"""
    tr, te = problem((500, 400, 60), 60000, seed=70)
    R = 32
    m = ct.Model(ctx, tr, None, te, rank=R)
    rng = np.random.default_rng(9)
    F = [rng.uniform(-1, 1, (d, R)).astype(np.float32) for d in tr.dims]
    m.set_factors(F)
    ind, vals = tr.arrays()
    oc = O.Coo(ind, vals, tr.dims)
    ind, vals = te.arrays()
    ot = O.Coo(ind, vals, te.dims)
    F64 = [f.astype(np.float64) for f in F]
    Fb = [f.copy() for f in F]
    opt = ct.Options(algorithm=ct.ALS, rank=R, lambda_=0.5, precision=ct.PREC_MIXED)
    for it in range(6):
        for mode in range(3):
            m.als_mode_update(mode, opt)
            assert O.als_mode(oc, mode, F64, 0.5) == 0
            assert O.als_mode_bf16(oc, mode, Fb, 0.5) == 0
        g = m.evaluate()["rmse_te"]
        o64, ob = O.rmse(ot, F64), O.rmse(ot, [f.astype(np.float64) for f in Fb])
        say(f"  sweep {it}: rmse_te mixed {g:.6f} bf16-oracle {ob:.6f} fp64 {o64:.6f}")
        assert abs(g - ob) < 2e-4
    assert abs(g - o64) < 1e-3


def test_mixed_rejects_unsupported_rank(ctx):
    tr, te = problem((50, 40, 30), 3000, seed=5)
    m = ct.Model(ctx, tr, None, te, rank=16)
    with pytest.raises(ct.CtcError):
        m.als_mode_update(0, ct.Options(algorithm=ct.ALS, rank=16, precision=ct.PREC_MIXED))
