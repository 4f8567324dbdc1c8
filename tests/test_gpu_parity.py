"""GPU parity tests: the HIP kernels of libctc (through the C ABI) against the
CPU oracle on the same seeded inputs.

Tolerances (from the reference's update rules as the oracle restates them):
  * MTTKRP, Gram-based ALS updates, CCD++ and the gradient: relative Frobenius
    error vs the FP64 oracle, FP32 on the GPU -> <= 1e-5 per factor.
  * SGD, default bulk-synchronous schedule (oracle.sgd_sync_epoch restates it):
    relative factor error <= 1e-5 per factor after every epoch checked, and
    per-epoch test RMSE within 1e-3 (and within 5% of how far the oracle's own
    test RMSE moves).  Its mixed-precision path (BF16-staged run deltas)
    against the oracle's mixed variant: <= 1e-4 per factor after two epochs
    (a BF16 rounding moves by a step when FP32 and FP64 sums straddle a
    midpoint), and test RMSE within 1e-3 of that oracle and of the FP64 oracle
    of the FP32 schedule.
  * SGD, opt-in Hogwild policies vs the sequential oracle: per-epoch test RMSE
    within 1e-3 on a 150K-nonzero tensor.
  * Loss / RMSE: FP64 accumulation on both sides -> 1e-6 relative.
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


def ocoo(t):
    ind, vals = t.arrays()
    return O.Coo(ind, vals, t.dims)


def init_factors(dims, R, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-1, 1, (d, R)).astype(np.float32) for d in dims]


def relerr(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def small_problem(nmodes=3, nnz=20000, zipf=None, seed=1, test_fraction=0.1, dims=None):
    dims = dims or ((300, 200, 60) if nmodes == 3 else (120, 80, 30, 12))
    zipf = zipf or tuple([1.0, 0.7] + [0.0] * (nmodes - 2))
    return ct.synth(dims, nnz, zipf=zipf, true_rank=4, value_map=ct.VALUES_AFFINE_1_5,
                    noise_std=0.1, test_fraction=test_fraction, seed=seed)


# --------------------------------------------------------------------------
@pytest.mark.parametrize("R", [8, 16, 32, 64, 128])
@pytest.mark.parametrize("nmodes", [3, 4])
def test_mttkrp(ctx, R, nmodes):
    tr, te = small_problem(nmodes, seed=R + nmodes)
    say(f"[mttkrp] R={R} N={nmodes} nnz={tr.nnz}")
    m = ct.Model(ctx, tr, None, te, rank=R)
    F = init_factors(tr.dims, R, 3)
    m.set_factors(F)
    oc = ocoo(tr)
    F64 = [f.astype(np.float64) for f in F]
    for mode in range(nmodes):
        got = m.mttkrp(mode)
        exp = O.mttkrp(oc, mode, F64)
        assert relerr(got, exp) < 1e-5, (mode, relerr(got, exp))


def test_sampled_profiling():
    """ctc_context_set_profiling(n > 1) times every n-th launch of each kernel
    name from the call on; 1 every launch; negative values are errors."""
    c = ct.Context(0)
    tr, _ = small_problem(3, seed=2)
    m = ct.Model(c, tr, None, None, rank=16)
    m.set_factors(init_factors(tr.dims, 16, 1))
    for every, want in ((3, 4), (1, 10), (4, 3)):
        c.set_profiling(True, every=every)
        c.reset_stats()
        for _ in range(10):
            m.mttkrp(0)
        n, ms = c.kernel_stats("mttkrp")
        assert n == want and ms > 0, (every, n, ms)
    c.set_profiling(False)
    with pytest.raises(ct.CtcError):
        c.set_profiling(True, every=-2)


@pytest.mark.parametrize("nmodes", [3, 4])
def test_evaluate(ctx, nmodes):
    tr, te = small_problem(nmodes, seed=9)
    say(f"[evaluate] N={nmodes}")
    m = ct.Model(ctx, tr, te, te, rank=16)
    F = init_factors(tr.dims, 16, 4)
    m.set_factors(F)
    F64 = [f.astype(np.float64) for f in F]
    ev = m.evaluate(lam=0.5)
    assert ev["loss"] == pytest.approx(O.loss_sq(ocoo(tr), F64), rel=1e-6)
    assert ev["rmse_te"] == pytest.approx(O.rmse(ocoo(te), F64), rel=1e-6)
    assert ev["rmse_vl"] == pytest.approx(ev["rmse_te"], rel=1e-12)
    assert ev["frob"] == pytest.approx(O.frob_sq(tr.dims, F64, 0.5), rel=1e-6)


ALS_CASES = ([(R, ct.ALS_SPANS) for R in (8, 16, 32, 64)] +
             [(R, ct.ALS_ROWS) for R in (32, 64, 128)])


@pytest.mark.parametrize("R,kernel", ALS_CASES)
@pytest.mark.parametrize("tile_len", [64, 256, 4096])
def test_als_mode_update(ctx, R, kernel, tile_len):
    """SPANS: tile_len nonzeros per span; ROWS: rows longer than tile_len are
    cut into chunks (64 makes most rows chunked)."""
    tr, te = small_problem(3, nnz=30000, seed=20 + R)
    say(f"[als_mode] R={R} kernel={kernel} tile_len={tile_len}")
    m = ct.Model(ctx, tr, None, te, rank=R)
    F = init_factors(tr.dims, R, 5)
    m.set_factors(F)
    oc = ocoo(tr)
    F64 = [f.astype(np.float64) for f in F]
    opt = ct.Options(algorithm=ct.ALS, rank=R, lambda_=0.05, tile_len=tile_len, als_kernel=kernel)
    for mode in range(3):
        m.als_mode_update(mode, opt)
        assert O.als_mode(oc, mode, F64, 0.05) == 0
        got = m.factor(mode)
        e = relerr(got, F64[mode])
        assert e < 1e-5, (mode, e)
        # continue from identical inputs: re-seed the GPU with the oracle's factors
        F64 = [f.astype(np.float32).astype(np.float64) for f in F64]
        m.set_factors([f.astype(np.float32) for f in F64])


@pytest.mark.parametrize("R,kernel", [(R, ct.ALS_SPANS) for R in (8, 16, 32, 64)] +
                         [(R, ct.ALS_ROWS) for R in (32, 64)])
def test_als_mode_update_four_modes(ctx, R, kernel):
    tr, te = small_problem(4, nnz=30000, seed=40 + R)
    say(f"[als_mode4] R={R} kernel={kernel}")
    m = ct.Model(ctx, tr, None, te, rank=R)
    F = init_factors(tr.dims, R, 6)
    m.set_factors(F)
    oc = ocoo(tr)
    F64 = [f.astype(np.float64) for f in F]
    opt = ct.Options(algorithm=ct.ALS, rank=R, lambda_=0.05, tile_len=256, als_kernel=kernel)
    for mode in range(4):
        m.als_mode_update(mode, opt)
        assert O.als_mode(oc, mode, F64, 0.05) == 0
        e = relerr(m.factor(mode), F64[mode])
        assert e < 1e-5, (mode, e)
        F64 = [f.astype(np.float32).astype(np.float64) for f in F64]
        m.set_factors([f.astype(np.float32) for f in F64])


@pytest.mark.parametrize("R,kernel", [(16, ct.ALS_SPANS), (32, ct.ALS_SPANS), (32, ct.ALS_ROWS), (64, ct.ALS_ROWS)])
@pytest.mark.parametrize("flags", [ct.QUIRK_TILE_HEAD | ct.QUIRK_LAST_LANE,
                                   ct.QUIRK_TILE_HEAD | ct.QUIRK_LAST_LANE | ct.QUIRK_TRUNC_VALUES])
def test_als_ref_quirks_match_oracle(ctx, R, kernel, flags):
    """ALS mode updates with the reference's reduction defects
    (options.ref_quirks) against the oracle's plain ALS on the nonzeros
    repeated by the oracle's quirk weights (values truncated under Q5),
    1e-5 per factor; the update differs from the intended one."""
    tr, te = small_problem(3, nnz=30000, seed=90 + R)
    say(f"[als_quirks] R={R} kernel={kernel} flags={flags}")
    m = ct.Model(ctx, tr, None, te, rank=R)
    F = init_factors(tr.dims, R, 8)
    m.set_factors(F)
    oc = ocoo(tr)
    F64 = [f.astype(np.float64) for f in F]
    opt = ct.Options(algorithm=ct.ALS, rank=R, lambda_=0.05, tile_len=256, als_kernel=kernel, ref_quirks=flags)
    for mode in range(3):
        plain = [f.copy() for f in F64]
        m.als_mode_update(mode, opt)
        assert O.als_mode(O.quirk_multiset(oc, mode, flags), mode, F64, 0.05) == 0
        e = relerr(m.factor(mode), F64[mode])
        assert e < 1e-5, (mode, e)
        assert O.als_mode(oc, mode, plain, 0.05) == 0
        assert relerr(F64[mode], plain[mode]) > 1e-3, mode  # the defects change the update
        F64 = [f.astype(np.float32).astype(np.float64) for f in F64]
        m.set_factors([f.astype(np.float32) for f in F64])
    # the plain CISS is untouched: an ordinary update afterwards matches the intended rule
    opt.ref_quirks = 0
    m.als_mode_update(0, opt)
    assert O.als_mode(oc, 0, F64, 0.05) == 0
    assert relerr(m.factor(0), F64[0]) < 1e-5


def test_ref_quirks_scope(ctx):
    """SGD takes no quirk bits; CCD's need Q4 (the reference's CCD has only
    that form); ALS ignores the CCD bits; unknown bits are errors."""
    tr, _ = small_problem(3, nnz=5000, seed=4)
    m = ct.Model(ctx, tr, None, None, rank=16)
    F = init_factors(tr.dims, 16, 1)
    m.set_factors(F)
    with pytest.raises(ct.CtcError, match="ALS and CCD kernels only"):
        m.sgd_epoch(ct.Options(algorithm=ct.SGD, rank=16, lambda_=0.05, ref_quirks=ct.QUIRK_TILE_HEAD), lr=1e-3)
    with pytest.raises(ct.CtcError, match="NO_ADDBACK"):
        m.ccd_epoch(ct.Options(algorithm=ct.CCD, rank=16, lambda_=0.05, ref_quirks=ct.QUIRK_TILE_HEAD))
    for alg in (ct.ALS, ct.CCD):
        with pytest.raises(ct.CtcError, match="unknown bits"):
            o = ct.Options(algorithm=alg, rank=16, lambda_=0.05, ref_quirks=32 | ct.QUIRK_NO_ADDBACK)
            m.als_mode_update(0, o) if alg == ct.ALS else m.ccd_epoch(o)
    m.als_mode_update(0, ct.Options(algorithm=ct.ALS, rank=16, lambda_=0.05,
                                    ref_quirks=ct.QUIRK_NO_ADDBACK | ct.QUIRK_CCD_GRID))
    F64 = [f.astype(np.float64) for f in F]
    assert O.als_mode(ocoo(tr), 0, F64, 0.05) == 0
    assert relerr(m.factor(0), F64[0]) < 1e-5  # the intended ALS update


def test_ref_quirks_every_term_dropped(ctx):
    """20 nonzeros, all on row 0 of mode 0: one partial tile whose only row
    segment Q1 drops.  The reference then zeroes the ALS factor and solves
    lambda I a = 0 (every row 0), and its CCD writes 0 / lambda to row 0 and
    keeps the rows without nonzeros; the oracle agrees on both."""
    rng = np.random.default_rng(5)
    n, dims, R = 20, (4, 9, 7), 8
    ind = [np.zeros(n, np.uint32), rng.integers(0, 9, n).astype(np.uint32), rng.integers(0, 7, n).astype(np.uint32)]
    vals = rng.uniform(1, 5, n)
    tr = ct.Coo.from_arrays(ind, vals, dims)
    oc = O.Coo(ind, vals, dims)
    F = init_factors(dims, R, 3)
    m = ct.Model(ctx, tr, None, None, rank=R)
    m.set_factors(F)
    m.als_mode_update(0, ct.Options(algorithm=ct.ALS, rank=R, lambda_=0.05, ref_quirks=ct.QUIRK_TILE_HEAD))
    assert np.all(m.factor(0) == 0)
    F64 = [f.astype(np.float64) for f in F]
    assert O.als_mode(O.quirk_multiset(oc, 0, ct.QUIRK_TILE_HEAD), 0, F64, 0.05) == 0
    assert np.all(F64[0] == 0)
    m.set_factors(F)
    m.ccd_epoch(ct.Options(algorithm=ct.CCD, rank=R, lambda_=5.0,
                           ref_quirks=ct.QUIRK_NO_ADDBACK | ct.QUIRK_TILE_HEAD))
    F64 = [f.astype(np.float64) for f in F]
    O.ccd_jacobi_epoch(oc, F64, 5.0, ct.QUIRK_NO_ADDBACK | ct.QUIRK_TILE_HEAD)
    assert np.all(F64[0][0] == 0) and np.array_equal(F64[0][1:], F[0][1:])
    for mode in range(3):
        assert relerr(m.factor(mode), F64[mode]) < 1e-6, mode


@pytest.mark.parametrize("R", [8, 32])
@pytest.mark.parametrize("flags", [ct.QUIRK_NO_ADDBACK,
                                   ct.QUIRK_NO_ADDBACK | ct.QUIRK_TILE_HEAD | ct.QUIRK_LAST_LANE,
                                   ct.QUIRK_NO_ADDBACK | ct.QUIRK_TILE_HEAD | ct.QUIRK_LAST_LANE
                                   | ct.QUIRK_TRUNC_VALUES | ct.QUIRK_CCD_GRID])
def test_ccd_ref_quirks_match_oracle(ctx, R, flags):
    """CCD epochs in the reference's form (options.ref_quirks with Q4, and the
    tile defects Q1/Q2, Q5 values and Q3 grid as asked) against the oracle's
    FP64 restatement from the same factors, two epochs, 1e-5 per factor;
    the trajectory differs from CCD++'s."""
    tr, te = small_problem(3, nnz=30000, seed=120 + R)
    say(f"[ccd_quirks] R={R} flags={flags}")
    m = ct.Model(ctx, tr, None, te, rank=R)
    F = init_factors(tr.dims, R, 9)
    m.set_factors(F)
    oc = ocoo(tr)
    F64 = [f.astype(np.float64) for f in F]
    opt = ct.Options(algorithm=ct.CCD, rank=R, lambda_=5.0, ref_quirks=flags)
    for ep in range(2):
        m.ccd_epoch(opt)
        O.ccd_jacobi_epoch(oc, F64, 5.0, flags)
        for mode in range(3):
            e = relerr(m.factor(mode), F64[mode])
            assert e < 1e-5, (ep, mode, e)
    plain = ct.Model(ctx, tr, None, te, rank=R)
    plain.set_factors(F)
    plain.ccd_epoch(ct.Options(algorithm=ct.CCD, rank=R, lambda_=5.0))
    plain.ccd_epoch(ct.Options(algorithm=ct.CCD, rank=R, lambda_=5.0))
    assert relerr(m.factor(0), plain.factor(0)) > 1e-3


def test_als_deterministic(ctx):
    tr, te = small_problem(3, nnz=30000, seed=77)
    say("[als_deterministic]")
    outs = []
    for _ in range(2):
        m = ct.Model(ctx, tr, None, te, rank=32)
        m.set_factors(init_factors(tr.dims, 32, 1))
        opt = ct.Options(algorithm=ct.ALS, rank=32, lambda_=0.05, tile_len=128)
        for mode in range(3):
            m.als_mode_update(mode, opt)
        outs.append(m.factors())
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_als_config1_ten_iterations(ctx):
    """Config 1 of the benchmark definition at full size: 2000x1500x300,
    1M nonzeros, rank-8 truth -> [1,5] + N(0,0.1), R=16, lambda=0.05, 90/10
    split, 10 iterations; every factor within 1e-5 of the FP64 oracle after
    every iteration."""
    tr, te = ct.synth((2000, 1500, 300), 1_000_000, true_rank=8, value_map=ct.VALUES_AFFINE_1_5,
                      noise_std=0.1, test_fraction=0.1, seed=0x5EED0001)
    say(f"[als_config1] train={tr.nnz} test={te.nnz}")
    R = 16
    m = ct.Model(ctx, tr, None, te, rank=R)
    F = init_factors(tr.dims, R, 11)
    m.set_factors(F)
    oc = ocoo(tr)
    ot = ocoo(te)
    F64 = [f.astype(np.float64) for f in F]
    orders = [O.mode_order(oc, md) for md in range(3)]
    opt = ct.Options(algorithm=ct.ALS, rank=R, lambda_=0.05)
    worst = 0.0
    for it in range(10):
        for mode in range(3):
            m.als_mode_update(mode, opt)
            assert O.als_mode(oc, mode, F64, 0.05, order=orders[mode], nthreads=8) == 0
        for mode in range(3):
            worst = max(worst, relerr(m.factor(mode), F64[mode]))
        rm_gpu = m.evaluate()["rmse_te"]
        rm_orc = O.rmse(ot, F64)
        say(f"  it={it} worst_rel={worst:.2e} rmse_te gpu={rm_gpu:.6f} oracle={rm_orc:.6f}")
        assert abs(rm_gpu - rm_orc) < 1e-4
    assert worst < 1e-5, worst


def test_context_handle_released_before_its_model():
    """The context handle may be destroyed first: the model keeps the context
    alive (reference count in the library) and still runs and frees cleanly."""
    from cutc_amd import _lib
    tr, te = small_problem(3, nnz=5000, seed=9)
    say("[ctx_lifetime]")
    c2 = ct.Context(0)
    m = ct.Model(c2, tr, None, te, rank=16)
    m.set_factors(init_factors(tr.dims, 16, 3))
    _lib.lib().ctc_context_destroy(c2._h)  # the caller's handle goes first
    c2._h = None
    m.als_mode_update(0, ct.Options(algorithm=ct.ALS, rank=16, lambda_=0.05))
    assert np.isfinite(m.evaluate(0.05)["rmse_te"])
    del m


def test_als_not_spd_is_reported(ctx):
    # lambda = 0 and rows with fewer nonzeros than R: singular normal systems
    tr, _ = ct.synth((500, 40, 30), 1500, test_fraction=0.0, seed=3)
    say("[als_not_spd]")
    m = ct.Model(ctx, tr, None, None, rank=16)
    m.set_factors(init_factors(tr.dims, 16, 2))
    with pytest.raises(ct.CtcError) as e:
        m.als_mode_update(0, ct.Options(algorithm=ct.ALS, rank=16, lambda_=0.0))
    assert e.value.status == 7
    # the model stays usable afterwards
    m.als_mode_update(0, ct.Options(algorithm=ct.ALS, rank=16, lambda_=0.1))


# --------------------------------------------------------------------------
# SGD, default schedule: bulk-synchronous walkers (oracle.sgd_sync_epoch)
SYNC_CASES = [  # nmodes, R, span, sub-step, nnz
    (3, 8, 96, 7, 20000), (3, 16, 512, 64, 40000), (3, 32, 256, 32, 40000), (3, 64, 2048, 128, 60000),
    (3, 128, 1024, 64, 20000), (4, 16, 512, 48, 30000), (4, 32, 256, 32, 20000), (4, 128, 640, 40, 20000),
]


def sync_problem(N, nnz, seed):
    dims = (600, 300, 50) if N == 3 else (200, 100, 30, 12)
    zipf = (1.0, 1.0, 0.0) if N == 3 else (1.0, 1.0, 0.0, 0.0)
    return ct.synth(dims, nnz, zipf=zipf, true_rank=4, value_map=ct.VALUES_INT_1_5, noise_std=0.3,
                    test_fraction=0.05, seed=seed)


@pytest.mark.parametrize("mixed", [False, True], ids=["fp32", "mixed"])
@pytest.mark.parametrize("N,R,span,sub,nnz", SYNC_CASES)
def test_sgd_sync_matches_oracle(ctx, N, R, span, sub, nnz, mixed):
    """Two epochs of the default schedule against its FP64 restatement, with
    hot rows (threshold 400) next to cold shared rows, walkers whose spans and
    sub-steps do not align with tiles or wave widths, 3- and 4-way tensors.
    mixed: the mixed-precision path (run deltas of modes 1.. staged in BF16)
    against the oracle's mixed variant, which rounds at the same points, with
    a bar of 1e-4: a delta summed in FP32 on the GPU and in FP64 in the oracle
    differs by about 1e-7 relative, which puts the two on different sides of a
    BF16 rounding midpoint for about one rounding in 1e5, and each such case
    moves one delta entry by a whole BF16 step (2^-8 relative); measured
    1.1e-5 on mode 2 of the rank-16 case."""
    tr, te = sync_problem(N, nnz, 11 + R + N)
    say(f"[sgd_sync] N={N} R={R} span={span} sub={sub} nnz={tr.nnz} mixed={mixed}")
    m = ct.Model(ctx, tr, None, te, rank=R)
    F = init_factors(tr.dims, R, 5)
    m.set_factors(F)
    F64 = [f.astype(np.float64) for f in F]
    otr, ote = ocoo(tr), ocoo(te)
    order, _ = O.mode_order(otr, 0)
    opt = ct.Options(algorithm=ct.SGD, rank=R, lambda_=0.05, sgd_scale=1.0, tile_len=span, sgd_substep=sub,
                     sgd_hot_rows=400, precision=ct.PREC_MIXED if mixed else ct.PREC_FP32)
    for ep in range(2):
        m.sgd_epoch(opt, lr=5e-3)
        O.sgd_sync_epoch(otr, F64, 5e-3, 0.05, span, sub, 400, order=order, nthreads=4, mixed=mixed)
        errs = [relerr(m.factor(k), F64[k]) for k in range(N)]
        say(f"  epoch {ep + 1}: rel err", ["%.2e" % e for e in errs])
        assert max(errs) < (1e-4 if mixed else 1e-5), errs
        assert abs(m.evaluate()["rmse_te"] - O.rmse(ote, F64)) < 1e-5


@pytest.mark.parametrize("mixed", [False, True], ids=["fp32", "mixed"])
def test_sgd_sync_is_bitwise_reproducible(ctx, mixed):
    """Deltas are summed in a fixed order (staged runs in FP64 slot by slot,
    mode 0's shared rows as fixed-point integers): two sweeps from the same
    factors give the same bits (2M nonzeros, thousands of walkers on shared
    rows), on the FP32 and the mixed-precision path."""
    tr, _ = ct.synth((50000, 8000, 500), 2_000_000, zipf=(1.0, 1.0, 0.0), true_rank=8,
                     value_map=ct.VALUES_INT_1_5, noise_std=0.3, test_fraction=0.0, seed=3)
    say(f"[sgd_sync_bits] nnz={tr.nnz} mixed={mixed}")
    m = ct.Model(ctx, tr, None, None, rank=32)
    F = init_factors(tr.dims, 32, 2)
    opt = ct.Options(algorithm=ct.SGD, rank=32, lambda_=0.05, sgd_scale=1.0, tile_len=2048, sgd_substep=32,
                     sgd_hot_rows=1000, precision=ct.PREC_MIXED if mixed else ct.PREC_FP32)
    outs = []
    for _ in range(2):
        m.set_factors(F)
        for _ in range(2):
            m.sgd_epoch(opt, lr=1e-3)
        outs.append([m.factor(k).copy() for k in range(3)])
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)


def test_sgd_sync_test_rmse_trajectory(ctx):
    """Twelve epochs of the default solve loop (exponential decay) on a skewed
    150K-nonzero tensor where the test RMSE falls by more than 0.3: per epoch
    within 1e-3 of the FP64 restatement and within 5% of the oracle's own
    movement, and factors after epoch 1 within 1e-5."""
    tr, te = ct.synth((3000, 1500, 200), 150_000, zipf=(1.0, 1.0, 0.0), true_rank=10,
                      value_map=ct.VALUES_INT_1_5, noise_std=0.3, test_fraction=0.01, seed=0x5EED0002)
    R, lr, span, sub = 32, 2e-2, 1024, 32
    say(f"[sgd_sync_traj] train={tr.nnz}")
    m = ct.Model(ctx, tr, None, te, rank=R)
    F = [0.5 * f for f in init_factors(tr.dims, R, 6)]
    otr, ote = ocoo(tr), ocoo(te)
    order, _ = O.mode_order(otr, 0)
    # factors after one epoch
    m.set_factors(F)
    F64 = [f.astype(np.float64) for f in F]
    opt = ct.Options(algorithm=ct.SGD, rank=R, lambda_=0.05, learning_rate=lr, lr_schedule=ct.LR_EXP_DECAY,
                     lr_decay=0.95, sgd_scale=1.0, max_epochs=12, max_bad_epochs=0, tile_len=span, sgd_substep=sub,
                     sgd_hot_rows=2000)
    m.sgd_epoch(opt, lr=lr)
    O.sgd_sync_epoch(otr, F64, lr, 0.05, span, sub, 2000, order=order, nthreads=4)
    assert max(relerr(m.factor(k), F64[k]) for k in range(3)) < 1e-5
    # the solve loop
    m.set_factors(F)
    F64 = [f.astype(np.float64) for f in F]
    rm_orc, _ = O.sgd_sync(otr, F64, lr, 0.05, 12, span, sub, 2000, test=ote, decay=0.95, order=order, nthreads=4)
    res = m.solve(opt)
    rm_gpu = [h["rmse_te"] for h in res.history[1:]]
    say("  oracle", np.round(rm_orc, 5))
    say("  gpu   ", np.round(rm_gpu, 5))
    assert len(rm_gpu) == 12
    moved = abs(O.rmse(ote, [f.astype(np.float64) for f in F]) - min(rm_orc))
    assert moved > 0.3, moved
    gap = np.max(np.abs(np.array(rm_gpu) - np.array(rm_orc)))
    assert gap < 1e-3 and gap < 0.05 * moved, gap


def test_sgd_sync_mixed_trajectory(ctx):
    """The mixed-precision SGD solve loop (run deltas of modes 1.. staged in
    BF16; FP32 factors and arithmetic) on the tensor of
    test_sgd_sync_test_rmse_trajectory: per epoch within 1e-3 of the FP64
    oracle of the FP32 schedule (the precision target) and of the oracle's
    mixed variant.  The step (lr 2e-2) takes the test RMSE from 2.18 to 0.68 in
    three epochs, and that descent amplifies the few deltas that BF16 rounds
    to a neighbouring value on one side and not the other (see
    test_sgd_sync_matches_oracle): measured 4.6e-4 from the mixed oracle at
    epoch 2, with the GPU between the two oracles."""
    tr, te = ct.synth((3000, 1500, 200), 150_000, zipf=(1.0, 1.0, 0.0), true_rank=10,
                      value_map=ct.VALUES_INT_1_5, noise_std=0.3, test_fraction=0.01, seed=0x5EED0002)
    R, lr, span, sub = 32, 2e-2, 1024, 32
    say(f"[sgd_sync_mixed_traj] train={tr.nnz}")
    m = ct.Model(ctx, tr, None, te, rank=R)
    F = [0.5 * f for f in init_factors(tr.dims, R, 6)]
    otr, ote = ocoo(tr), ocoo(te)
    order, _ = O.mode_order(otr, 0)
    rm = {}
    for mixed in (False, True):
        F64 = [f.astype(np.float64) for f in F]
        rm[mixed], _ = O.sgd_sync(otr, F64, lr, 0.05, 12, span, sub, 2000, test=ote, decay=0.95, order=order,
                                  nthreads=4, mixed=mixed)
    m.set_factors(F)
    opt = ct.Options(algorithm=ct.SGD, rank=R, lambda_=0.05, learning_rate=lr, lr_schedule=ct.LR_EXP_DECAY,
                     lr_decay=0.95, sgd_scale=1.0, max_epochs=12, max_bad_epochs=0, tile_len=span, sgd_substep=sub,
                     sgd_hot_rows=2000, precision=ct.PREC_MIXED)
    res = m.solve(opt)
    rm_gpu = np.array([h["rmse_te"] for h in res.history[1:]])
    say("  oracle fp64 ", np.round(rm[False], 5))
    say("  oracle mixed", np.round(rm[True], 5))
    say("  gpu mixed   ", np.round(rm_gpu, 5))
    gap_mixed = np.max(np.abs(rm_gpu - np.array(rm[True])))
    gap_fp64 = np.max(np.abs(rm_gpu - np.array(rm[False])))
    say(f"  gap to the mixed oracle {gap_mixed:.2e}, to the FP64 oracle {gap_fp64:.2e}")
    assert gap_mixed < 1e-3, gap_mixed
    assert gap_fp64 < 1e-3, gap_fp64


def test_sgd_mixed_needs_the_sync_schedule(ctx):
    tr, _ = sync_problem(3, 2000, 5)
    m = ct.Model(ctx, tr, None, None, rank=16)
    m.set_factors(init_factors(tr.dims, 16, 1))
    opt = ct.Options(algorithm=ct.SGD, rank=16, lambda_=0.05, sgd_update=ct.SGD_AUTO, precision=ct.PREC_MIXED)
    with pytest.raises(ct.CtcError, match="bulk-synchronous"):
        m.sgd_epoch(opt, lr=1e-3)


# --------------------------------------------------------------------------
# SGD, opt-in Hogwild policies against the sequential rule
@pytest.mark.parametrize("policy,tile_len", [(ct.SGD_AUTO, 2048), (ct.SGD_AUTO, 1024), (ct.SGD_ATOMIC, 256),
                                             (ct.SGD_HOT_ATOMIC, 2048)])
def test_sgd_test_rmse_trajectory(ctx, policy, tile_len):
    """SGD (lr 1e-3, x0.95 per epoch, lambda 0.05, scale 1) test-RMSE per
    epoch within 1e-3 of the sequential FP64 oracle."""
    tr, te = ct.synth((3000, 1500, 200), 150_000, zipf=(1.0, 1.0, 0.0), true_rank=10,
                      value_map=ct.VALUES_INT_1_5, noise_std=0.3, test_fraction=0.01, seed=0x5EED0002)
    R = 32
    say(f"[sgd] train={tr.nnz}")
    m = ct.Model(ctx, tr, None, te, rank=R)
    F = [0.5 * f for f in init_factors(tr.dims, R, 6)]
    m.set_factors(F)
    F64 = [f.astype(np.float64) for f in F]
    epochs = 12
    rm_orc, _ = O.sgd(ocoo(tr), F64, 1e-3, 0.05, epochs, test=ocoo(te), scale=1.0, decay=0.95)
    opt = ct.Options(algorithm=ct.SGD, rank=R, lambda_=0.05, learning_rate=1e-3,
                     lr_schedule=ct.LR_EXP_DECAY, lr_decay=0.95, sgd_scale=1.0, max_epochs=epochs,
                     max_bad_epochs=0, tile_len=tile_len, sgd_update=policy)
    res = m.solve(opt)
    rm_gpu = [h["rmse_te"] for h in res.history[1:]]
    say("  oracle", np.round(rm_orc, 5))
    say("  gpu   ", np.round(rm_gpu, 5))
    assert len(rm_gpu) == epochs
    gap = np.max(np.abs(np.array(rm_gpu) - np.array(rm_orc)))
    assert gap < 1e-3, gap
    assert rm_gpu[-1] < rm_gpu[0]


def test_sgd_uniform_mode_never_hot(ctx):
    """A row is hot only above the threshold and above 4x its mode's mean.  Mode
    2 here has 8 rows of ~18.6K nonzeros each (uniform) and mode 1 rows of ~100:
    with a threshold of 1000 no row qualifies, so AUTO runs exactly the policy of
    the default threshold (no hot rows) and the two test-RMSE trajectories agree
    to 2e-3 (twice the trajectory test's bar: mode-0 rows split between walkers
    are stored racily, so runs differ slightly).  This checks that the bound
    keeps the uniform mode out of the hot set; the stall a racy uniform mode
    causes is shown on the 100M-nonzero tensor, not here (this tensor levels off
    within an epoch, so even the RACY policy stays within 5e-3 of AUTO)."""
    tr, te = ct.synth((3000, 1500, 8), 150_000, zipf=(1.0, 0.0, 0.0), true_rank=10,
                      value_map=ct.VALUES_INT_1_5, noise_std=0.3, test_fraction=0.01, seed=0x5EED0012)
    R = 32
    say(f"[sgd_uniform] train={tr.nnz}")
    m = ct.Model(ctx, tr, None, te, rank=R)
    F = [0.5 * f for f in init_factors(tr.dims, R, 7)]
    runs = []
    for policy, thr in ((ct.SGD_AUTO, 0), (ct.SGD_AUTO, 1000)):
        m.set_factors(F)
        opt = ct.Options(algorithm=ct.SGD, rank=R, lambda_=0.05, learning_rate=5e-3,
                         lr_schedule=ct.LR_EXP_DECAY, lr_decay=0.95, sgd_scale=1.0, max_epochs=6,
                         max_bad_epochs=0, sgd_update=policy, sgd_hot_rows=thr, tile_len=2048)
        runs.append(np.array([h["rmse_te"] for h in m.solve(opt).history]))
        say(f"  policy {policy} threshold {thr or 'default'}:", np.round(runs[-1], 5))
    assert np.max(np.abs(runs[0] - runs[1])) < 2e-3
    assert runs[1][-1] < runs[1][0]


@pytest.mark.parametrize("R", [8, 16, 64, 128])
@pytest.mark.parametrize("policy", [ct.SGD_AUTO, ct.SGD_ATOMIC, ct.SGD_RACY, ct.SGD_HOT_ATOMIC, ct.SGD_SYNC])
def test_sgd_single_walker_matches_sequential(ctx, R, policy):
    """With one tile the GPU walk is one sequential walker: factors match the
    sequential oracle to FP32 rounding (the bulk-synchronous schedule with
    1-nonzero sub-steps, so that every nonzero starts from the merged rows)."""
    tr, _ = ct.synth((10, 8, 6), 64, test_fraction=0.0, seed=R)
    say(f"[sgd_single] R={R}")
    m = ct.Model(ctx, tr, None, None, rank=R)
    F = init_factors(tr.dims, R, 8)
    m.set_factors(F)
    F64 = [f.astype(np.float64) for f in F]
    O.sgd_epoch(ocoo(tr), F64, 1e-2, 0.1, 1.0)
    m.sgd_epoch(ct.Options(algorithm=ct.SGD, rank=R, lambda_=0.1, sgd_scale=1.0, tile_len=64, sgd_update=policy,
                           sgd_substep=1), lr=1e-2)
    for md in range(3):
        assert relerr(m.factor(md), F64[md]) < 1e-5


# --------------------------------------------------------------------------
@pytest.mark.parametrize("nmodes", [3, 4])
@pytest.mark.parametrize("per_occ", [False, True])
def test_grad(ctx, nmodes, per_occ):
    import torch
    tr, _ = small_problem(nmodes, seed=30 + nmodes, test_fraction=0.0)
    R = 32
    say(f"[grad] N={nmodes} per_occ={per_occ}")
    m = ct.Model(ctx, tr, None, None, rank=R)
    F = init_factors(tr.dims, R, 9)
    dev = torch.device("cuda:0")
    Ft = [torch.from_numpy(f).to(dev) for f in F]
    Gt = [torch.empty_like(f) for f in Ft]
    torch.cuda.synchronize()
    loss = m.grad_device([f.data_ptr() for f in Ft], [g.data_ptr() for g in Gt], 0.3, per_occ)
    torch.cuda.synchronize()
    oloss, OG = O.grad(ocoo(tr), [f.astype(np.float64) for f in F], 0.3, per_occ)
    assert loss == pytest.approx(oloss, rel=1e-5)
    for md in range(nmodes):
        assert relerr(Gt[md].cpu().numpy(), OG[md]) < 1e-5


# --------------------------------------------------------------------------
def test_solve_loop_and_reference_shaped_entry(ctx):
    tr, te = small_problem(3, nnz=20000, seed=41)
    say("[solve_loop]")
    m = ct.Model(ctx, tr, te, te, rank=16)
    m.init_random(3)
    res = m.solve(ct.Options(algorithm=ct.ALS, rank=16, lambda_=0.05, max_epochs=30))
    assert res.epochs >= 2 and res.history[0]["rmse_vl"] > res.best_rmse
    assert res.ms_per_epoch > 0 and res.nnz_per_s > 0
    best = m.factors(best=True)
    m2 = ct.Model(ctx, tr, te, te, rank=16)
    m2.set_factors(best)
    assert m2.evaluate()["rmse_vl"] == pytest.approx(res.best_rmse, rel=1e-9)
    mats = [np.ascontiguousarray(f, dtype=np.float64) for f in init_factors(tr.dims, 16, 2)]
    bm = [np.zeros_like(f) for f in mats]
    out = ct.tc_als(tr, None, None, te, te, mats, bm, regularization_index=0.05)
    assert out["best_rmse"] < 2.0 and np.abs(bm[0]).sum() > 0


# --------------------------------------------------------------------------
# CCD++
# --------------------------------------------------------------------------
@pytest.mark.parametrize("R,tile_len,inner,nmodes", [
    (8, 64, 1, 3), (32, 256, 1, 3), (32, 4096, 2, 3), (16, 128, 1, 4), (64, 512, 1, 3)])
def test_ccd_epochs(ctx, R, tile_len, inner, nmodes):
    """Canonical CCD++: every factor within 1e-5 of the FP64 oracle after each
    of 4 epochs, from the same initial factors."""
    tr, te = small_problem(nmodes, nnz=25000, seed=50 + R + nmodes, test_fraction=0.05)
    say(f"[ccd] R={R} tile_len={tile_len} inner={inner} N={nmodes}")
    m = ct.Model(ctx, tr, None, te, rank=R)
    F = init_factors(tr.dims, R, 12)
    m.set_factors(F)
    F64 = [f.astype(np.float64) for f in F]
    oc = ocoo(tr)
    orders = [O.mode_order(oc, md) for md in range(nmodes)]
    res = O.residual(oc, F64)
    opt = ct.Options(algorithm=ct.CCD, rank=R, lambda_=0.05, tile_len=tile_len, inner_iters=inner)
    for ep in range(4):
        m.ccd_epoch(opt)
        O.ccd_epoch(oc, F64, 0.05, res, inner=inner, orders=orders)
        errs = [relerr(m.factor(md), F64[md]) for md in range(nmodes)]
        say(f"  epoch {ep} errs {['%.1e' % e for e in errs]}")
        assert max(errs) < 1e-5, errs


def test_ccd_deterministic_and_solve(ctx):
    tr, te = small_problem(3, nnz=20000, seed=61)
    say("[ccd_det]")
    outs = []
    for _ in range(2):
        m = ct.Model(ctx, tr, te, te, rank=16)
        m.set_factors(init_factors(tr.dims, 16, 3))
        res = m.solve(ct.Options(algorithm=ct.CCD, rank=16, lambda_=0.05, max_epochs=5, max_bad_epochs=0))
        outs.append(m.factors())
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    assert res.history[-1]["rmse_tr"] < res.history[0]["rmse_tr"]


# --------------------------------------------------------------------------
# differentiation: autograd.Function over the adjoint kernel
# --------------------------------------------------------------------------
def test_autograd_finite_differences_and_sgd_direction(ctx):
    """4-way tensor shaped like the differentiation configuration (Zipf modes
    0-1, uniform 2-3), gradient through torch.autograd against central finite
    differences (eps = 1e-3, FP64 oracle loss) on a 1000-nonzero subsample,
    and against the summed reference SGD step (per-occurrence L2)."""
    import torch
    from cutc_amd.autograd import completion_loss
    tr, _ = ct.synth((2000, 1000, 100, 24), 1000, zipf=(1.0, 1.0, 0.0, 0.0), true_rank=16,
                     value_map=ct.VALUES_RAW, noise_std=0.2, test_fraction=0.0, seed=0x5EED0003)
    say("[autograd] nnz", tr.nnz)
    R, lam = 32, 0.1
    m = ct.Model(ctx, tr, None, None, rank=R)
    F = [0.3 * f for f in init_factors(tr.dims, R, 13)]
    dev = torch.device("cuda:0")
    oc = ocoo(tr)
    for per_occ in (False, True):
        Ft = [torch.from_numpy(f).to(dev).requires_grad_(True) for f in F]
        loss = completion_loss(m, Ft, lam, per_occ)
        loss.backward()
        G = [f.grad.detach().cpu().numpy().astype(np.float64) for f in Ft]
        F64 = [f.astype(np.float64) for f in F]
        oloss, _ = O.grad(oc, F64, lam, per_occ)
        assert float(loss.detach()) == pytest.approx(oloss, rel=1e-5)
        rng = np.random.default_rng(5)
        eps = 1e-3
        for md in range(4):
            rows = np.unique(tr.ind(md))
            for _ in range(5):
                i, r = int(rng.choice(rows)), int(rng.integers(R))
                Fp = [f.copy() for f in F64]
                Fm = [f.copy() for f in F64]
                Fp[md][i, r] += eps
                Fm[md][i, r] -= eps
                fd = (O.grad(oc, Fp, lam, per_occ)[0] - O.grad(oc, Fm, lam, per_occ)[0]) / (2 * eps)
                assert G[md][i, r] == pytest.approx(fd, rel=1e-4, abs=1e-5), (md, i, r, per_occ)
        if per_occ:
            # summed reference SGD direction (SG/sgd.cu:93-95) = -gradient
            acc = [np.zeros_like(f) for f in F64]
            ind = [tr.ind(q) for q in range(4)]
            for p in range(tr.nnz):
                rows = [F64[q][ind[q][p]] for q in range(4)]
                e = tr.vals[p] - np.sum(rows[0] * rows[1] * rows[2] * rows[3])
                for q in range(4):
                    o = np.prod([rows[s] for s in range(4) if s != q], axis=0)
                    acc[q][ind[q][p]] += e * o - lam * rows[q]
            for q in range(4):
                assert relerr(-G[q], acc[q]) < 1e-5


def test_autograd_training_decreases_loss(ctx):
    import torch
    from cutc_amd.autograd import completion_loss
    tr, te = ct.synth((400, 300, 50, 24), 60000, zipf=(1.0, 1.0, 0.0, 0.0), true_rank=16,
                      value_map=ct.VALUES_RAW, noise_std=0.2, test_fraction=0.01, seed=0x5EED0003)
    say("[autograd_train]")
    m = ct.Model(ctx, tr, None, te, rank=32)
    dev = torch.device("cuda:0")
    Ft = [torch.from_numpy(0.5 * f).to(dev).requires_grad_(True) for f in init_factors(tr.dims, 32, 2)]
    opt = torch.optim.Adam(Ft, lr=0.01)
    losses = []
    for _ in range(40):
        opt.zero_grad()
        loss = completion_loss(m, Ft, 0.05)
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
    assert all(b < a for a, b in zip(losses, losses[1:])), losses[::5]
    assert losses[-1] < 0.8 * losses[0], losses[::5]
