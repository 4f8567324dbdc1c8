"""The drop-in surface against the oracle: the reference-shaped entry points
(tc_main_ciss, tc_sgd, tc_ccd; SG/completion.c:422-497, SG/sgd.cuh:18-32,
SG/ccd.cuh:20-33), the `ctc` command line solving from the reference's five
.tns files (SG/tc.c:17-60), the bold-driver learning rate (SG/sgd.cu:314-323)
and the seeded rotating ALS mode order.

Each GPU run is replayed by the FP64 oracle from the same factors with the
same epoch loop: evaluation of epoch 0, the convergence rule of
SG/completion.c:296-362 (oracle.Converge) on the validation RMSE, and the
learning-rate schedule.  Bars: factors within 1e-5 relative per mode,
RMSEs within 1e-5 relative (the CLI and progress lines print 6 significant
digits: 2e-5)."""
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import cutc_amd as ct
import oracle as O
from mt64 import MT19937_64

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ocoo(t):
    ind, vals = t.arrays()
    return O.Coo(ind, vals, t.dims)


def relerr(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def problem(seed, nnz=20000, dims=(300, 200, 60)):
    """train / validation / test from one synthetic tensor (80/10/10)."""
    full, _ = ct.synth(dims, nnz, zipf=(1.0, 0.7, 0.0), true_rank=4, value_map=ct.VALUES_AFFINE_1_5,
                       noise_std=0.1, test_fraction=0.0, seed=seed)
    return full.split(seed)


def oracle_solve(alg, tr, va, F, lam, epochs=None, tol=1e-4, max_bad=3, lr=None, scale=1.0, schedule=None,
                 sgd_shape=None, rotate_seed=None, max_epochs=1000):
    """The epoch loop of ctc_solve (SG/completion.c:296-362) on the FP64
    oracle: evaluation and convergence after epoch 0 and every epoch, the bold
    driver for SGD, the seeded mode rotation for ALS.  -> (per-epoch history of
    (rmse_vl, lr), best factors, best epoch, epochs run)."""
    otr = ocoo(tr)
    ova = ocoo(va)
    orders = [O.mode_order(otr, m) for m in range(3)]
    conv = O.Converge(tolerance=tol, max_badepochs=max_bad)
    hist = []
    r0 = O.rmse(ova, F)
    hist.append((r0, lr))
    conv.step(r0, 0)
    best, best_e = [f.copy() for f in F], 0
    res = O.residual(otr, F) if alg == ct.CCD else None
    rot = MT19937_64(rotate_seed ^ 0xA15) if rotate_seed is not None else None
    prev_obj = None
    e = 0
    while e < (epochs or max_epochs):
        e += 1
        lr_e = lr
        if alg == ct.ALS:
            start = rot() % 3 if rot else 0
            for k in range(3):
                md = (start + k) % 3
                assert O.als_mode(otr, md, F, lam, order=orders[md], nthreads=4) == 0
        elif alg == ct.SGD:
            span, sub, hot = sgd_shape
            O.sgd_sync_epoch(otr, F, lr, lam, span, sub, hot, scale=scale, order=orders[0][0], nthreads=4)
        else:
            O.ccd_epoch(otr, F, lam, res, orders=orders, nthreads=4)
        r = O.rmse(ova, F)
        if alg == ct.SGD and schedule == ct.LR_BOLD_DRIVER:
            obj = O.loss_sq(otr, F) + O.frob_sq(otr.dims, F, lam)
            if e > 1:
                lr = lr * 1.05 if obj < prev_obj else lr * 0.5
            prev_obj = obj
        elif alg == ct.SGD and schedule == ct.LR_EXP_DECAY:
            lr *= 0.95
        hist.append((r, lr_e))
        done, improved = conv.step(r, e)
        if improved:
            best, best_e = [f.copy() for f in F], e
        if done and epochs is None:
            break
    return hist, best, best_e, e


# --------------------------------------------------------------------------
def test_tc_main_ciss_progress_matches_oracle(capfd):
    """tc_main_ciss (ALS): seeded random factors, the reference's progress line
    per epoch; every printed validation and test RMSE against the oracle."""
    tr, va, te = problem(61)
    R, lam, seed = 8, 0.05, 3
    ctx = ct.Context(0)
    m = ct.Model(ctx, tr, None, None, rank=R)
    m.init_random(seed)  # the factors tc_main_ciss starts from
    F = [f.astype(np.float64) for f in m.factors()]
    say("[tc_main_ciss]")
    capfd.readouterr()
    ct.tc_main_ciss(tr, None, None, va, te, ct.ALS,
                    ct.Options(algorithm=ct.ALS, rank=R, lambda_=lam, max_epochs=6, seed=seed))
    out = capfd.readouterr().out
    lines = re.findall(r"epoch:(\d+)\s+loss: (\S+)\s+RMSE-tr: (\S+)\s+RMSE-vl: (\S+)\s+RMSE-te: (\S+)", out)
    assert len(lines) >= 2, out
    epochs = int(lines[-1][0])
    ote = ocoo(te)
    otr, ova = ocoo(tr), ocoo(va)
    orders = [O.mode_order(otr, md) for md in range(3)]
    for e, (_, loss, rtr, rvl, rte) in enumerate(lines):
        if e > 0:
            for md in range(3):
                O.als_mode(otr, md, F, lam, order=orders[md], nthreads=4)
        assert float(rvl) == pytest.approx(O.rmse(ova, F), rel=2e-5), e
        assert float(rte) == pytest.approx(O.rmse(ote, F), rel=2e-5), e
        assert float(rtr) == pytest.approx(O.rmse(otr, F), rel=2e-5), e
    say(f"  {epochs} epochs, final RMSE-te {lines[-1][4]}")


def test_tc_sgd_reference_shaped_matches_oracle():
    """tc_sgd with the reference's defaults (scale 1e-5, bold driver,
    tolerance 1e-4, 3 bad epochs): final and best factors, best RMSE and
    best epoch against the oracle's run of the same loop."""
    tr, va, te = problem(62)
    R, lam, lr = 16, 0.05, 200.0  # lr x scale 1e-5 = 2e-3 per step
    rng = np.random.default_rng(5)
    mats = [np.ascontiguousarray(rng.uniform(-1, 1, (d, R))) for d in tr.dims]
    mats = [m.astype(np.float32).astype(np.float64) for m in mats]  # the library holds FP32 factors
    F = [m.copy() for m in mats]
    bm = [np.zeros_like(m) for m in mats]
    c = ct.Options(algorithm=ct.SGD).to_c()
    say("[tc_sgd]")
    out = ct.tc_sgd(tr, None, None, va, te, mats, bm, regularization_index=lam, learning_rate=lr)
    hist, best, best_e, epochs = oracle_solve(ct.SGD, tr, va, F, lam, lr=lr, scale=c.sgd_scale,
                                              schedule=ct.LR_BOLD_DRIVER,
                                              sgd_shape=(c.tile_len, c.sgd_substep, c.sgd_hot_rows))
    say(f"  gpu best epoch {out['best_epoch']} rmse {out['best_rmse']:.7f}; oracle {best_e} "
        f"{min(h[0] for h in hist):.7f} after {epochs} epochs")
    assert out["best_epoch"] == best_e
    assert out["best_rmse"] == pytest.approx(min(h[0] for h in hist), rel=1e-5)
    for md in range(3):
        assert relerr(mats[md], F[md]) < 1e-5
        assert relerr(bm[md], best[md]) < 1e-5


def test_tc_ccd_reference_shaped_matches_oracle():
    """tc_ccd (canonical CCD++, SG/ccd.cu:532-554): the same comparison."""
    tr, va, te = problem(63)
    R, lam = 8, 0.05
    rng = np.random.default_rng(6)
    mats = [rng.uniform(-1, 1, (d, R)).astype(np.float32).astype(np.float64) for d in tr.dims]
    F = [m.copy() for m in mats]
    bm = [np.zeros_like(m) for m in mats]
    say("[tc_ccd]")
    out = ct.tc_ccd(tr, None, None, va, te, mats, bm, regularization_index=lam)
    hist, best, best_e, epochs = oracle_solve(ct.CCD, tr, va, F, lam)
    say(f"  gpu best epoch {out['best_epoch']} rmse {out['best_rmse']:.7f}; oracle {best_e} after {epochs}")
    assert out["best_epoch"] == best_e
    assert out["best_rmse"] == pytest.approx(min(h[0] for h in hist), rel=1e-5)
    for md in range(3):
        assert relerr(mats[md], F[md]) < 1e-5
        assert relerr(bm[md], best[md]) < 1e-5


def test_cli_solves_from_five_tns_files(tmp_path):
    """`ctc <train_a> <train_b> <train_c> <validate> <test>` (SG/tc.c:40-50):
    the reference's five positional files (the two extra mode-sorted training
    copies from `ctc prep`), ALS; the printed summary against the oracle."""
    full, _ = ct.synth((300, 200, 60), 20000, zipf=(1.0, 0.7, 0.0), true_rank=4, value_map=ct.VALUES_AFFINE_1_5,
                       noise_std=0.1, test_fraction=0.0, seed=64)
    src = tmp_path / "t.tns"
    full.write(str(src), base=1, decimals=-1)
    exe = os.path.join(ROOT, "cutc_amd", "ctc")
    subprocess.run([exe, "prep", str(src), str(tmp_path / "d"), "--seed", "9", "--decimals", "-1"], check=True)
    files = [str(tmp_path / f"d_{s}.tns") for s in ("tr", "tr2", "tr3", "v", "te")]
    assert all(os.path.exists(f) for f in files)
    r = subprocess.run([exe, *files, "--alg", "als", "--rank", "8", "--lambda", "0.05", "--max-epochs", "5",
                        "--seed", "3", "--json"], check=True, capture_output=True, text=True)
    say("[cli]", r.stdout.strip().splitlines()[-1])
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    tr, va, te = (ct.Coo.read(f) for f in (files[0], files[3], files[4]))
    ctx = ct.Context(0)
    m = ct.Model(ctx, tr, None, None, rank=8)
    m.init_random(3)
    F = [f.astype(np.float64) for f in m.factors()]
    otr = ocoo(tr)
    ote = O.Coo(te.arrays()[0], te.arrays()[1], tr.dims)
    orders = [O.mode_order(otr, md) for md in range(3)]
    for _ in range(summary["epochs"]):
        for md in range(3):
            O.als_mode(otr, md, F, 0.05, order=orders[md], nthreads=4)
    assert summary["rmse_te"] == pytest.approx(O.rmse(ote, F), rel=2e-5)


def test_cli_reference_ccd_quirks(tmp_path):
    """`ctc ... --alg ccd --ref-quirks 11` (Q1, Q2 and the reference's CCD
    form Q4): the printed test RMSE against the oracle's restatement of the
    reference's CCD epochs (orc_ccd_jacobi_mode, modes 0, 1, 2) from the same
    init."""
    full, _ = ct.synth((300, 200, 60), 20000, zipf=(1.0, 0.7, 0.0), true_rank=4, value_map=ct.VALUES_AFFINE_1_5,
                       noise_std=0.1, test_fraction=0.0, seed=65)
    src = tmp_path / "t.tns"
    full.write(str(src), base=1, decimals=-1)
    exe = os.path.join(ROOT, "cutc_amd", "ctc")
    subprocess.run([exe, "prep", str(src), str(tmp_path / "d"), "--seed", "9", "--decimals", "-1"], check=True)
    files = [str(tmp_path / f"d_{s}.tns") for s in ("tr", "tr2", "tr3", "v", "te")]
    flags = ct.QUIRK_TILE_HEAD | ct.QUIRK_LAST_LANE | ct.QUIRK_NO_ADDBACK
    r = subprocess.run([exe, *files, "--alg", "ccd", "--rank", "8", "--lambda", "5", "--max-epochs", "4",
                        "--seed", "3", "--ref-quirks", str(flags), "--json"], check=True, capture_output=True,
                       text=True)
    say("[cli ccd quirks]", r.stdout.strip().splitlines()[-1])
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    tr, te = ct.Coo.read(files[0]), ct.Coo.read(files[4])
    m = ct.Model(ct.Context(0), tr, None, None, rank=8)
    m.init_random(3)
    F = [f.astype(np.float64) for f in m.factors()]
    otr = ocoo(tr)
    ote = O.Coo(te.arrays()[0], te.arrays()[1], tr.dims)
    for _ in range(summary["epochs"]):
        O.ccd_jacobi_epoch(otr, F, 5.0, flags)
    assert summary["rmse_te"] == pytest.approx(O.rmse(ote, F), rel=2e-5)


def test_sgd_bold_driver_schedule():
    """The bold driver (SG/sgd.cu:314-323): after each epoch from the second
    on, lr x1.05 if loss + lambda * sum ||F||^2 fell, else x0.5.  Per-epoch lr
    and validation RMSE of ctc_solve against the oracle's loop."""
    tr, va, te = problem(65)
    R, lam, lr = 16, 0.05, 2e-2
    ctx = ct.Context(0)
    m = ct.Model(ctx, tr, va, te, rank=R)
    rng = np.random.default_rng(7)
    F0 = [rng.uniform(-1, 1, (d, R)).astype(np.float32) for d in tr.dims]
    m.set_factors(F0)
    opt = ct.Options(algorithm=ct.SGD, rank=R, lambda_=lam, learning_rate=lr, lr_schedule=ct.LR_BOLD_DRIVER,
                     sgd_scale=1.0, max_epochs=10, max_bad_epochs=100, tile_len=512, sgd_substep=16,
                     sgd_hot_rows=300)
    res = m.solve(opt)
    hist, _, _, _ = oracle_solve(ct.SGD, tr, va, [f.astype(np.float64) for f in F0], lam, epochs=10, lr=lr,
                                 schedule=ct.LR_BOLD_DRIVER, sgd_shape=(512, 16, 300), max_bad=100)
    g_lr = [h["lr"] for h in res.history[1:11]]
    o_lr = [h[1] for h in hist[1:]]
    say("[bold] gpu lr", np.round(g_lr, 6), "\n        oracle  ", np.round(o_lr, 6))
    assert len(set(np.round(o_lr, 12))) > 2  # the schedule both grew and cut the rate
    np.testing.assert_allclose(g_lr, o_lr, rtol=1e-12)
    np.testing.assert_allclose([h["rmse_vl"] for h in res.history[1:11]], [h[0] for h in hist[1:]], rtol=1e-5)


def test_als_rotating_mode_order():
    """CTC_ORDER_ROTATE: each epoch starts at mode rot() % N of a
    std::mt19937_64 seeded with seed ^ 0xA15 (the reference rotates at random,
    SG/als.cu:684-685); replayed in the oracle, factors within 1e-5."""
    tr, va, te = problem(66)
    R, lam, seed = 16, 0.05, 11
    ctx = ct.Context(0)
    m = ct.Model(ctx, tr, va, te, rank=R)
    rng = np.random.default_rng(8)
    F0 = [rng.uniform(-1, 1, (d, R)).astype(np.float32) for d in tr.dims]
    m.set_factors(F0)
    res = m.solve(ct.Options(algorithm=ct.ALS, rank=R, lambda_=lam, max_epochs=5, max_bad_epochs=100,
                             mode_order=ct.ORDER_ROTATE, seed=seed))
    F = [f.astype(np.float64) for f in F0]
    starts = [MT19937_64(seed ^ 0xA15)() % 3]
    oracle_solve(ct.ALS, tr, va, F, lam, epochs=5, rotate_seed=seed, max_bad=100)
    say(f"[rotate] first start mode {starts[0]}")
    for md in range(3):
        assert relerr(m.factor(md), F[md]) < 1e-5
    assert res.epochs == 5


def test_sgd_autotune_restores_model_and_tuned_shape_matches_oracle():
    """ctc_autotune on the default SGD schedule (the GA's target,
    GA/fitness.c:67-79): the tuned (span, sub-step, block) is one of the
    searched shapes, the factors and the best-model copy are left as they were,
    and an epoch with the tuned shape matches the oracle's epoch of that
    schedule."""
    tr, va, te = problem(67, nnz=60000)
    R, lam = 32, 0.05
    ctx = ct.Context(0)
    m = ct.Model(ctx, tr, va, te, rank=R)
    rng = np.random.default_rng(9)
    F0 = [rng.uniform(-1, 1, (d, R)).astype(np.float32) for d in tr.dims]
    m.set_factors(F0)  # the best-model copy = F0
    opt = ct.Options(algorithm=ct.SGD, rank=R, lambda_=lam, sgd_scale=1.0)
    m.sgd_epoch(opt, lr=5e-3)  # the factors move, the best copy stays F0
    F1 = m.factors()
    assert any(not np.array_equal(a, b) for a, b in zip(F1, F0))
    t = m.autotune(opt)
    say(f"[sgd_autotune] {t}")
    assert t["tile_len"] in (1024, 2048, 4096, 8192) and t["sgd_substep"] in (16, 32, 64)
    assert t["tile_len"] >= 64 * t["sgd_substep"]  # at least 64 sub-steps per epoch
    assert t["block"] in (64, 256) and t["trials"] == 18  # 9 shapes x 2 blocks
    for md in range(3):
        np.testing.assert_array_equal(m.factor(md), F1[md])
        np.testing.assert_array_equal(m.factors(best=True)[md], F0[md])
    tuned = ct.Options(algorithm=ct.SGD, rank=R, lambda_=lam, sgd_scale=1.0, tile_len=t["tile_len"],
                       sgd_substep=t["sgd_substep"], block=t["block"])
    m.sgd_epoch(tuned, lr=5e-3)
    F = [f.astype(np.float64) for f in F1]
    otr = ocoo(tr)
    O.sgd_sync_epoch(otr, F, 5e-3, lam, t["tile_len"], t["sgd_substep"], tuned.to_c().sgd_hot_rows,
                     order=O.mode_order(otr, 0)[0], nthreads=4)
    for md in range(3):
        assert relerr(m.factor(md), F[md]) < 1e-5
