"""Pins for the CPU oracle (oracle/): every routine is checked against something
that is not a retyping of its own formula — hand-worked fixtures, dense numpy /
scipy library routines, closed forms, invariants and finite differences."""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rand_coo(rng, dims, nnz):
    tot = int(np.prod(dims))
    lin = rng.choice(tot, size=nnz, replace=False)
    ind, rem = [], lin
    for d in reversed(dims):
        ind.append(rem % d)
        rem = rem // d
    ind = ind[::-1]
    vals = rng.normal(size=nnz) + 3.0
    return O.Coo(ind, vals, dims)


def dense(t: O.Coo):
    X = np.zeros(t.dims)
    M = np.zeros(t.dims, bool)
    X[tuple(t.ind)] = t.vals
    M[tuple(t.ind)] = True
    return X, M


def rand_factors(rng, dims, R, dtype=np.float64):
    return [rng.uniform(-1, 1, size=(d, R)).astype(dtype) for d in dims]


# --------------------------------------------------------------------------
# .tns reader (SG/io.c)
# --------------------------------------------------------------------------

def test_tns_one_based_with_comments_and_blank_lines():
    t = O.tns_read(os.path.join(GOLD, "tns", "base1.tns"))
    assert t.nmodes == 3 and t.nnz == 7
    assert t.dims == (4, 3, 3)
    assert t.offset == (1, 1, 1)
    assert list(t.ind[0]) == [0, 0, 0, 1, 1, 3, 3]
    assert list(t.ind[2]) == [0, 2, 0, 1, 1, 0, 2]
    assert list(t.vals) == [1.5, 2.5, 3.5, 4.5, 5.5, 1.25, 2.0]


def test_tns_zero_based_matches_one_based():
    a = O.tns_read(os.path.join(GOLD, "tns", "base1.tns"))
    b = O.tns_read(os.path.join(GOLD, "tns", "base0.tns"))
    assert b.offset == (0, 0, 0)
    assert a.dims == b.dims
    for m in range(3):
        assert np.array_equal(a.ind[m], b.ind[m])
    assert np.array_equal(a.vals, b.vals)


def test_tns_four_way():
    t = O.tns_read(os.path.join(GOLD, "tns", "fourway.tns"))
    assert t.nmodes == 4 and t.nnz == 3
    assert t.dims == (3, 3, 2, 4)
    assert list(t.vals) == [0.5, -1.5, 2.25]


def test_tns_base_above_one_reads_as_one_based_and_missing_file():
    # SG/io.c:263-281 starts the per-mode base at 1 and keeps the minimum, so a
    # smallest index of 2 leaves the base at 1: the exit(1) branch at :291-296
    # is unreachable and the file reads as 1-based with an empty first row.
    t = O.tns_read(os.path.join(GOLD, "tns", "badbase.tns"))
    assert t.offset == (1, 1, 1)
    assert t.dims == (3, 2, 2)
    assert list(t.ind[0]) == [1, 2]
    with pytest.raises(OSError) as e:
        O.tns_read(os.path.join(GOLD, "tns", "does_not_exist.tns"))
    assert e.value.args[0] == 2


# --------------------------------------------------------------------------
# CISS reference layout (SG/ciss.c)
# --------------------------------------------------------------------------

def _case_coo(c):
    a = np.array(c["coo"], dtype=np.float64)
    return O.Coo([a[:, 0].astype(np.int64) - 1, a[:, 1].astype(np.int64) - 1,
                  a[:, 2].astype(np.int64) - 1], a[:, 3], c["dims"])


def _same(expected, got):
    exp = np.array([[np.nan if v is None else v for v in r] for r in expected], dtype=np.float64)
    assert exp.shape == got.shape, (exp.shape, got.shape)
    both_nan = np.isnan(exp) & np.isnan(got)
    eq = (exp == got) | both_nan
    assert eq.all(), np.argwhere(~eq)
    # -0.0 in header[0] of a tile whose first slot is 0 (SG/ciss.c:139)
    zero = (exp == 0) & ~np.isnan(exp)
    assert np.array_equal(np.signbit(exp[zero]), np.signbit(got[zero]))


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_ciss_reference_layout_hand_worked(idx):
    cases = json.load(open(os.path.join(GOLD, "ciss_cases.json")))["cases"]
    c = cases[idx]
    t = _case_coo(c)
    d, dc, ent = O.ciss_reference_layout(t, c["mode"], c["tile_len"])
    assert list(d) == c["directory"]
    assert list(dc) == c["dcounter"]
    _same(c["entries"], ent)


def test_ciss_truncation_quirk():
    t = O.Coo([[0, 0], [0, 1], [0, 1]], [2.7, 4.2], (1, 2, 2))
    _, _, ent = O.ciss_reference_layout(t, 0, 4, truncate_vals=False)
    assert ent[1, 2] == 2.7 and ent[2, 2] == 4.2
    _, _, ent = O.ciss_reference_layout(t, 0, 4, truncate_vals=True)
    assert ent[1, 2] == 2.0 and ent[2, 2] == 4.0


def _decode_reference(bitmap, lane, f_id):
    """The kernels' decode (SG/als.cu:230-243) written with Python ints."""
    b = int("{:064b}".format(bitmap)[::-1], 2)  # __brevll
    while b & 1 == 0:
        b >>= 1
    b >>= 1
    mask = (b << (63 - lane)) & ((1 << 64) - 1)
    myfid = f_id + lane - bin(mask).count("1") + 1
    mybit = (b >> lane) & 1
    return myfid, mybit


@pytest.mark.parametrize("L", [1, 2, 3, 5, 8, 12])
def test_ciss_bitmap_decode_exhaustive(L):
    """Every continuation pattern of a full tile: the reference's popcount
    decode equals a naive count of row changes."""
    for pattern in range(1 << L):
        cont = [(pattern >> e) & 1 for e in range(L)]  # 1 = entry continues the row
        bitmap = 1
        for e in range(L):
            bitmap = (bitmap << 1) | cont[e]
        f0 = 5
        slot = f0
        for e in range(L):
            if not cont[e]:
                slot += 1
            fid, bit = _decode_reference(bitmap, e, f0)
            assert bit == cont[e]
            assert fid == slot


def test_ciss_bitmap_decode_random_L32():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        cont = rng.integers(0, 2, 32)
        bitmap = 1
        for e in range(32):
            bitmap = (bitmap << 1) | int(cont[e])
        slot = 11
        for e in range(32):
            slot += 0 if cont[e] else 1
            assert _decode_reference(bitmap, e, 11) == (slot, int(cont[e]))


def test_ciss_layout_random_roundtrip():
    """Decoding the emitted layout with the reference's own decode returns the
    sorted COO (directory slot -> row, records -> next indices and values)."""
    rng = np.random.default_rng(5)
    t = rand_coo(rng, (13, 7, 5), 150)
    for mode in range(3):
        ts = O.sort_by_mode(t, mode)
        for L in (1, 4, 7, 32):
            d, dc, ent = O.ciss_reference_layout(ts, mode, L)
            got = []
            n = ts.nnz
            for tile in range(n // L + 1):
                base = tile * (L + 1)
                hdr = ent[base]
                if tile * L >= n:
                    assert tuple(ent[base + 1]) == (-1, -1, -1)
                    continue
                f_id = int(-hdr[0])
                for e in range(L):
                    rec = ent[base + 1 + e]
                    if rec[0] == -1 and rec[1] == -1:
                        break
                    fid, _ = _decode_reference(int(hdr[2]), e, f_id)
                    got.append((int(d[fid]) - 1, int(rec[0]) - 1, int(rec[1]) - 1, rec[2]))
            m1, m2 = (mode + 1) % 3, (mode + 2) % 3
            exp = [(int(ts.ind[mode][p]), int(ts.ind[m1][p]), int(ts.ind[m2][p]), ts.vals[p])
                   for p in range(n)]
            assert got == exp
            # dcounter: last nonzero index of each slot
            rows = ts.ind[mode]
            last = np.flatnonzero(np.r_[rows[1:] != rows[:-1], True])
            assert list(dc[1:]) == list(last)


def test_mode_order_is_lexicographic():
    rng = np.random.default_rng(9)
    t = rand_coo(rng, (6, 5, 4), 60)
    for mode in range(3):
        perm, rowptr = O.mode_order(t, mode)
        keys = [tuple(int(t.ind[(mode + s) % 3][p]) for s in range(3)) for p in perm]
        assert keys == sorted(keys)
        counts = np.bincount(t.ind[mode], minlength=t.dims[mode])
        assert np.array_equal(np.diff(rowptr), counts)


# --------------------------------------------------------------------------
# MTTKRP / Gram vs dense numpy
# --------------------------------------------------------------------------

@pytest.mark.parametrize("nmodes", [3, 4])
def test_mttkrp_matches_dense_einsum(nmodes):
    rng = np.random.default_rng(11)
    dims = (5, 4, 6, 3)[:nmodes]
    t = rand_coo(rng, dims, 40)
    F = rand_factors(rng, dims, 7)
    X, _ = dense(t)
    letters = "ijkl"[:nmodes]
    for mode in range(nmodes):
        others = [m for m in range(nmodes) if m != mode]
        spec = letters + "," + ",".join(letters[m] + "r" for m in others) + "->" + letters[mode] + "r"
        exp = np.einsum(spec, X, *[F[m] for m in others])
        got = O.mttkrp(t, mode, F)
        np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12)


def test_gram_rows_matches_dense():
    rng = np.random.default_rng(12)
    dims = (4, 6, 5)
    t = rand_coo(rng, dims, 50)
    F = rand_factors(rng, dims, 5)
    _, M = dense(t)
    G = O.gram_rows(t, 1, F)
    for j in range(dims[1]):
        H = np.einsum("ik,ir,kr->ikr", M[:, j, :], F[0], F[2]).reshape(-1, 5)
        np.testing.assert_allclose(G[j], H.T @ H, rtol=1e-12, atol=1e-12)
        w = np.linalg.eigvalsh(G[j])
        assert w.min() > -1e-10  # symmetric PSD
        np.testing.assert_array_equal(G[j], G[j].T)


# --------------------------------------------------------------------------
# dense Cholesky vs LAPACK (scipy)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 16, 33])
def test_cholesky_matches_lapack(n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n))
    A = A @ A.T + n * np.eye(n)
    L = O.cholesky(A)
    Ls = scipy.linalg.cholesky(A, lower=True)
    np.testing.assert_allclose(L, Ls, rtol=1e-12, atol=1e-12)
    b = rng.normal(size=n)
    x = O.cholesky_solve(L, b)
    np.testing.assert_allclose(x, scipy.linalg.cho_solve((Ls, True), b), rtol=1e-10, atol=1e-12)


def test_cholesky_rejects_non_spd():
    A = np.array([[1.0, 2.0], [2.0, 1.0]])
    with pytest.raises(np.linalg.LinAlgError):
        O.cholesky(A)


# --------------------------------------------------------------------------
# ALS
# --------------------------------------------------------------------------

def test_als_mode_equals_per_row_ridge_lstsq():
    rng = np.random.default_rng(21)
    dims = (6, 7, 5)
    t = rand_coo(rng, dims, 90)
    R, lam = 4, 0.3
    F = rand_factors(rng, dims, R)
    X, M = dense(t)
    for mode in range(3):
        F0 = [f.copy() for f in F]
        assert O.als_mode(t, mode, F, lam) == 0
        others = [m for m in range(3) if m != mode]
        for i in range(dims[mode]):
            sel = np.flatnonzero(t.ind[mode] == i)
            if sel.size == 0:
                assert np.all(F[mode][i] == 0)
                continue
            H = F0[others[0]][t.ind[others[0]][sel]] * F0[others[1]][t.ind[others[1]][sel]]
            A = np.vstack([H, math.sqrt(lam) * np.eye(R)])
            b = np.r_[t.vals[sel], np.zeros(R)]
            x = np.linalg.lstsq(A, b, rcond=None)[0]
            np.testing.assert_allclose(F[mode][i], x, rtol=1e-9, atol=1e-10)
        F = [f.copy() for f in F]


def test_als_recovers_exact_rank_one():
    rng = np.random.default_rng(4)
    dims = (5, 4, 3)
    a, b, c = (rng.uniform(0.5, 1.5, d) for d in dims)
    X = np.einsum("i,j,k->ijk", a, b, c)
    idx = np.nonzero(np.ones(dims))
    t = O.Coo(list(idx), X[idx], dims)
    F = [rng.uniform(0.1, 1, (d, 1)) for d in dims]
    O.als(t, F, 1e-12, 1)
    P = np.einsum("ir,jr,kr->ijk", *F)
    np.testing.assert_allclose(P, X, rtol=1e-8)


# --------------------------------------------------------------------------
# BF16-operand ALS (mixed-precision path)
# --------------------------------------------------------------------------

def test_bf16_round_matches_torch_round_to_nearest_even():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    xs = (rng.standard_normal(3000) * 10.0 ** rng.integers(-6, 6, 3000)).astype(np.float32)
    xs = np.concatenate([xs, np.float32([1 + 2**-8, 1 + 3 * 2**-8, 1 + 2**-9, -1 - 2**-8, 0.0, 65504.0])])
    ours = np.array([O.bf16_round(x) for x in xs], np.float32)
    ref = torch.tensor(xs).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours, ref)
    assert O.bf16_round(1 + 2**-8) == 1.0          # tie -> even
    assert O.bf16_round(1 + 3 * 2**-8) == 1 + 2**-6  # tie -> even (upwards)


def test_als_mode_bf16_equals_fp64_on_bf16_exact_inputs():
    # factor entries k/8 (3 significant bits): every row and every product h is
    # exact in BF16, so the BF16-operand update must equal the FP64 one
    rng = np.random.default_rng(31)
    dims = (9, 8, 7)
    t = rand_coo(rng, dims, 200)
    R, lam = 8, 0.25
    for mode in range(3):
        F32 = [(rng.integers(-8, 9, (d, R)) / 8.0).astype(np.float32) for d in dims]
        F64 = [f.astype(np.float64) for f in F32]
        assert O.als_mode_bf16(t, mode, F32, lam) == 0
        assert O.als_mode(t, mode, F64, lam) == 0
        np.testing.assert_allclose(F32[mode], F64[mode].astype(np.float32), rtol=2e-7, atol=1e-7)


def test_als_rows_equals_full_mode_update():
    rng = np.random.default_rng(33)
    dims = (30, 20, 10)
    t = rand_coo(rng, dims, 1500)
    R = 8
    F32 = rand_factors(rng, dims, R, np.float32)
    rows = np.array([0, 5, 7, 29])
    sub = O.als_rows(t, 0, F32, 0.1, rows)
    F64 = [f.astype(np.float64) for f in F32]
    assert O.als_mode(t, 0, F64, 0.1) == 0
    np.testing.assert_array_equal(sub, F64[0][rows])


def test_als_mode_bf16_rounding_error_is_bf16_sized():
    rng = np.random.default_rng(32)
    dims = (10, 9, 8)
    t = rand_coo(rng, dims, 400)
    R, lam = 8, 0.1
    F32 = rand_factors(rng, dims, R, np.float32)
    F64 = [f.astype(np.float64) for f in F32]
    assert O.als_mode_bf16(t, 0, F32, lam) == 0
    assert O.als_mode(t, 0, F64, lam) == 0
    rel = np.linalg.norm(F32[0] - F64[0]) / np.linalg.norm(F64[0])
    assert 1e-5 < rel < 5e-2, rel  # rounding to 8 significant bits is visible but small


# --------------------------------------------------------------------------
# SGD
# --------------------------------------------------------------------------

def test_sgd_single_nonzero_closed_form():
    R = 3
    t = O.Coo([[1], [0], [2]], [2.0], (2, 1, 3))
    rng = np.random.default_rng(2)
    F = rand_factors(rng, (2, 1, 3), R)
    a, b, c = F[0][1].copy(), F[1][0].copy(), F[2][2].copy()
    lr, lam, s = 0.1, 0.01, 0.5
    e = 2.0 - np.sum(a * b * c)
    O.sgd_epoch(t, F, lr, lam, s)
    np.testing.assert_allclose(F[0][1], a + lr * s * (e * b * c - lam * a), rtol=1e-14)
    np.testing.assert_allclose(F[1][0], b + lr * s * (e * a * c - lam * b), rtol=1e-14)
    np.testing.assert_allclose(F[2][2], c + lr * s * (e * a * b - lam * c), rtol=1e-14)
    np.testing.assert_array_equal(F[0][0], rand_factors(np.random.default_rng(2), (2, 1, 3), R)[0][0])


def test_sgd_two_nonzeros_are_sequential():
    """The second update sees the first one's result on a shared row."""
    R = 2
    t = O.Coo([[0, 0], [0, 1], [0, 0]], [1.0, -1.0], (1, 2, 1))
    F = [np.array([[0.5, -0.2]]), np.array([[0.3, 0.1], [0.7, 0.4]]), np.array([[0.9, 0.6]])]
    G = [f.copy() for f in F]
    lr, lam = 0.2, 0.05
    O.sgd_epoch(t, F, lr, lam, 1.0)
    for p, j in ((0, 0), (1, 1)):
        a, b, c = G[0][0].copy(), G[1][j].copy(), G[2][0].copy()
        e = t.vals[p] - np.sum(a * b * c)
        G[0][0] = a + lr * (e * b * c - lam * a)
        G[1][j] = b + lr * (e * a * c - lam * b)
        G[2][0] = c + lr * (e * a * b - lam * c)
    for f, g in zip(F, G):
        np.testing.assert_allclose(f, g, rtol=1e-14)


def test_sgd_four_way_step():
    t = O.Coo([[0], [0], [0], [0]], [1.0], (1, 1, 1, 1))
    F = [np.array([[0.5, 0.2]]), np.array([[0.3, -0.1]]), np.array([[0.9, 0.4]]),
         np.array([[0.2, 0.8]])]
    a, b, c, d = (f[0].copy() for f in F)
    e = 1.0 - np.sum(a * b * c * d)
    O.sgd_epoch(t, F, 0.1, 0.0, 1.0)
    np.testing.assert_allclose(F[3][0], d + 0.1 * e * a * b * c, rtol=1e-14)


# --------------------------------------------------------------------------
# bulk-synchronous SGD schedule (libctc's default SGD)
# --------------------------------------------------------------------------
NO_HOT = 2**63


def _sgd_step(x, rows, lr, lam):
    """Reference step on one nonzero from the given rows (SG/sgd.cu:86-99):
    returns the deltas of every row."""
    prod = np.prod(rows, axis=0)
    e = x - np.sum(prod)
    out = []
    for m, r in enumerate(rows):
        others = np.prod([rows[o] for o in range(len(rows)) if o != m], axis=0)
        out.append(lr * (e * others - lam * r))
    return out


@pytest.mark.parametrize("hot", [NO_HOT, 0])
def test_sgd_sync_one_walker_unit_substeps_is_sequential(hot):
    """One walker over the whole stream with 1-nonzero sub-steps is the
    sequential rule, whatever the hot-row threshold (one run per sub-step: its
    sum and mean agree)."""
    rng = np.random.default_rng(41)
    dims = (9, 7, 5)
    t = rand_coo(rng, dims, 150)
    order, _ = O.mode_order(t, 0)
    F = rand_factors(rng, dims, 4)
    G = [f.copy() for f in F]
    O.sgd_epoch(t, F, 0.05, 0.1, 1.0, order=order)
    O.sgd_sync_epoch(t, G, 0.05, 0.1, span=t.nnz, sub=1, hot_rows=hot, order=order)
    for f, g in zip(F, G):
        np.testing.assert_allclose(g, f, rtol=1e-13, atol=1e-15)


def test_sgd_sync_walkers_on_disjoint_rows_are_sequential():
    """Walkers whose spans share no row of any mode do not interact: the
    schedule equals the sequential rule over the whole stream."""
    rng = np.random.default_rng(42)
    blk, nb, per = (4, 3, 5), 4, 20
    ind = [[], [], []]
    vals = []
    for b in range(nb):  # block b: rows [b*d, (b+1)*d) of every mode, 20 nonzeros
        lin = rng.choice(int(np.prod(blk)), size=per, replace=False)
        i0, r = np.divmod(lin, blk[1] * blk[2])
        i1, i2 = np.divmod(r, blk[2])
        for m, ii in enumerate((i0, i1, i2)):
            ind[m].extend((ii + b * blk[m]).tolist())
        vals.extend((rng.normal(size=per) + 3).tolist())
    dims = tuple(nb * d for d in blk)
    t = O.Coo(ind, vals, dims)
    order, _ = O.mode_order(t, 0)
    F = rand_factors(rng, dims, 3)
    G = [f.copy() for f in F]
    O.sgd_epoch(t, F, 0.03, 0.05, 1.0, order=order)
    O.sgd_sync_epoch(t, G, 0.03, 0.05, span=per, sub=1, hot_rows=NO_HOT, order=order)
    for f, g in zip(F, G):
        np.testing.assert_allclose(g, f, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("hot", [False, True])
def test_sgd_sync_shared_row_sum_or_mean(hot):
    """Two walkers, one nonzero each, sharing row 0 of mode 1: both steps read
    the sub-step's factors; the row takes the sum of the two deltas (cold) or
    their mean (hot: 2 nonzeros > hot_rows = 1 and > 4 * 2 / 10)."""
    rng = np.random.default_rng(43)
    dims, R = (2, 10, 2), 3
    t = O.Coo([[0, 1], [0, 0], [0, 1]], [2.0, -1.0], dims)
    F = rand_factors(rng, dims, R)
    F0 = [f.copy() for f in F]
    lr, lam = 0.1, 0.02
    d0 = _sgd_step(2.0, [F0[0][0], F0[1][0], F0[2][0]], lr, lam)
    d1 = _sgd_step(-1.0, [F0[0][1], F0[1][0], F0[2][1]], lr, lam)
    O.sgd_sync_epoch(t, F, lr, lam, span=1, sub=1, hot_rows=1 if hot else NO_HOT)
    np.testing.assert_allclose(F[0][0], F0[0][0] + d0[0], rtol=1e-14)
    np.testing.assert_allclose(F[0][1], F0[0][1] + d1[0], rtol=1e-14)
    np.testing.assert_allclose(F[2][1], F0[2][1] + d1[2], rtol=1e-14)
    shared = (d0[1] + d1[1]) / 2 if hot else d0[1] + d1[1]
    np.testing.assert_allclose(F[1][0], F0[1][0] + shared, rtol=1e-14)
    np.testing.assert_array_equal(F[1][1:], F0[1][1:])


def test_sgd_sync_runs_restart_from_the_substep_factors():
    """One walker, one 3-nonzero sub-step: the mode-0 row is one run (the
    walker's copy carries its updates), mode-2 row 0 is visited twice with
    another row between, so its second run starts again from the sub-step's
    factors and the row takes both runs' deltas."""
    rng = np.random.default_rng(44)
    dims, R = (1, 3, 2), 4
    t = O.Coo([[0, 0, 0], [0, 1, 2], [0, 1, 0]], [1.5, 2.5, 0.5], dims)
    F = rand_factors(rng, dims, R)
    F0 = [f.copy() for f in F]
    lr, lam = 0.07, 0.03
    a = F0[0][0].copy()
    da1 = _sgd_step(1.5, [a, F0[1][0], F0[2][0]], lr, lam)
    a = a + da1[0]
    da2 = _sgd_step(2.5, [a, F0[1][1], F0[2][1]], lr, lam)
    a = a + da2[0]
    da3 = _sgd_step(0.5, [a, F0[1][2], F0[2][0]], lr, lam)
    a = a + da3[0]
    O.sgd_sync_epoch(t, F, lr, lam, span=3, sub=3, hot_rows=NO_HOT)
    np.testing.assert_allclose(F[0][0], a, rtol=1e-14)
    np.testing.assert_allclose(F[2][0], F0[2][0] + da1[2] + da3[2], rtol=1e-14)
    np.testing.assert_allclose(F[2][1], F0[2][1] + da2[2], rtol=1e-14)
    for j, d in enumerate((da1, da2, da3)):
        np.testing.assert_allclose(F[1][j], F0[1][j] + d[1], rtol=1e-14)


def test_sgd_sync_threads_agree():
    """Walkers of a sub-step run in parallel: the result differs from the
    one-thread run by FP64 rounding only."""
    rng = np.random.default_rng(45)
    dims = (40, 30, 20)
    t = rand_coo(rng, dims, 3000)
    order, _ = O.mode_order(t, 0)
    F = rand_factors(rng, dims, 8)
    G = [f.copy() for f in F]
    O.sgd_sync_epoch(t, F, 0.01, 0.05, span=128, sub=16, hot_rows=100, order=order, nthreads=1)
    O.sgd_sync_epoch(t, G, 0.01, 0.05, span=128, sub=16, hot_rows=100, order=order, nthreads=4)
    for f, g in zip(F, G):
        np.testing.assert_allclose(g, f, rtol=1e-12, atol=1e-14)


def _bf16(x):
    """FP64 -> FP32 -> BF16 (round to nearest even) -> FP64, elementwise."""
    u = np.asarray(x, dtype=np.float64).astype(np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_bf16_rounding_known_values():
    """Round to nearest even at the BF16 half-way points, 8 significant bits."""
    one_ulp = 2.0 ** -7  # BF16 spacing in [1, 2)
    assert _bf16(1.0 + one_ulp / 2) == 1.0                   # tie -> even (1.0)
    assert _bf16(1.0 + 3 * one_ulp / 2) == 1.0 + 2 * one_ulp  # tie -> even (1 + 2 ulp)
    assert _bf16(1.0 + 0.6 * one_ulp) == 1.0 + one_ulp
    assert _bf16(-3.0) == -3.0 and _bf16(0.0) == 0.0
    assert O.bf16_round(1.0 + 3 * one_ulp / 2) == 1.0 + 2 * one_ulp  # the oracle's C rounding agrees


def test_sgd_sync_mixed_rounding_points():
    """The mixed-precision variant, one walker and one sub-step over two
    nonzeros that share their mode-0 and mode-1 rows: rows are read unrounded;
    each mode's run delta is accumulated unrounded along its run and, for modes
    1 and 2, rounded to BF16 once, when the run is handed in (mode 1: one run
    over both nonzeros; mode 2: two runs); mode 0's delta is never rounded.
    Worked step by step with numpy's own rounding."""
    rng = np.random.default_rng(47)
    dims, R = (1, 1, 2), 4
    t = O.Coo([[0, 0], [0, 0], [0, 1]], [2.5, -1.25], dims)
    F = rand_factors(rng, dims, R)
    F0 = [f.copy() for f in F]
    lr, lam = 0.3, 0.05
    # nonzero 0 from the sub-step's factors
    v0, v1, v2a = F0[0][0].copy(), F0[1][0].copy(), F0[2][0].copy()
    d = _sgd_step(2.5, [v0, v1, v2a], lr, lam)
    v0, v1 = v0 + d[0], v1 + d[1]
    del0, del1, del2a = d[0].copy(), d[1].copy(), d[2].copy()
    # nonzero 1: the mode-2 run ends (row 0 handed in), row 1 read
    v2b = F0[2][1].copy()
    d = _sgd_step(-1.25, [v0, v1, v2b], lr, lam)
    del0, del1, del2b = del0 + d[0], del1 + d[1], d[2]
    O.sgd_sync_epoch(t, F, lr, lam, span=2, sub=2, hot_rows=NO_HOT, mixed=True)
    np.testing.assert_allclose(F[0][0], F0[0][0] + del0, rtol=1e-14)
    np.testing.assert_allclose(F[1][0], F0[1][0] + _bf16(del1), rtol=1e-14)
    np.testing.assert_allclose(F[2][0], F0[2][0] + _bf16(del2a), rtol=1e-14)
    np.testing.assert_allclose(F[2][1], F0[2][1] + _bf16(del2b), rtol=1e-14)
    # the roundings matter in this case: the FP64 variant gives other values
    G = [f.copy() for f in F0]
    O.sgd_sync_epoch(t, G, lr, lam, span=2, sub=2, hot_rows=NO_HOT)
    assert np.abs(G[1][0] - F[1][0]).max() > 1e-6


def test_sgd_sync_mixed_is_fp64_on_bf16_exact_data():
    """Factors of BF16 numbers and a step (a power of two) that keeps every
    delta on the BF16 grid: the roundings change nothing, so the mixed variant
    equals the FP64 one."""
    dims, R = (3, 2, 2), 2
    t = O.Coo([[0, 1, 2], [0, 1, 1], [1, 0, 1]], [1.0, 2.0, 0.5], dims)
    F = [np.full((d, R), 0.5) for d in dims]
    G = [f.copy() for f in F]
    # lambda 0, unit values and rows of 0.5: err * others and step are powers of two
    O.sgd_sync_epoch(t, F, 2.0 ** -4, 0.0, span=1, sub=1, hot_rows=NO_HOT, mixed=True)
    O.sgd_sync_epoch(t, G, 2.0 ** -4, 0.0, span=1, sub=1, hot_rows=NO_HOT)
    for f, g in zip(F, G):
        np.testing.assert_array_equal(f, g)


# --------------------------------------------------------------------------
# CCD++
# --------------------------------------------------------------------------

def _objective(t, F, lam):
    return 0.5 * O.loss_sq(t, F) + 0.5 * O.frob_sq(t.dims, F, lam)


def test_ccd_last_column_is_stationary_and_objective_decreases():
    rng = np.random.default_rng(31)
    dims = (8, 6, 7)
    t = rand_coo(rng, dims, 200)
    R, lam = 3, 0.2
    F = rand_factors(rng, dims, R)
    res = O.residual(t, F)
    objs = [_objective(t, F, lam)]
    orders = [O.mode_order(t, m) for m in range(3)]
    for _ in range(4):
        O.ccd_epoch(t, F, lam, res, orders=orders)
        objs.append(_objective(t, F, lam))
        # the residual array tracks x - prediction
        np.testing.assert_allclose(res, O.residual(t, F), rtol=1e-9, atol=1e-9)
    assert all(b <= a + 1e-9 for a, b in zip(objs, objs[1:])), objs
    # the last coordinate block updated (mode 2, column R-1) minimises the
    # objective: its partial gradient is zero.
    _, G = O.grad(t, F, lam)
    assert np.max(np.abs(G[2][:, R - 1])) < 1e-9


def test_ccd_one_column_single_mode_is_ridge():
    """R=1, mode 0 update equals the 1-D ridge closed form from the add-back
    residual (which is x itself for a one-column model)."""
    rng = np.random.default_rng(8)
    dims = (4, 5, 3)
    t = rand_coo(rng, dims, 30)
    F = rand_factors(rng, dims, 1)
    F0 = [f.copy() for f in F]
    lam = 0.7
    res = O.residual(t, F)
    O.ccd_epoch(t, F, lam, res)
    for i in range(dims[0]):
        sel = np.flatnonzero(t.ind[0] == i)
        h = F0[1][t.ind[1][sel], 0] * F0[2][t.ind[2][sel], 0]
        exp = np.sum(t.vals[sel] * h) / (lam + np.sum(h * h))
        assert F[0][i, 0] == pytest.approx(exp, rel=1e-12, abs=1e-14)


# --------------------------------------------------------------------------
# gradient (adjoint) vs central finite differences
# --------------------------------------------------------------------------

@pytest.mark.parametrize("per_occ", [False, True])
@pytest.mark.parametrize("nmodes", [3, 4])
def test_grad_matches_finite_differences(per_occ, nmodes):
    rng = np.random.default_rng(40 + nmodes)
    dims = (5, 4, 6, 3)[:nmodes]
    t = rand_coo(rng, dims, 45)
    R, lam, eps = 3, 0.3, 1e-6
    F = rand_factors(rng, dims, R)
    loss, G = O.grad(t, F, lam, per_occ)

    def f(FF):
        return O.grad(t, FF, lam, per_occ)[0]

    for m in range(nmodes):
        for _ in range(6):
            i, r = rng.integers(dims[m]), rng.integers(R)
            Fp = [x.copy() for x in F]
            Fm = [x.copy() for x in F]
            Fp[m][i, r] += eps
            Fm[m][i, r] -= eps
            fd = (f(Fp) - f(Fm)) / (2 * eps)
            assert G[m][i, r] == pytest.approx(fd, rel=1e-6, abs=1e-7)


def test_grad_per_occurrence_is_summed_sgd_direction():
    """-grad with per-occurrence L2 equals the sum over nonzeros of the
    reference SGD step direction computed at the same point."""
    rng = np.random.default_rng(44)
    dims = (4, 3, 5)
    t = rand_coo(rng, dims, 25)
    R, lam = 2, 0.1
    F = rand_factors(rng, dims, R)
    _, G = O.grad(t, F, lam, True)
    acc = [np.zeros_like(f) for f in F]
    for p in range(t.nnz):
        rows = [F[m][t.ind[m][p]] for m in range(3)]
        e = t.vals[p] - np.sum(rows[0] * rows[1] * rows[2])
        for m in range(3):
            o = [rows[q] for q in range(3) if q != m]
            acc[m][t.ind[m][p]] += e * o[0] * o[1] - lam * rows[m]
    for m in range(3):
        np.testing.assert_allclose(-G[m], acc[m], rtol=1e-12, atol=1e-12)


# --------------------------------------------------------------------------
# loss / RMSE / convergence
# --------------------------------------------------------------------------

def test_loss_rmse_closed_form():
    t = O.Coo([[0, 1], [0, 0], [0, 1]], [3.0, 1.0], (2, 1, 2))
    F = [np.array([[1.0], [2.0]]), np.array([[1.0]]), np.array([[1.0], [0.25]])]
    # predictions: 1*1*1 = 1 and 2*1*0.25 = 0.5 -> errors 2 and 0.5
    assert O.loss_sq(t, F) == pytest.approx(4.25)
    assert O.rmse(t, F) == pytest.approx(math.sqrt(4.25 / 2))
    assert O.frob_sq(t.dims, F, 0.5) == pytest.approx(0.5 * (5.0 + 1.0 + 1.0625))


def test_converge_sequence():
    c = O.Converge(tolerance=1e-4, max_badepochs=3)
    seq = [1.0, 0.9, 0.89995, 0.8, 0.80001, 0.79999, 0.81]
    out = [c.step(v, e) for e, v in enumerate(seq)]
    assert [o[1] for o in out] == [True, True, False, True, False, False, False]
    assert [o[0] for o in out] == [False, False, False, False, False, False, True]
    assert c.s.bestepoch == 3 and c.s.best_rmse == 0.8


def test_f32_twin_tracks_f64():
    rng = np.random.default_rng(50)
    dims = (20, 15, 10)
    t = rand_coo(rng, dims, 600)
    F64 = rand_factors(rng, dims, 4)
    F32 = [f.astype(np.float32) for f in F64]
    O.als(t, F64, 0.5, 2)
    O.als(t, F32, 0.5, 2)
    for a, b in zip(F64, F32):
        assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-4


def test_mt19937_64_standard_check_value():
    """The replay of libctc's seeded mode rotation uses a Python
    std::mt19937_64; the C++ standard requires the 10000th output of a
    default-constructed engine to be 9981545732273789042."""
    from mt64 import MT19937_64
    g = MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


# --------------------------------------------------------------------------
# CSF layout and its MTTKRP (SG/csf.c, SG/mttkrp.c)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_csf_expands_back_to_the_tensor(mode):
    """Level order (i,j,k) / (j,i,k) / (k,i,j) (SG/csf.c:30-34); walking the
    tree gives back every (index, value) of the COO exactly once, level-0 ids
    strictly increase, and fiber ids strictly increase within their node."""
    rng = np.random.default_rng(50 + mode)
    dims = (17, 13, 11)
    t = rand_coo(rng, dims, 600)
    c = O.Csf(t, mode)
    lv = c.level_mode
    assert lv == ((0, 1, 2), (1, 0, 2), (2, 0, 1))[mode]
    f0, p0, f1, p1, f2, v = c.fids(0), c.fptr(0), c.fids(1), c.fptr(1), c.fids(2), c.vals
    assert np.all(np.diff(f0.astype(np.int64)) > 0)
    got = set()
    for a in range(c.nfibs[0]):
        fib = f1[p0[a]:p0[a + 1]]
        assert len(fib) > 0 and np.all(np.diff(fib.astype(np.int64)) > 0)
        for f in range(p0[a], p0[a + 1]):
            assert p1[f + 1] > p1[f]
            for q in range(p1[f], p1[f + 1]):
                idx = [0, 0, 0]
                idx[lv[0]], idx[lv[1]], idx[lv[2]] = int(f0[a]), int(f1[f]), int(f2[q])
                got.add((*idx, float(v[q])))
    want = {(int(t.ind[0][p]), int(t.ind[1][p]), int(t.ind[2][p]), float(t.vals[p])) for p in range(t.nnz)}
    assert got == want and c.nfibs[2] == t.nnz


@pytest.mark.parametrize("n0,expect", [(5, [0, 1, 2, 3, 4, 5] + [5] * 59),
                                       (130, [3 * g for g in range(44)] + [130] * 21)])
def test_csf_blockref_groups(n0, expect):
    """64 groups of ceil(n0 / 64) level-0 nodes (SG/csf.c:75-92): n0 = 5 gives
    one node per group; n0 = 130 groups of 3 (44 groups, the last partial)."""
    t = O.Coo([np.arange(n0), np.zeros(n0, int), np.zeros(n0, int)], np.ones(n0), (n0, 1, 1))
    assert list(O.Csf(t, 0).blockref) == expect


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_mttkrp_csf_equals_coo_mttkrp(mode):
    rng = np.random.default_rng(60 + mode)
    dims = (40, 30, 20)
    t = rand_coo(rng, dims, 3000)
    F = rand_factors(rng, dims, 8)
    c = O.Csf(t, mode)
    perm, rowptr = O.mode_order(t, mode)
    ref = O.mttkrp(t, mode, F, order=(perm, rowptr))
    np.testing.assert_allclose(O.mttkrp_csf(c, F, nthreads=4), ref, rtol=1e-12, atol=1e-12)
    F32 = [f.astype(np.float32) for f in F]
    np.testing.assert_allclose(O.mttkrp_csf(c, F32, nthreads=1), ref, rtol=2e-5, atol=2e-5)


# --------------------------------------------------------------------------
# The reference's tile-reduction defects (quirks Q1, Q2) as weights
# --------------------------------------------------------------------------
def _reference_tile_walk(rows, terms, q1, q2, L=32):
    """The reference's warp-per-tile segmented reduction (SG/als.cu:226-312)
    run lane by lane: per tile of L consecutive nonzeros, mybit = the lane
    continues the previous nonzero's row (lane 0: forced to 1 under Q1, else
    a head), mylbit = mybit except lane 0's 0; itercounter rounds of
    t = shfl_down(nt, 1) + (1 - mylbit) t, nt = mybit t, where the last lane
    of a full tile reads its own nt under Q2 (0 otherwise) and the last lane of
    a partial tile reads 0; lanes with mybit = 0 add t to their row."""
    out = {}
    n = len(rows)
    for b in range(0, n, L):
        nl = min(L, n - b)
        cont = [int(b + l > 0 and rows[b + l] == rows[b + l - 1]) for l in range(nl)]
        iters = sum(cont[1:])
        mybit, mylbit = list(cont), list(cont)
        mybit[0], mylbit[0] = (1, 0) if q1 else (0, 0)
        t = [terms[b + l] for l in range(nl)]
        nt = [mybit[l] * t[l] for l in range(nl)]
        for _ in range(iters):
            inc = [nt[l + 1] if l + 1 < nl else (nt[l] if (q2 and nl == L) else 0) for l in range(nl)]
            t = [inc[l] + (1 - mylbit[l]) * t[l] for l in range(nl)]
            nt = [mybit[l] * t[l] for l in range(nl)]
        for l in range(nl):
            if not mybit[l]:
                out[rows[b + l]] = out.get(rows[b + l], 0) + t[l]
    return out


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_quirk_weights_match_the_lane_by_lane_walk(flags):
    """Per-row sums of integer terms weighted by orc_quirk_weights equal the
    lane-by-lane simulation of the reference's reduction, for every flag
    combination, on tensors whose rows run across tiles, end exactly at lane
    31, fill whole tiles, and end in a partial tile."""
    rng = np.random.default_rng(70 + flags)
    q1, q2 = bool(flags & 1), bool(flags & 2)
    for trial in range(40):
        dims = (int(rng.integers(3, 40)), 7, 5)
        lens = rng.geometric(1.0 / rng.choice([2, 8, 40]), size=dims[0])
        lens[rng.random(dims[0]) < 0.2] = 0
        rows = np.repeat(np.arange(dims[0]), lens)
        n = len(rows)
        if n == 0:
            continue
        t = O.Coo([rows, rng.integers(0, 7, n), rng.integers(0, 5, n)], np.ones(n), dims)
        perm, _ = O.mode_order(t, 0)
        terms = rng.integers(-50, 50, n)  # by nonzero position in t
        w = O.quirk_weights(t, 0, flags)
        want = _reference_tile_walk([int(t.ind[0][p]) for p in perm], [int(terms[p]) for p in perm], q1, q2)
        got = {}
        for p in range(n):
            if w[p]:
                got[int(t.ind[0][p])] = got.get(int(t.ind[0][p]), 0) + int(w[p]) * int(terms[p])
        want = {k: v for k, v in want.items() if v != 0}
        got = {k: v for k, v in got.items() if v != 0}
        assert got == want, (trial, flags)
    if flags == 0:
        assert np.all(w == 1)


def test_quirk_weights_hand_worked_tile():
    """One full tile: lanes 0-4 on row 1 (lane 0 continuing row 1 from the
    previous tile), lanes 5-31 on row 2.  Q1 drops lanes 0-4; under Q2 the
    continuation lane 31 (segment head 5) is re-added itercounter - (31 - 5)
    = (4 + 26) - 26 = 4 extra times, weight 5.  The previous tile (row 0 on
    lanes 0-30, row 1 on lane 31) has a single-lane head at 31: no extra."""
    rows = np.array([0] * 31 + [1] + [1] * 5 + [2] * 27)
    n = len(rows)
    # mode-1 ids increase with the position, so the mode-0 order is t's order
    t = O.Coo([rows, np.arange(n), np.zeros(n, int)], np.ones(n), (3, n, 1))
    w = O.quirk_weights(t, 0, 3)
    assert list(w[:32]) == [0] * 31 + [1]
    assert list(w[32:37]) == [0] * 5
    assert list(w[37:63]) == [1] * 26 and w[63] == 5
    assert list(O.quirk_weights(t, 0, 1)[32:]) == [0] * 5 + [1] * 27
    assert O.quirk_weights(t, 0, 2)[63] == 5


def test_quirk_multiset_truncates_and_repeats():
    t = O.Coo([[0, 0, 1], [0, 1, 0], [0, 0, 0]], [2.7, 3.2, -1.5], (2, 2, 1))
    q = O.quirk_multiset(t, 0, O.QUIRK_TRUNC_VALUES)
    assert list(q.vals) == [2.0, 3.0, -1.0] and q.nnz == 3
    q = O.quirk_multiset(t, 0, O.QUIRK_TILE_HEAD)  # one partial tile: its first segment (row 0) dropped
    assert q.nnz == 1 and list(q.ind[0]) == [1]


def test_quirk_ccd_grid_weights():
    """Q3: the CCD launch has tilenum / 256 + 1 blocks of 8 warps, one tile of
    32 each, tilenum = nnz / 32 + 1 (SG/ccd.cu:520).  1000 nonzeros: 32 tiles,
    one block, 8 tiles = the first 256 nonzeros of the mode order; 100
    nonzeros: all 4 tiles covered; 70000: 2188 tiles, 9 blocks, 72 tiles."""
    rng = np.random.default_rng(12)
    for n, kept in ((1000, 256), (100, 100), (70000, 72 * 32)):
        t = O.Coo([rng.integers(0, 50, n), rng.integers(0, 9, n), rng.integers(0, 4, n)], np.ones(n), (50, 9, 4))
        perm, _ = O.mode_order(t, 0)
        w = O.quirk_weights(t, 0, O.QUIRK_CCD_GRID)
        assert list(w[perm]) == [1] * kept + [0] * (n - kept), n
        # with Q1 as well, both apply: the dropped tiles stay dropped
        w2 = O.quirk_weights(t, 0, O.QUIRK_CCD_GRID | O.QUIRK_TILE_HEAD)
        assert np.all(w2[perm][kept:] == 0) and np.array_equal(w2[perm][:kept], O.quirk_weights(t, 0, 1)[perm][:kept])


def test_ccd_jacobi_hand_worked():
    """R = 1, row 0 holds (0,0,0) = 3 and (0,1,0) = 5, B = [1, 2], C = [1],
    a = 0.5, lambda = 1: h = (1, 2), e = (3 - 0.5, 5 - 1) = (2.5, 4), num =
    2.5 + 8 = 10.5, den = 1 + 4 = 5, a <- 10.5 / 6 = 1.75.  Row 1 has no
    nonzero and keeps its value; with Q1 the single partial tile's first (and
    only) segment is dropped, so row 0 becomes 0 / (1 + 0) = 0."""
    t = O.Coo([[0, 0], [0, 1], [0, 0]], [3.0, 5.0], (2, 2, 1))
    F = [np.array([[0.5], [7.0]]), np.array([[1.0], [2.0]]), np.array([[1.0]])]
    O.ccd_jacobi_mode(t, 0, F, 1.0)
    assert F[0][0, 0] == 1.75 and F[0][1, 0] == 7.0
    F[0][0, 0] = 0.5
    O.ccd_jacobi_mode(t, 0, F, 1.0, O.QUIRK_TILE_HEAD)
    assert F[0][0, 0] == 0.0 and F[0][1, 0] == 7.0


def _reference_ccd_mode(t, mode, F, lam, q1, q2, q3, trunc):
    """update_frac_gpu_as + update_ccd_gpu (SG/ccd.cu:82-206,318-348) run lane
    by lane: in the mode's sorted order each lane computes h and localtbuffer[2]
    = x - <a_row, h> from the factors before the update; per column m the
    terms h_m^2 and h_m * e go through the tile walk; the launch covers
    tilenum / 256 + 1 blocks of 8 tiles under Q3 (else every tile); then every
    row with a nonzero takes num / (lambda + den), 0 / lambda if nothing
    reached it."""
    perm, _ = O.mode_order(t, mode)
    n, R = t.nnz, F[0].shape[1]
    others = [m for m in range(t.nmodes) if m != mode]
    rows = [int(t.ind[mode][p]) for p in perm]
    h = np.ones((n, R))
    for m in others:
        h *= F[m][np.asarray(t.ind[m])[perm]]
    x = np.asarray(t.vals, np.float64)[perm]
    if trunc:
        x = np.trunc(x)
    e = x - np.einsum("pr,pr->p", F[mode][rows], h)
    if q3:
        covered = ((n // 32 + 1) // 256 + 1) * 8 * 32
        rows, h, e = rows[:covered], h[:covered], e[:covered]
    new = F[mode].copy()
    for r in set(int(i) for i in t.ind[mode]):
        new[r] = 0.0
    for m in range(R):
        den = _reference_tile_walk(rows, list(h[:, m] * h[:, m]), q1, q2)
        num = _reference_tile_walk(rows, list(h[:, m] * e), q1, q2)
        for r in set(int(i) for i in t.ind[mode]):
            new[r, m] = num.get(r, 0.0) / (lam + den.get(r, 0.0))
    F[mode][:] = new


@pytest.mark.parametrize("flags", [0, 1, 2, 3, 7, 16, 19, 23])
def test_ccd_jacobi_matches_the_lane_by_lane_kernel(flags):
    """orc_ccd_jacobi_mode with the quirk weights equals the lane-by-lane
    simulation of the reference's CCD kernels for every combination of Q1,
    Q2, Q3, Q5, on tensors with rows across tiles and more tiles than the
    CCD grid covers, every mode, to 1e-12."""
    rng = np.random.default_rng(300 + flags)
    q1, q2, q3, tr = bool(flags & 1), bool(flags & 2), bool(flags & 16), bool(flags & 4)
    for trial in range(3):
        dims = (int(rng.integers(5, 30)), int(rng.integers(4, 12)), 6)
        n = int(rng.choice([70, 400, 9000]))
        lens = rng.multinomial(n, rng.dirichlet(np.ones(dims[0]) * 0.3))
        rows = np.repeat(np.arange(dims[0]), lens)
        t = O.Coo([rows, rng.integers(0, dims[1], n), rng.integers(0, dims[2], n)],
                  rng.normal(2.0, 1.5, n), dims)
        R = 3
        F0 = [rng.uniform(-1, 1, (d, R)) for d in dims]
        for mode in range(3):
            want = [f.copy() for f in F0]
            _reference_ccd_mode(t, mode, want, 2.0, q1, q2, q3, tr)
            got = [f.copy() for f in F0]
            O.ccd_jacobi_mode(t, mode, got, 2.0, flags | O.QUIRK_NO_ADDBACK)
            np.testing.assert_allclose(got[mode], want[mode], rtol=1e-12, atol=1e-12)
            for m in range(3):
                if m != mode:
                    assert np.array_equal(got[m], F0[m])


def test_ccd_jacobi_is_a_ridge_step_on_the_residual():
    """Without weights the update equals (m - G a) / (lambda + diag G) per row,
    with m the MTTKRP and G the Gram of the oracle's ALS path (the row's own
    value enters through G a: no add-back)."""
    rng = np.random.default_rng(8)
    n, dims, R = 3000, (40, 30, 20), 4
    t = O.Coo([rng.integers(0, d, n) for d in dims], rng.normal(0, 1, n), dims)
    F = [rng.uniform(-1, 1, (d, R)) for d in dims]
    mode = 1
    perm, rowptr = O.mode_order(t, mode)
    m = O.mttkrp(t, mode, F)
    G = O.gram_rows(t, mode, F)
    a = F[mode]
    want = a.copy()
    has = np.diff(rowptr.astype(np.int64)) > 0
    ga = np.einsum("irs,is->ir", G, a)
    want[has] = (m[has] - ga[has]) / (0.7 + np.einsum("irr->ir", G)[has])
    O.ccd_jacobi_mode(t, mode, F, 0.7)
    np.testing.assert_allclose(F[mode], want, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("per_occ", [False, True])
def test_grad_rows_equal_the_full_gradient(per_occ):
    """orc_grad_rows_f64 (sampled rows of a mode's gradient, for full-size
    checks) equals those rows of orc_grad_f64 from the same float factors,
    for every mode of a 4-way tensor, rows with and without nonzeros."""
    rng = np.random.default_rng(21)
    dims, n, R = (40, 20, 10, 4), 3000, 8
    # mode-0 rows 30..39 hold no nonzero
    t = O.Coo([rng.integers(0, min(d, 30), n) for d in dims], rng.normal(0, 1, n), dims)
    F32 = [rng.uniform(-1, 1, (d, R)).astype(np.float32) for d in dims]
    _, G = O.grad(t, [f.astype(np.float64) for f in F32], 0.3, per_occ)
    for mode in range(4):
        rows = np.arange(dims[mode], dtype=np.uint32)
        got = O.grad_rows(t, mode, F32, 0.3, rows, per_occ)
        np.testing.assert_allclose(got, G[mode], rtol=1e-12, atol=1e-12)
