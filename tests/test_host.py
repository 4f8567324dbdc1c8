"""CPU tests of libctc's host side: the C ABI exports, the .tns reader, the
synthetic generator and the CISS builder (checked against the oracle)."""
import os
import re

import numpy as np
import pytest

import cutc_amd as ct
import oracle as O
from cutc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def header_symbols():
    src = open(os.path.join(ROOT, "include", "ctc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ctc_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_header_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(L, s), s
    # and the ctypes table covers the header exactly
    assert sorted(_lib.SIGNATURES) == syms


def test_status_strings():
    L = _lib.lib()
    for code in range(10):
        assert L.ctc_status_string(code)
    assert L.ctc_version().startswith(b"ctc")


@pytest.mark.parametrize("name", ["base1.tns", "base0.tns", "fourway.tns", "badbase.tns"])
def test_tns_reader_matches_oracle(name):
    path = os.path.join(GOLD, "tns", name)
    a = ct.tt_read(path)
    b = O.tns_read(path)
    assert a.dims == b.dims and a.nnz == b.nnz
    assert a.base == b.offset
    for m in range(a.nmodes):
        assert np.array_equal(a.ind(m), b.ind[m])
    assert np.array_equal(a.vals, b.vals)


def test_tns_reader_errors(tmp_path):
    with pytest.raises(ct.CtcError) as e:
        ct.tt_read(tmp_path / "missing.tns")
    assert e.value.status == 2
    p = tmp_path / "empty.tns"
    p.write_text("# nothing\n\n")
    with pytest.raises(ct.CtcError) as e:
        ct.tt_read(p)
    assert e.value.status == 3
    p = tmp_path / "bad.tns"
    p.write_text("1 2 3 4.0\n1 x 3 2.0\n")
    with pytest.raises(ct.CtcError) as e:
        ct.tt_read(p)
    assert e.value.status == 3


def test_tns_write_roundtrip(tmp_path):
    tr, _ = ct.synth((30, 20, 10), 500, test_fraction=0.0, seed=3)
    p = tmp_path / "x.tns"
    tr.write(p, base=1)
    b = ct.tt_read(p)
    assert b.base == (1, 1, 1)
    for m in range(3):
        assert np.array_equal(b.ind(m), tr.ind(m))
    assert np.array_equal(b.vals, tr.vals)
    tr.write(p, base=0, decimals=1)  # tvtensor.py-style %.1f
    c = ct.tt_read(p)
    assert np.allclose(c.vals, np.round(tr.vals, 1), atol=1e-12)


def _splatt_bin(path, iw, vw, dims, ind, vals):
    """A SPLATT binary coordinate file packed by hand from its published layout
    (int32 magic 0, uint64 index width, uint64 value width, then nmodes, dims,
    nnz, the index arrays and the values)."""
    import struct
    fi, fv = ("<I" if iw == 4 else "<Q"), ("<f" if vw == 4 else "<d")
    with open(path, "wb") as f:
        f.write(struct.pack("<iQQ", 0, iw, vw))
        for v in [len(dims), *dims, len(vals)]:
            f.write(struct.pack(fi, v))
        for arr in ind:
            for v in arr:
                f.write(struct.pack(fi, v))
        for v in vals:
            f.write(struct.pack(fv, v))


@pytest.mark.parametrize("iw,vw", [(4, 4), (8, 8), (4, 8), (8, 4)])
def test_binary_reader_known_answer(tmp_path, iw, vw):
    path = str(tmp_path / "t.bin")
    ind = [[1, 3, 3], [2, 1, 2], [1, 2, 2]]  # 1-based, as the reference stores them
    vals = [1.5, -2.25, 4.0]
    _splatt_bin(path, iw, vw, (3, 2, 2), ind, vals)
    t = ct.Coo.read(path)
    got_ind, got_vals = t.arrays()
    assert t.dims == (3, 2, 2) and t.nnz == 3 and t.base == (1, 1, 1)
    for m in range(3):
        assert list(got_ind[m]) == [v - 1 for v in ind[m]]
    assert list(got_vals) == vals


def test_binary_reader_zero_based_and_errors(tmp_path):
    path = str(tmp_path / "z.bin")
    _splatt_bin(path, 8, 8, (4, 3, 2), [[0, 3], [2, 0], [1, 0]], [0.1, 0.2])
    t = ct.Coo.read(path)
    assert t.base == (0, 0, 0) and t.dims == (4, 3, 2)
    assert [list(x) for x in t.arrays()[0]] == [[0, 3], [2, 0], [1, 0]]
    with open(path, "r+b") as f:  # truncate the values
        f.truncate(os.path.getsize(path) - 4)
    with pytest.raises(ct.CtcError):
        ct.Coo.read(path)
    bad = str(tmp_path / "bad.bin")
    with open(bad, "wb") as f:
        f.write(b"\x01\x00\x00\x00" + b"\x00" * 16)
    with pytest.raises(ct.CtcError):
        ct.Coo.read(bad)


def test_binary_roundtrip_widths(tmp_path):
    tr, _ = ct.synth((50, 40, 30), 2000, test_fraction=0.0, seed=8)
    path = str(tmp_path / "r.bin")
    tr.write_binary(path)
    back = ct.Coo.read(path)
    a, b = tr.arrays(), back.arrays()
    assert back.dims == tr.dims
    for m in range(3):
        assert np.array_equal(a[0][m], b[0][m])
    assert np.array_equal(a[1], b[1])
    # 4-byte indices and, for float-exact values, 4-byte values
    assert struct_header(path) == (0, 4, 8)
    ct.Coo.from_arrays(a[0], np.round(a[1] * 4) / 4, tr.dims).write_binary(path)
    assert struct_header(path) == (0, 4, 4)


def struct_header(path):
    import struct
    with open(path, "rb") as f:
        return struct.unpack("<iQQ", f.read(20))


def test_coo_sort_is_stable_lexicographic():
    rng = np.random.default_rng(3)
    n = 5000
    ind = [rng.integers(0, d, n) for d in (40, 30, 20)]
    vals = np.arange(n, dtype=np.float64)  # records the input position
    t = ct.Coo.from_arrays(ind, vals, (40, 30, 20))
    for keys in ((1, 2), (2, 0), (0, 1), (2,)):
        got_ind, got_vals = t.sorted(keys).arrays()
        # numpy's lexsort is stable; its last key is the most significant
        order = np.lexsort([ind[k] for k in reversed(keys)])
        assert np.array_equal(got_vals, vals[order])
        for m in range(3):
            assert np.array_equal(got_ind[m], ind[m][order])


def test_coo_split_fractions_cover_and_seed():
    tr, _ = ct.synth((300, 200, 100), 100000, test_fraction=0.0, seed=4)
    a, b, c = tr.split(seed=1)
    n = tr.nnz
    assert a.nnz + b.nnz + c.nnz == n
    assert abs(a.nnz / n - 0.8) < 0.01 and abs(b.nnz / n - 0.1) < 0.01 and abs(c.nnz / n - 0.1) < 0.01
    key = lambda t: set(zip(*[x.tolist() for x in t.arrays()[0]]))
    assert key(a) | key(b) | key(c) == key(tr)
    for part in (a, b, c):  # each part keeps the (mode 0, mode 1) order
        i0, i1 = part.arrays()[0][:2]
        assert np.all((np.diff(i0.astype(np.int64)) > 0) | ((np.diff(i0.astype(np.int64)) == 0) & (np.diff(i1.astype(np.int64)) >= 0)))
    a2, _, _ = tr.split(seed=1)
    a3, _, _ = tr.split(seed=2)
    assert np.array_equal(a.arrays()[1], a2.arrays()[1])  # same seed, same split
    same = a3.nnz == a.nnz and np.array_equal(a.arrays()[1], a3.arrays()[1])
    assert not same  # another seed, another split


def test_cli_prep_writes_reference_file_set(tmp_path):
    import subprocess
    src = str(tmp_path / "x.tns")
    tr, _ = ct.synth((60, 50, 40), 3000, test_fraction=0.0, value_map=ct.VALUES_INT_1_5, seed=6)
    tr.write(src, base=1, decimals=1)
    cli = os.path.join(os.path.dirname(ct.__file__), "ctc")
    r = subprocess.run([cli, "prep", src, str(tmp_path / "x"), "--seed", "3"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    parts = {k: ct.Coo.read(str(tmp_path / f"x_{k}.tns")) for k in ("tr", "v", "te", "tr2", "tr3")}
    assert parts["tr"].nnz + parts["v"].nnz + parts["te"].nnz == 3000
    assert parts["tr2"].nnz == parts["tr"].nnz == parts["tr3"].nnz
    i = parts["tr2"].arrays()[0]
    assert np.all(np.diff(i[1].astype(np.int64) * 100 + i[2]) >= 0)  # sorted by (1, 2)
    with open(str(tmp_path / "x_tr.tns")) as f:
        assert re.match(r"^\d+ \d+ \d+ \d\.\d$", f.readline().strip())  # %.1f values


def test_synth_is_seeded_distinct_and_exact():
    a, at = ct.synth((200, 150, 30), 20000, zipf=(1.0, 1.0, 0.0), true_rank=4,
                     value_map=ct.VALUES_INT_1_5, noise_std=0.3, test_fraction=0.05, seed=7)
    b, bt = ct.synth((200, 150, 30), 20000, zipf=(1.0, 1.0, 0.0), true_rank=4,
                     value_map=ct.VALUES_INT_1_5, noise_std=0.3, test_fraction=0.05, seed=7)
    assert a.nnz + at.nnz == 20000 and at.nnz == 1000
    for m in range(3):
        assert np.array_equal(a.ind(m), b.ind(m))
    assert np.array_equal(a.vals, b.vals)
    keys = np.concatenate([(a.ind(0).astype(np.int64) * 150 + a.ind(1)) * 30 + a.ind(2),
                           (at.ind(0).astype(np.int64) * 150 + at.ind(1)) * 30 + at.ind(2)])
    assert np.unique(keys).size == 20000
    ka = (a.ind(0).astype(np.int64) * 150 + a.ind(1)) * 30 + a.ind(2)
    assert np.all(np.diff(ka) > 0)  # train sorted by mode 0
    assert set(np.unique(a.vals)) <= {1.0, 2.0, 3.0, 4.0, 5.0}
    # Zipf skew on mode 0: row 0 is the heaviest
    c = np.bincount(a.ind(0), minlength=200)
    assert c[0] == c.max() and c[0] > 5 * np.median(c)
    d, _ = ct.synth((200, 150, 30), 20000, seed=8, test_fraction=0.0)
    assert not np.array_equal(d.ind(0), a.ind(0)) or not np.array_equal(d.ind(1), a.ind(1))


def test_synth_affine_range():
    tr, te = ct.synth((100, 80, 60), 5000, true_rank=8, value_map=ct.VALUES_AFFINE_1_5,
                      noise_std=0.0, test_fraction=0.1, seed=1)
    v = np.concatenate([tr.vals, te.vals])
    assert v.min() == pytest.approx(1.0) and v.max() == pytest.approx(5.0)


def _oracle_coo(t):
    ind, vals = t.arrays()
    return O.Coo(ind, vals, t.dims)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_ciss_native_structure(mode):
    tr, _ = ct.synth((300, 200, 50), 30000, zipf=(1.0, 0.5, 0.0), test_fraction=0.0, seed=11)
    c = ct.ciss_alloc(tr, mode)
    oc = _oracle_coo(tr)
    perm, rowptr = O.mode_order(oc, mode)
    assert np.array_equal(c.perm, perm)
    rows = tr.ind(mode)[c.perm]
    np.testing.assert_array_equal(c.directory, np.unique(rows))
    np.testing.assert_array_equal(np.diff(c.rowptr), np.bincount(rows)[np.unique(rows)])
    for s in range(2):
        np.testing.assert_array_equal(c.oidx(s), tr.ind((mode + 1 + s) % 3)[c.perm])
    np.testing.assert_array_equal(c.vals, tr.vals[c.perm].astype(np.float32))
    # tile headers: slot of each nonzero from (tile_slot, tile_mask) equals the
    # slot from the row pointers
    slot_true = np.repeat(np.arange(c.dlength), np.diff(c.rowptr))
    n = c.nnz
    heads = np.zeros(n, bool)
    heads[c.rowptr[:-1]] = True
    for t in range(c.ntiles):
        b = t * 64
        e = min(n, b + 64)
        mask = int(c.tile_mask[t])
        exp_mask = sum(1 << (p - b) for p in range(b, e) if heads[p])
        assert mask == exp_mask
        assert c.tile_slot[t] == slot_true[b]


@pytest.mark.parametrize("L", [1, 3, 8, 32, 52])
def test_ciss_reference_layout_equals_oracle(L):
    tr, _ = ct.synth((40, 30, 20), 2000, zipf=(1.0, 0.0, 0.0), test_fraction=0.0, seed=5)
    oc = _oracle_coo(tr)
    for mode in range(3):
        c = ct.ciss_alloc(tr, mode)
        d, dc, ent = c.reference_layout(L)
        od, odc, oent = O.ciss_reference_layout(O.sort_by_mode(oc, mode), mode, L)
        assert np.array_equal(d, od)
        assert np.array_equal(dc, odc)
        same = (ent == oent) | (np.isnan(ent) & np.isnan(oent))
        assert same.all()
        z = (oent == 0)
        assert np.array_equal(np.signbit(ent[z]), np.signbit(oent[z]))


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_ciss_reference_layout_hand_worked(idx):
    import json
    c = json.load(open(os.path.join(GOLD, "ciss_cases.json")))["cases"][idx]
    a = np.array(c["coo"], dtype=np.float64)
    t = ct.Coo.from_arrays([a[:, m].astype(np.int64) - 1 for m in range(3)], a[:, 3], c["dims"])
    d, dc, ent = ct.ciss_alloc(t, c["mode"]).reference_layout(c["tile_len"])
    assert list(d) == c["directory"] and list(dc) == c["dcounter"]
    exp = np.array([[np.nan if v is None else v for v in r] for r in c["entries"]])
    assert ((exp == ent) | (np.isnan(exp) & np.isnan(ent))).all()


def test_ciss_unsorted_input_is_sorted_internally():
    rng = np.random.default_rng(2)
    tr, _ = ct.synth((50, 40, 30), 3000, test_fraction=0.0, seed=2)
    ind, vals = tr.arrays()
    p = rng.permutation(tr.nnz)
    sh = ct.Coo.from_arrays([a[p] for a in ind], vals[p], tr.dims)
    for mode in range(3):
        a = ct.ciss_alloc(tr, mode)
        b = ct.ciss_alloc(sh, mode)
        assert np.array_equal(a.directory, b.directory)
        assert np.array_equal(a.oidx(0), b.oidx(0)) and np.array_equal(a.vals, b.vals)


def test_ciss_nnz_multiple_of_64_and_single_row():
    t = ct.Coo.from_arrays([np.zeros(128, np.int64), np.arange(128) % 16, np.arange(128) // 16],
                           np.arange(128, dtype=float), (1, 16, 8))
    c = ct.ciss_alloc(t, 0)
    assert c.ntiles == 2 and c.dlength == 1
    assert list(c.tile_slot) == [0, 0]
    assert list(c.tile_mask) == [1, 0]


def test_options_defaults_follow_reference():
    for alg, lam in ((ct.ALS, 10.0), (ct.SGD, 1e-3), (ct.CCD, 100.0)):
        o = ct.Options(algorithm=alg).to_c()
        assert o.lambda_ == lam and o.rank == 16 and o.tolerance == 1e-4 and o.max_bad_epochs == 3
    assert ct.Options(algorithm=ct.SGD).to_c().sgd_scale == 1e-5


def test_work_unit_defaults():
    """Measured defaults: SGD walkers of 2048 nonzeros in 32-nonzero sub-steps of
    the bulk-synchronous schedule, CCD++ work units of 1024; ALS 0, so that each
    ALS kernel family applies its own measured length; SGD hot rows above 65536
    nonzeros."""
    import ctypes as C
    from cutc_amd import _lib
    for alg, tl in ((ct.ALS, 0), (ct.SGD, 2048), (ct.CCD, 1024)):
        o = _lib.Options()
        assert _lib.lib().ctc_options_default(alg, C.byref(o)) == 0
        assert o.tile_len == tl, (alg, o.tile_len)
        if alg == ct.SGD:
            assert o.sgd_hot_rows == 65536 and o.sgd_substep == 32 and o.sgd_update == ct.SGD_SYNC


def test_context_without_gpu_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(ct.CtcError) as e:
        ct.Context(0)
    assert e.value.status in (6, 9)


@pytest.mark.parametrize("flags", [1, 2, 3, 16, 19])
def test_ref_quirk_weights_equal_oracle(flags):
    """The library's weights of the reference's tile-reduction defects
    (ctc_ref_quirk_weights) equal the oracle's, which the lane-by-lane walk
    of test_oracle.py pins, on a skewed tensor in every mode."""
    tr, _ = ct.synth((300, 120, 40), 30000, zipf=(1.0, 0.8, 0.0), true_rank=4, value_map=ct.VALUES_AFFINE_1_5,
                     noise_std=0.1, test_fraction=0.0, seed=17)
    ind, vals = tr.arrays()
    oc = O.Coo(ind, vals, tr.dims)
    for mode in range(3):
        w = tr.ref_quirk_weights(mode, flags)
        assert np.array_equal(w, O.quirk_weights(oc, mode, flags)), mode
        assert (w == 0).any() if flags & 17 else (w >= 1).all()
