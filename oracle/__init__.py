"""CPU oracle for the tensor-completion hot path — TEST INFRASTRUCTURE ONLY.

This package restates the reference's update rules on the CPU (FP64, plus an
FP32 twin used as the timed CPU baseline).  It may be imported by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` only;
the product (``cutc_amd``) never imports it.

Reference files restated (SG = ``software/single GPU/sgpu``):
  SG/io.c:58-106,234-308     .tns reader           -> tns_read
  SG/ciss.c:68-195           CISS layout           -> ciss_reference_layout
  SG/als.cu:186-317,467-491  ALS mode update       -> als_mode / als
  SG/sgd.cu:21-143,264-324   SGD epoch             -> sgd_epoch / sgd
  SG/sgd.cu:21-143 + NC/sgd.cu:229-262,523-525 (walker schedule, replica
                             averaging)               -> sgd_sync_epoch / sgd_sync
  SG/ccd.cu:532-554, NC/ccd.cu:28-180  CCD++       -> ccd_epoch / ccd
  SG/completion.c:123-214,296-362      loss/RMSE   -> loss_sq / rmse / converge
  SG/csf.c:11-331, SG/mttkrp.c:16-93   CSF + serial MTTKRP -> Csf / mttkrp_csf

Pins (each function is checked against something it is not a retyping of):
  tns_read                 hand-written files (tests/golden/tns/*)
  ciss_reference_layout    hand-worked layouts (tests/golden/ciss_*.json) and a
                           brute-force decode of every bitmap
  mttkrp / gram_rows       dense numpy einsum over the densified tensor
  csf_build / mttkrp_csf   the CSF expands back to the COO (round trip), its levels are
                           strictly increasing within their parent, blockref by hand;
                           mttkrp_csf equals mttkrp (itself pinned by einsum)
  cholesky / solve         scipy.linalg.cho_factor / cho_solve
  als_mode                 per-row ridge regression via numpy.linalg.lstsq on the
                           stacked [H; sqrt(lambda) I] system; exact rank-1 recovery
  sgd_epoch                single-nonzero closed form; two-nonzero sequential form
  sgd_sync_epoch           equals sgd_epoch with one walker and 1-nonzero sub-steps,
                           and with walkers on disjoint rows; hand-computed sums and
                           means of snapshot steps on a shared row
  ccd_epoch                per-column 1-D ridge optimality (zero partial gradient)
                           and monotone decrease of the objective
  grad                     central finite differences in FP64
  loss_sq / converge       closed forms
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
MAX_MODES = 4


class OrcCoo(C.Structure):
    _fields_ = [
        ("nmodes", C.c_uint32),
        ("nnz", C.c_uint64),
        ("dims", C.c_uint64 * MAX_MODES),
        ("ind", C.POINTER(C.c_uint32) * MAX_MODES),
        ("vals", C.POINTER(C.c_double)),
        ("offset", C.c_uint64 * MAX_MODES),
    ]


class OrcCsf(C.Structure):
    _fields_ = [
        ("mode", C.c_uint32),
        ("level_mode", C.c_uint32 * 3),
        ("nnz", C.c_uint64),
        ("dims", C.c_uint64 * 3),
        ("nfibs", C.c_uint64 * 3),
        ("fptr", C.POINTER(C.c_uint64) * 2),
        ("fids", C.POINTER(C.c_uint32) * 3),
        ("vals", C.POINTER(C.c_double)),
        ("blockref", C.c_uint64 * 65),
    ]


class OrcConv(C.Structure):
    _fields_ = [
        ("best_rmse", C.c_double),
        ("tolerance", C.c_double),
        ("nbadepochs", C.c_uint64),
        ("bestepoch", C.c_uint64),
        ("max_badepochs", C.c_uint64),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: run `make -C {_HERE}`")
        _LIB = C.CDLL(path)
    return _LIB


@dataclass
class Coo:
    """COO tensor with 0-based uint32 indices and float64 values."""

    ind: list
    vals: np.ndarray
    dims: tuple
    offset: tuple = field(default=None)

    @property
    def nmodes(self):
        return len(self.ind)

    @property
    def nnz(self):
        return int(self.vals.shape[0])

    def __post_init__(self):
        self.ind = [np.ascontiguousarray(a, dtype=np.uint32) for a in self.ind]
        self.vals = np.ascontiguousarray(self.vals, dtype=np.float64)
        self.dims = tuple(int(d) for d in self.dims)
        if self.offset is None:
            self.offset = tuple(0 for _ in self.ind)
        self._c = None

    def c(self):
        s = OrcCoo()
        s.nmodes = self.nmodes
        s.nnz = self.nnz
        for m in range(self.nmodes):
            s.dims[m] = self.dims[m]
            s.ind[m] = self.ind[m].ctypes.data_as(C.POINTER(C.c_uint32))
            s.offset[m] = self.offset[m]
        s.vals = self.vals.ctypes.data_as(C.POINTER(C.c_double))
        self._c = s
        return C.byref(s)

    def permuted(self, perm):
        return Coo([a[perm] for a in self.ind], self.vals[perm], self.dims, self.offset)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _u64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def _ptrs(arrs, kind):
    T = C.POINTER(C.c_double) if kind == "f64" else C.POINTER(C.c_float)
    arr = (T * MAX_MODES)()
    for i, a in enumerate(arrs):
        arr[i] = a.ctypes.data_as(T)
    return arr


def _prec(F):
    return "f64" if F[0].dtype == np.float64 else "f32"


def tns_read(path):
    s = OrcCoo()
    st = lib().orc_tns_read(path.encode(), C.byref(s))
    if st != 0:
        raise OSError(st, f"orc_tns_read({path}) failed with status {st}")
    try:
        n = s.nnz
        ind = [np.ctypeslib.as_array(s.ind[m], shape=(n,)).copy() for m in range(s.nmodes)]
        vals = np.ctypeslib.as_array(s.vals, shape=(n,)).copy()
        return Coo(ind, vals, tuple(s.dims[m] for m in range(s.nmodes)),
                   tuple(s.offset[m] for m in range(s.nmodes)))
    finally:
        lib().orc_coo_free(C.byref(s))


def mode_order(t: Coo, mode: int):
    perm = np.empty(t.nnz, np.uint32)
    rowptr = np.empty(t.dims[mode] + 1, np.uint64)
    st = lib().orc_mode_order(t.c(), C.c_uint32(mode), _u32p(perm), _u64p(rowptr))
    assert st == 0
    return perm, rowptr


QUIRK_TILE_HEAD = 1     # Q1: a tile's first row segment is not written
QUIRK_LAST_LANE = 2     # Q2: the last lane's term re-added in a full tile
QUIRK_TRUNC_VALUES = 4  # Q5: values stored through an integer (SG/ciss.c:123,157,176)
QUIRK_NO_ADDBACK = 8    # Q4: the reference's CCD update form (SG/ccd.cu:157-195,318-348)
QUIRK_CCD_GRID = 16     # Q3: only the tiles the CCD launch covers (SG/ccd.cu:520)
_WEIGHT_BITS = QUIRK_TILE_HEAD | QUIRK_LAST_LANE | QUIRK_CCD_GRID


def quirk_weights(t: Coo, mode: int, flags: int, tile_len: int = 32):
    """Per-nonzero weights (indexed like t) of the reference's tile-reduction
    defects Q1 / Q2 in the mode-`mode` walk (SG/als.cu:226-312) and of the
    CCD grid Q3 (SG/ccd.cu:520)."""
    perm, _ = mode_order(t, mode)
    w = np.empty(t.nnz, np.uint32)
    st = lib().orc_quirk_weights(t.c(), C.c_uint32(mode), _u32p(perm), C.c_uint32(tile_len),
                                 C.c_uint32(flags), _u32p(w))
    assert st == 0, st
    return w


def quirk_multiset(t: Coo, mode: int, flags: int, tile_len: int = 32) -> Coo:
    """The tensor whose plain ALS update of `mode` equals the reference's
    update with the given defects: each nonzero repeated by its weight (0
    drops it), values truncated toward zero under QUIRK_TRUNC_VALUES."""
    w = quirk_weights(t, mode, flags & _WEIGHT_BITS, tile_len).astype(np.int64)
    ind = [np.repeat(np.asarray(x), w) for x in t.ind]
    vals = np.repeat(np.asarray(t.vals, dtype=np.float64), w)
    if flags & QUIRK_TRUNC_VALUES:
        vals = np.trunc(vals)
    return Coo(ind, vals, t.dims)


def ccd_jacobi_mode(t: Coo, mode: int, F, lam, flags=0, tile_len: int = 32):
    """In-place FP64 update of F[mode] by the reference's CCD form (quirk Q4,
    orc_ccd_jacobi_mode), its sums weighted by quirk_weights(flags) and the
    values truncated under QUIRK_TRUNC_VALUES."""
    assert F[0].dtype == np.float64
    perm, rowptr = mode_order(t, mode)
    w = quirk_weights(t, mode, flags & _WEIGHT_BITS, tile_len)
    st = lib().orc_ccd_jacobi_mode(t.c(), C.c_uint32(mode), C.c_uint32(F[0].shape[1]), _ptrs(F, "f64"),
                                   C.c_double(lam), _u32p(w), C.c_int(1 if flags & QUIRK_TRUNC_VALUES else 0),
                                   _u32p(perm), _u64p(rowptr))
    assert st == 0, st


def ccd_jacobi_epoch(t: Coo, F, lam, flags=0, start=0):
    """One epoch of the reference's CCD form: modes start, start + 1, ... in turn."""
    for k in range(t.nmodes):
        ccd_jacobi_mode(t, (start + k) % t.nmodes, F, lam, flags)


def sort_by_mode(t: Coo, mode: int) -> Coo:
    perm, _ = mode_order(t, mode)
    return t.permuted(perm)


def ciss_reference_layout(t_sorted: Coo, mode: int, tile_len: int = 32, truncate_vals=False):
    """Reference TB-COO arrays (SG/ciss.c:68-195).  Fields the reference leaves
    unwritten are NaN."""
    n = t_sorted.nnz
    directory = np.zeros(n + 1, np.uint64)
    dcounter = np.zeros(n + 2, np.uint64)
    entries = np.full(3 * (n // tile_len + 1 + n + 1), np.nan)
    dl = C.c_uint64()
    sz = C.c_uint64()
    st = lib().orc_ciss_reference_layout(t_sorted.c(), C.c_uint32(mode), C.c_uint32(tile_len),
                                         C.c_int(1 if truncate_vals else 0), _u64p(directory),
                                         _u64p(dcounter), _dp(entries), C.byref(dl), C.byref(sz))
    if st != 0:
        raise ValueError(f"orc_ciss_reference_layout status {st}")
    d = dl.value
    return directory[:d].copy(), dcounter[:d + 1].copy(), entries[: 3 * sz.value].reshape(-1, 3).copy()


def _fn(name, F):
    return getattr(lib(), f"{name}_{_prec(F)}")


def loss_sq(t: Coo, F, nthreads=1):
    f = _fn("orc_loss_sq", F)
    f.restype = C.c_double
    return f(t.c(), C.c_uint32(F[0].shape[1]), _ptrs(F, _prec(F)), C.c_int(nthreads))


def rmse(t: Coo, F, nthreads=1):
    return float(np.sqrt(loss_sq(t, F, nthreads) / t.nnz))


def frob_sq(dims, F, lam):
    f = _fn("orc_frob_sq", F)
    f.restype = C.c_double
    d = (C.c_uint64 * MAX_MODES)(*dims)
    return f(C.c_uint32(len(F)), d, C.c_uint32(F[0].shape[1]), _ptrs(F, _prec(F)), C.c_double(lam))


def mttkrp(t: Coo, mode, F, order=None, nthreads=1):
    perm, rowptr = order if order is not None else mode_order(t, mode)
    R = F[0].shape[1]
    out = np.zeros((t.dims[mode], R), F[0].dtype)
    p = _fp(out) if out.dtype == np.float32 else _dp(out)
    st = _fn("orc_mttkrp", F)(t.c(), C.c_uint32(mode), C.c_uint32(R), _ptrs(F, _prec(F)),
                              _u32p(perm), _u64p(rowptr), p, C.c_int(nthreads))
    assert st == 0
    return out


def gram_rows(t: Coo, mode, F, row0=0, nrows=None, order=None):
    perm, rowptr = order if order is not None else mode_order(t, mode)
    R = F[0].shape[1]
    if nrows is None:
        nrows = t.dims[mode] - row0
    G = np.zeros((nrows, R, R), np.float64)
    st = _fn("orc_gram_rows", F)(t.c(), C.c_uint32(mode), C.c_uint32(R), _ptrs(F, _prec(F)),
                                 _u32p(perm), _u64p(rowptr), C.c_uint64(row0), C.c_uint64(nrows),
                                 _dp(G))
    assert st == 0
    return G


def als_mode(t: Coo, mode, F, lam, order=None, nthreads=1):
    """In-place ALS update of F[mode].  Returns the status (0 ok, 7 not SPD)."""
    perm, rowptr = order if order is not None else mode_order(t, mode)
    R = F[0].shape[1]
    return _fn("orc_als_mode", F)(t.c(), C.c_uint32(mode), C.c_uint32(R), _ptrs(F, _prec(F)),
                                  C.c_double(lam), _u32p(perm), _u64p(rowptr), C.c_int(nthreads))


def bf16_round(x):
    """float32 -> bfloat16 -> float32, round to nearest even (scalar)."""
    f = lib().orc_bf16_round
    f.restype = C.c_float
    f.argtypes = [C.c_float]
    return float(f(float(x)))


def als_mode_bf16(t: Coo, mode, F, lam, order=None, nthreads=1):
    """In-place ALS update of F[mode] (float32 factors) in the arithmetic of the
    BF16-operand path: BF16-rounded factor rows, BF16-rounded h in the Gram and the right-hand side,
    FP64 accumulation and solve.  Returns the status (0 ok, 7 not SPD)."""
    perm, rowptr = order if order is not None else mode_order(t, mode)
    R = F[0].shape[1]
    assert all(f.dtype == np.float32 for f in F)
    fn = lib().orc_als_mode_bf16
    fn.restype = C.c_int
    return fn(t.c(), C.c_uint32(mode), C.c_uint32(R), _ptrs(F, "f32"), C.c_double(lam), _u32p(perm),
              _u64p(rowptr), C.c_int(nthreads))


def als_rows(t: Coo, mode, F, lam, rows, order=None, nthreads=1):
    """FP64 ALS update of the given rows of F[mode] from float32 factors (F is
    not modified) -> (len(rows), R) float64 array."""
    perm, rowptr = order if order is not None else mode_order(t, mode)
    R = F[0].shape[1]
    assert all(f.dtype == np.float32 for f in F)
    rows = np.ascontiguousarray(rows, np.uint32)
    out = np.zeros((len(rows), R), np.float64)
    fn = lib().orc_als_rows_f64
    fn.restype = C.c_int
    st = fn(t.c(), C.c_uint32(mode), C.c_uint32(R), _ptrs(F, "f32"), C.c_double(lam), _u32p(perm),
            _u64p(rowptr), _u32p(rows), C.c_uint64(len(rows)), _dp(out), C.c_int(nthreads))
    assert st == 0, st
    return out


def grad_rows(t: Coo, mode, F, lam, rows, reg_per_occurrence=False, order=None, nthreads=1):
    """Rows `rows` of mode `mode`'s loss gradient in FP64 from float32 factors
    (orc_grad_rows_f64) -> (len(rows), R) float64 array."""
    perm, rowptr = order if order is not None else mode_order(t, mode)
    R = F[0].shape[1]
    assert all(f.dtype == np.float32 for f in F)
    rows = np.ascontiguousarray(rows, np.uint32)
    out = np.zeros((len(rows), R), np.float64)
    fn = lib().orc_grad_rows_f64
    fn.restype = C.c_int
    st = fn(t.c(), C.c_uint32(mode), C.c_uint32(R), _ptrs(F, "f32"), C.c_double(lam),
            C.c_int(1 if reg_per_occurrence else 0), _u32p(perm), _u64p(rowptr), _u32p(rows),
            C.c_uint64(len(rows)), _dp(out), C.c_int(nthreads))
    assert st == 0, st
    return out


def sgd_epoch(t: Coo, F, lr, lam, scale=1.0, order=None):
    R = F[0].shape[1]
    op = _u32p(order) if order is not None else None
    _fn("orc_sgd_epoch", F)(t.c(), op, C.c_uint32(R), _ptrs(F, _prec(F)), C.c_double(lr),
                            C.c_double(lam), C.c_double(scale))


def sgd_sync_epoch(t: Coo, F, lr, lam, span, sub, hot_rows, scale=1.0, order=None, nthreads=1, mixed=False):
    """One epoch of the bulk-synchronous walker schedule (libctc's default SGD):
    walkers of `span` positions of `order`, sub-steps of `sub` nonzeros from
    the sub-step's factors, per-run deltas summed on cold rows and averaged on
    hot rows (more than hot_rows and 4x the mode's mean nonzeros).  Restates
    SG/sgd.cu:21-143 (walker per tile, atomicAdd merge) with the replica
    averaging of NC/sgd.cu:229-262,523-525 for hot rows.  mixed=True: the
    mixed-precision variant, whose run deltas of modes 1.. are rounded
    (through FP32) to BF16 before they are summed."""
    R = F[0].shape[1]
    op = _u32p(order) if order is not None else None
    _fn("orc_sgd_sync_epoch", F)(t.c(), op, C.c_uint32(R), _ptrs(F, _prec(F)), C.c_double(lr),
                                 C.c_double(lam), C.c_double(scale), C.c_uint64(span), C.c_uint64(sub),
                                 C.c_uint64(hot_rows), C.c_int(1 if mixed else 0), C.c_int(nthreads))


class Csf:
    """CSF of one mode of a 3-way tensor (SG/csf.c:11-331), built by the
    oracle; numpy views of its arrays."""

    def __init__(self, t: Coo, mode: int):
        self.s = OrcCsf()
        self._t = t
        st = lib().orc_csf_build(t.c(), C.c_uint32(mode), C.byref(self.s))
        assert st == 0, st

    def __del__(self):
        if getattr(self, "s", None) is not None and _LIB is not None:
            _LIB.orc_csf_free(C.byref(self.s))
            self.s = None

    def _arr(self, p, n):
        return np.ctypeslib.as_array(p, shape=(int(n),)).copy() if n else np.zeros(0)

    @property
    def level_mode(self):
        return tuple(int(x) for x in self.s.level_mode)

    @property
    def nfibs(self):
        return tuple(int(x) for x in self.s.nfibs)

    def fptr(self, level):
        return self._arr(self.s.fptr[level], self.s.nfibs[level] + 1)

    def fids(self, level):
        return self._arr(self.s.fids[level], self.s.nfibs[level])

    @property
    def vals(self):
        return self._arr(self.s.vals, self.s.nnz)

    @property
    def blockref(self):
        return np.array(self.s.blockref[:], dtype=np.int64)


def mttkrp_csf(csf: Csf, F, nthreads=1):
    """MTTKRP of the CSF's mode (SG/mttkrp.c:16-93) -> (dims[mode], R) array
    in F's precision."""
    R = F[0].shape[1]
    out = np.empty((int(csf.s.dims[csf.s.mode]), R), F[0].dtype)
    p = _fp(out) if out.dtype == np.float32 else _dp(out)
    st = _fn("orc_mttkrp_csf", F)(C.byref(csf.s), C.c_uint32(R), _ptrs(F, _prec(F)), p, C.c_int(nthreads))
    assert st == 0, st
    return out


def residual(t: Coo, F):
    res = np.empty(t.nnz, F[0].dtype)
    p = _fp(res) if res.dtype == np.float32 else _dp(res)
    _fn("orc_residual", F)(t.c(), C.c_uint32(F[0].shape[1]), _ptrs(F, _prec(F)), p)
    return res


def ccd_epoch(t: Coo, F, lam, res, inner=1, orders=None, nthreads=1):
    R = F[0].shape[1]
    if orders is None:
        orders = [mode_order(t, m) for m in range(t.nmodes)]
    P = (C.POINTER(C.c_uint32) * MAX_MODES)()
    RP = (C.POINTER(C.c_uint64) * MAX_MODES)()
    for m, (perm, rowptr) in enumerate(orders):
        P[m] = _u32p(perm)
        RP[m] = _u64p(rowptr)
    p = _fp(res) if res.dtype == np.float32 else _dp(res)
    st = _fn("orc_ccd_epoch", F)(t.c(), C.c_uint32(R), _ptrs(F, _prec(F)), C.c_double(lam),
                                 C.c_uint32(inner), p, P, RP, C.c_int(nthreads))
    assert st == 0


def grad(t: Coo, F, lam, reg_per_occurrence=False):
    G = [np.zeros_like(f) for f in F]
    f = _fn("orc_grad", F)
    f.restype = C.c_double
    loss = f(t.c(), C.c_uint32(F[0].shape[1]), _ptrs(F, _prec(F)), C.c_double(lam),
             C.c_int(1 if reg_per_occurrence else 0), _ptrs(G, _prec(F)))
    return loss, G


def cholesky(A):
    A = np.array(A, np.float64, order="C", copy=True)
    st = lib().orc_cholesky(_dp(A), C.c_uint32(A.shape[0]))
    if st != 0:
        raise np.linalg.LinAlgError("not SPD")
    return A


def cholesky_solve(L, b):
    x = np.array(b, np.float64, copy=True)
    lib().orc_cholesky_solve(_dp(np.ascontiguousarray(L)), C.c_uint32(L.shape[0]), _dp(x))
    return x


class Converge:
    """SG/completion.c:296-362 bookkeeping."""

    def __init__(self, tolerance=1e-4, max_badepochs=3, best_rmse=1e7):
        self.s = OrcConv(best_rmse, tolerance, 0, 0, max_badepochs)

    def step(self, rmse_value, epoch):
        imp = C.c_int()
        conv = lib().orc_converge(C.byref(self.s), C.c_double(rmse_value), C.c_uint64(epoch),
                                  C.byref(imp))
        return bool(conv), bool(imp.value)


# ---------------------------------------------------------------------------
# whole-solver drivers (the reference's epoch loops, SG/als.cu:676-810,
# SG/sgd.cu:276-324, SG/ccd.cu:524-626), fixed mode order 0..N-1.
# ---------------------------------------------------------------------------

def als(t: Coo, F, lam, iters, test: Coo | None = None, nthreads=1):
    """Runs `iters` ALS sweeps in place.  Returns per-iteration factor copies and
    test RMSE."""
    orders = [mode_order(t, m) for m in range(t.nmodes)]
    hist, rm = [], []
    for _ in range(iters):
        for m in range(t.nmodes):
            st = als_mode(t, m, F, lam, order=orders[m], nthreads=nthreads)
            if st != 0:
                raise np.linalg.LinAlgError(f"ALS system not SPD in mode {m}")
        hist.append([f.copy() for f in F])
        if test is not None:
            rm.append(rmse(test, F))
    return hist, rm


def sgd(t: Coo, F, lr, lam, epochs, test: Coo | None = None, scale=1.0, decay=None,
        bold_driver=False, order=None):
    """Sequential SGD in mode-0 order.  decay: multiply lr by it after each
    epoch (exponential schedule).  bold_driver: the reference schedule
    (SG/sgd.cu:314-323)."""
    if order is None:
        order, _ = mode_order(t, 0)
    rm, lrs = [], []
    prev_obj = None
    for e in range(1, epochs + 1):
        lrs.append(lr)
        sgd_epoch(t, F, lr, lam, scale, order=order)
        if test is not None:
            rm.append(rmse(test, F))
        if bold_driver:
            obj = loss_sq(t, F) + frob_sq(t.dims, F, lam)
            if e > 1:
                lr = lr * 1.05 if obj < prev_obj else lr * 0.5
            prev_obj = obj
        elif decay is not None:
            lr *= decay
    return rm, lrs


def sgd_sync(t: Coo, F, lr, lam, epochs, span, sub, hot_rows, test: Coo | None = None, scale=1.0, decay=None,
             order=None, nthreads=1, mixed=False):
    """Epochs of sgd_sync_epoch in mode-0 order with an exponential lr decay."""
    if order is None:
        order, _ = mode_order(t, 0)
    rm, lrs = [], []
    for _ in range(epochs):
        lrs.append(lr)
        sgd_sync_epoch(t, F, lr, lam, span, sub, hot_rows, scale, order=order, nthreads=nthreads, mixed=mixed)
        if test is not None:
            rm.append(rmse(test, F))
        if decay is not None:
            lr *= decay
    return rm, lrs


def ccd(t: Coo, F, lam, epochs, test: Coo | None = None, inner=1, nthreads=1):
    orders = [mode_order(t, m) for m in range(t.nmodes)]
    res = residual(t, F)
    hist, rm = [], []
    for _ in range(epochs):
        ccd_epoch(t, F, lam, res, inner=inner, orders=orders, nthreads=nthreads)
        hist.append([f.copy() for f in F])
        if test is not None:
            rm.append(rmse(test, F))
    return hist, rm
