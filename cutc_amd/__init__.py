"""cutc_amd — MI355X-native sparse tensor completion (ALS, CCD++, SGD over CISS).

Python mirror of the reference's operator surface (abovekumo/cuTC), calling
libctc.so (HIP kernels for gfx950) through its C ABI:

  reference (SG = software/single GPU/sgpu)          here
  tt_read            SG/sptensor.c:64-70             tt_read / Coo.read
  ciss_alloc         SG/ciss.c:68-195                ciss_alloc / Ciss
  tc_main_ciss       SG/completion.c:422-497         tc_main_ciss
  tc_als/sgd/ccd     SG/als.cuh:20-33 etc.           tc_als / tc_sgd / tc_ccd
  tc_rmse/loss_sq    SG/completion.c:123-214         Model.evaluate
  GA autotuner       GA/fitness.c:31-100             Model.autotune

There is no CPU fallback: every compute call runs the HIP kernels and fails
if libctc.so or the GPU is missing.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, fields

import numpy as np

from . import _lib
from ._lib import (ALS, ALS_AUTO, ALS_ROWS, ALS_SPANS, CCD, LR_BOLD_DRIVER, LR_CONSTANT, LR_EXP_DECAY, ORDER_FIXED,  # noqa: F401
                   ORDER_ROTATE, PREC_FP32, PREC_MIXED, QUIRK_CCD_GRID, QUIRK_LAST_LANE, QUIRK_NO_ADDBACK,
                   QUIRK_TILE_HEAD, QUIRK_TRUNC_VALUES, SGD,
                   SGD_ATOMIC, SGD_AUTO, SGD_HOT_ATOMIC, SGD_NEXT_RACY, SGD_RACY, SGD_SYNC,
                   VALUES_AFFINE_1_5, VALUES_INT_1_5, VALUES_RAW, CtcError, check, lib)

__all__ = ["Coo", "Ciss", "Context", "Model", "Batch", "Options", "synth", "tt_read", "ciss_alloc",
           "tc_main_ciss", "tc_als", "tc_sgd", "tc_ccd", "CtcError", "ALS", "SGD", "CCD"]


# ---------------------------------------------------------------------------
# COO tensors
# ---------------------------------------------------------------------------
class Coo:
    """A COO tensor owned by libctc (0-based uint32 indices, float64 values)."""

    def __init__(self, ptr):
        self._p = ptr  # POINTER(_lib.Coo)

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None and _lib.alive():
            lib().ctc_coo_free(self._p)
            self._p = None

    @property
    def c(self):
        return self._p

    @property
    def nmodes(self):
        return int(self._p.contents.nmodes)

    @property
    def nnz(self):
        return int(self._p.contents.nnz)

    @property
    def dims(self):
        s = self._p.contents
        return tuple(int(s.dims[m]) for m in range(s.nmodes))

    @property
    def base(self):
        s = self._p.contents
        return tuple(int(s.base[m]) for m in range(s.nmodes))

    def ind(self, m):
        return np.ctypeslib.as_array(self._p.contents.ind[m], shape=(self.nnz,))

    @property
    def vals(self):
        return np.ctypeslib.as_array(self._p.contents.vals, shape=(self.nnz,))

    def ref_quirk_weights(self, mode, flags):
        """How many times each nonzero's term reaches its row's sums in the
        reference's ALS and CCD reductions of `mode` under QUIRK_* `flags`
        (ctc_ref_quirk_weights)."""
        w = np.empty(self.nnz, np.uint32)
        check(lib().ctc_ref_quirk_weights(self._p, mode, flags, w.ctypes.data_as(C.POINTER(C.c_uint32))))
        return w

    @classmethod
    def read(cls, path):
        p = C.POINTER(_lib.Coo)()
        check(lib().ctc_coo_read(str(path).encode(), C.byref(p)))
        return cls(p)

    @classmethod
    def from_arrays(cls, ind, vals, dims=None):
        ind = [np.asarray(a) for a in ind]
        vals = np.asarray(vals, dtype=np.float64)
        n = len(ind)
        if dims is None:
            dims = [int(a.max()) + 1 if a.size else 0 for a in ind]
        d = (C.c_uint64 * _lib.MAX_MODES)(*dims)
        p = C.POINTER(_lib.Coo)()
        check(lib().ctc_coo_alloc(n, vals.shape[0], d, C.byref(p)))
        t = cls(p)
        for m in range(n):
            t.ind(m)[:] = ind[m].astype(np.uint32)
        t.vals[:] = vals
        return t

    def sorted(self, keys):
        """Stable lexicographic sort by the given modes (stensor.py's orders)."""
        k = (C.c_uint32 * len(keys))(*keys)
        out = C.POINTER(_lib.Coo)()
        check(lib().ctc_coo_sort(self.c, k, len(keys), C.byref(out)))
        return Coo(out)

    def split(self, seed=0):
        """Seeded 80/10/10 split -> (train, validation, test) (tvtensor.py)."""
        tr, va, te = C.POINTER(_lib.Coo)(), C.POINTER(_lib.Coo)(), C.POINTER(_lib.Coo)()
        check(lib().ctc_coo_split(self.c, int(seed), C.byref(tr), C.byref(va), C.byref(te)))
        return Coo(tr), Coo(va), Coo(te)

    def write_binary(self, path, base=1):
        """SPLATT binary coordinate file (read back by Coo.read for *.bin)."""
        check(lib().ctc_coo_write_binary(self.c, path.encode(), int(base)))

    def write(self, path, base=1, decimals=-1):
        check(lib().ctc_coo_write(self._p, str(path).encode(), base, decimals))

    def arrays(self):
        return [self.ind(m).copy() for m in range(self.nmodes)], self.vals.copy()


def tt_read(path):
    """Reads a .tns file (the reference's tt_read, SG/sptensor.c:64-70)."""
    return Coo.read(path)


def synth(dims, nnz, zipf=None, true_rank=8, value_map=VALUES_AFFINE_1_5, noise_std=0.1,
          test_fraction=0.1, seed=0):
    """Seeded synthetic tensor -> (train sorted by mode 0, test)."""
    s = _lib.SynthSpec()
    s.nmodes = len(dims)
    for m, d in enumerate(dims):
        s.dims[m] = int(d)
        s.zipf[m] = float(zipf[m]) if zipf else 0.0
    s.nnz = int(nnz)
    s.true_rank = int(true_rank)
    s.value_map = int(value_map)
    s.noise_std = float(noise_std)
    s.test_fraction = float(test_fraction)
    s.seed = int(seed)
    tr = C.POINTER(_lib.Coo)()
    te = C.POINTER(_lib.Coo)()
    check(lib().ctc_synth_generate(C.byref(s), C.byref(tr), C.byref(te)))
    return Coo(tr), (Coo(te) if te else None)


# ---------------------------------------------------------------------------
# CISS
# ---------------------------------------------------------------------------
class Ciss:
    """Host CISS of one mode (ciss_alloc, SG/ciss.c:68-195)."""

    def __init__(self, coo: Coo, mode: int):
        self._coo = coo
        self._h = C.c_void_p()
        check(lib().ctc_ciss_build(coo.c, mode, C.byref(self._h)))
        v = _lib.CissView()
        check(lib().ctc_ciss_get_view(self._h, C.byref(v)))
        self._v = v

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None and _lib.alive():
            lib().ctc_ciss_free(self._h)
            self._h = None

    def _arr(self, ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,)) if n else np.zeros(0)

    @property
    def nnz(self):
        return int(self._v.nnz)

    @property
    def dlength(self):
        return int(self._v.dlength)

    @property
    def ntiles(self):
        return int(self._v.ntiles)

    @property
    def directory(self):
        return self._arr(self._v.directory, self.dlength)

    @property
    def rowptr(self):
        return self._arr(self._v.rowptr, self.dlength + 1)

    def oidx(self, s):
        return self._arr(self._v.oidx[s], self.nnz)

    @property
    def vals(self):
        return self._arr(self._v.vals, self.nnz)

    @property
    def tile_slot(self):
        return self._arr(self._v.tile_slot, self.ntiles)

    @property
    def tile_mask(self):
        return self._arr(self._v.tile_mask, self.ntiles)

    @property
    def perm(self):
        return self._arr(self._v.perm, self.nnz)

    def reference_layout(self, tile_len=32, truncate_vals=False):
        """The reference's (directory, dcounter, entries) arrays; fields the
        reference leaves unwritten are NaN."""
        n = self.nnz
        d = np.zeros(max(self.dlength, 1), np.uint64)
        dc = np.zeros(self.dlength + 1, np.uint64)
        ent = np.full(3 * (n // tile_len + 1 + n + 1), np.nan)
        dl = C.c_uint64()
        sz = C.c_uint64()
        check(lib().ctc_ciss_reference_layout(
            self._h, self._coo.c, tile_len, 1 if truncate_vals else 0,
            d.ctypes.data_as(C.POINTER(C.c_uint64)), dc.ctypes.data_as(C.POINTER(C.c_uint64)),
            ent.ctypes.data_as(C.POINTER(C.c_double)), C.byref(dl), C.byref(sz)))
        return d[: dl.value], dc[: dl.value + 1], ent[: 3 * sz.value].reshape(-1, 3)


def ciss_alloc(coo: Coo, mode: int) -> Ciss:
    """Reference name (SG/ciss.c:68).  `mode` is 0-based here (the reference's
    is 1-based)."""
    return Ciss(coo, mode)


# ---------------------------------------------------------------------------
# options
# ---------------------------------------------------------------------------
@dataclass
class Options:
    """Run-time form of the reference's compile-time constants (SG/base.h)."""

    algorithm: int = ALS
    rank: int = 16
    lambda_: float = None
    learning_rate: float = 1e-3
    lr_schedule: int = LR_BOLD_DRIVER
    lr_decay: float = 0.95
    sgd_scale: float = 1e-5
    max_epochs: int = None
    tolerance: float = 1e-4
    max_bad_epochs: int = 3
    tile_len: int = None
    block: int = None
    seed: int = 0
    mode_order: int = ORDER_FIXED
    precision: int = PREC_FP32
    inner_iters: int = 1
    refine: int = 0
    eval_every_epoch: int = 1
    verbose: int = 0
    sgd_update: int = None  # None: the library default (SGD_SYNC)
    spread: int = 0
    sgd_hot_rows: int = None  # None: the library default (65536)
    als_kernel: int = 0
    sgd_substep: int = None
    ref_quirks: int = None  # QUIRK_* bits: the reference's ALS / CCD kernel defects (diagnostic)

    def to_c(self):
        o = _lib.Options()
        check(lib().ctc_options_default(self.algorithm, C.byref(o)))
        for f in fields(self):
            v = getattr(self, f.name)
            if v is not None:
                setattr(o, f.name, v)
        return o


# ---------------------------------------------------------------------------
# device context and model
# ---------------------------------------------------------------------------
class Context:
    def __init__(self, device=0):
        self._h = C.c_void_p()
        check(lib().ctc_context_create(device, C.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None and _lib.alive():
            lib().ctc_context_destroy(self._h)
            self._h = None

    @property
    def stream(self):
        s = C.c_void_p()
        check(lib().ctc_context_stream(self._h, C.byref(s)))
        return s.value

    def synchronize(self):
        check(lib().ctc_context_synchronize(self._h))

    def set_profiling(self, on=True, every=1):
        """Time kernels with HIP events: every launch, or every `every`-th
        launch of each kernel name (a sample that barely perturbs the run)."""
        check(lib().ctc_context_set_profiling(self._h, int(every) if on else 0))

    def kernel_stats(self, name):
        n = C.c_uint64()
        ms = C.c_double()
        check(lib().ctc_context_kernel_stats(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)

    def reset_stats(self):
        check(lib().ctc_context_reset_stats(self._h))

    def copy_bandwidth(self, nbytes=1 << 30):
        g = C.c_double()
        check(lib().ctc_context_copy_bandwidth(self._h, int(nbytes), C.byref(g)))
        return float(g.value)


@dataclass
class SolveResult:
    epochs: int
    best_epoch: int
    best_rmse: float
    ms_per_epoch: float
    nnz_per_s: float
    train_nnz: int
    converged: bool
    history: list = field(default_factory=list)


class Model:
    """Device-resident completion problem: training CISS per mode,
    validation/test tensors and the factor matrices."""

    def __init__(self, ctx: Context, train: Coo, validation: Coo | None = None,
                 test: Coo | None = None, rank=16):
        self.ctx = ctx
        self.rank = int(rank)
        self.dims = train.dims
        self.nmodes = train.nmodes
        self._h = C.c_void_p()
        check(lib().ctc_model_create(ctx._h, train.c, validation.c if validation else None,
                                     test.c if test else None, self.rank, C.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None and _lib.alive():
            lib().ctc_model_destroy(self._h)
            self._h = None

    def init_random(self, seed=0):
        check(lib().ctc_model_init_random(self._h, int(seed)))

    def set_factor(self, mode, F):
        F = np.ascontiguousarray(F, dtype=np.float32)
        assert F.shape == (self.dims[mode], self.rank), F.shape
        check(lib().ctc_model_set_factor(self._h, mode, F.ctypes.data_as(C.POINTER(C.c_float))))

    def set_factors(self, Fs):
        for m, F in enumerate(Fs):
            self.set_factor(m, F)

    def factor(self, mode, best=False):
        out = np.empty((self.dims[mode], self.rank), np.float32)
        check(lib().ctc_model_get_factor(self._h, mode, 1 if best else 0,
                                         out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def factors(self, best=False):
        return [self.factor(m, best) for m in range(self.nmodes)]

    def factor_device_ptr(self, mode):
        p = C.c_void_p()
        check(lib().ctc_model_factor_device(self._h, mode, C.byref(p)))
        return p.value

    def prepare(self, opt: Options):
        o = opt.to_c()
        check(lib().ctc_model_prepare(self._h, C.byref(o)))

    def solve(self, opt: Options) -> SolveResult:
        o = opt.to_c()
        cap = int(o.max_epochs) + 1
        hist = (_lib.EpochStats * cap)()
        r = _lib.Result()
        r.history = C.cast(hist, C.POINTER(_lib.EpochStats))
        r.history_cap = cap
        check(lib().ctc_solve(self._h, C.byref(o), C.byref(r)))
        h = [{k: getattr(hist[e], k) for k, _ in _lib.EpochStats._fields_} for e in range(r.epochs + 1)]
        return SolveResult(r.epochs, r.best_epoch, r.best_rmse, r.ms_per_epoch, r.nnz_per_s,
                           r.train_nnz, bool(r.converged), h)

    def als_mode_update(self, mode, opt: Options):
        o = opt.to_c()
        check(lib().ctc_als_mode_update(self._h, mode, C.byref(o)))

    def sgd_epoch(self, opt: Options, lr=None):
        o = opt.to_c()
        check(lib().ctc_sgd_epoch(self._h, C.byref(o), o.learning_rate if lr is None else lr))

    def ccd_epoch(self, opt: Options):
        o = opt.to_c()
        check(lib().ctc_ccd_epoch(self._h, C.byref(o)))

    def mttkrp(self, mode):
        out = np.empty((self.dims[mode], self.rank), np.float32)
        check(lib().ctc_mttkrp(self._h, mode, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def evaluate(self, lam=0.0):
        v = [C.c_double() for _ in range(5)]
        check(lib().ctc_evaluate(self._h, lam, *[C.byref(x) for x in v]))
        return dict(zip(["loss", "frob", "rmse_tr", "rmse_vl", "rmse_te"], [x.value for x in v]))

    def grad_device(self, factor_ptrs, grad_ptrs, lam, reg_per_occurrence=False, stream=None):
        """Gradient on device buffers (addresses as ints); returns the loss."""
        n = self.nmodes
        fp = (C.c_void_p * n)(*factor_ptrs)
        gp = (C.c_void_p * n)(*grad_ptrs)
        loss = C.c_double()
        check(lib().ctc_grad(self._h, fp, gp, lam, 1 if reg_per_occurrence else 0, C.byref(loss),
                             C.c_void_p(stream) if stream else None))
        return loss.value

    def grad_device_noalloc(self, factor_ptrs, grad_ptrs, lam, loss_ptr=None, reg_per_occurrence=False,
                            stream=None):
        """ctc_grad_noalloc: the same gradient on device buffers without any
        device allocation or synchronisation; the loss goes to the device
        double at loss_ptr (None: not written).  Enqueued on `stream`."""
        n = self.nmodes
        fp = (C.c_void_p * n)(*factor_ptrs)
        gp = (C.c_void_p * n)(*grad_ptrs)
        check(lib().ctc_grad_noalloc(self._h, fp, gp, lam, 1 if reg_per_occurrence else 0,
                                     C.c_void_p(loss_ptr) if loss_ptr else None,
                                     C.c_void_p(stream) if stream else None))

    def autotune(self, opt: Options):
        o = opt.to_c()
        r = _lib.TuneResult()
        check(lib().ctc_autotune(self._h, C.byref(o), C.byref(r)))
        return {"tile_len": r.tile_len, "block": r.block, "ms": r.ms, "trials": r.trials,
                "sgd_substep": r.sgd_substep}


# ---------------------------------------------------------------------------
# batched ALS over many independent problems
# ---------------------------------------------------------------------------
class Batch:
    """Independent completion problems solved together with ALS.

    Problems are grouped by rank (8/16/32/64); each rank class is one
    block-diagonal tensor, so every ALS phase is one kernel launch over all
    problems of the class, and `als(iters, graph=True)` replays a hipGraph of
    the whole run with one branch per class.  A problem's factor of mode m has
    max(train.dims[m], test.dims[m]) rows."""

    def __init__(self, ctx: Context, trains, tests, ranks):
        n = len(trains)
        assert len(ranks) == n and (tests is None or len(tests) == n)
        self.ctx = ctx
        self._keep = list(trains) + list(tests or [])
        tr = (C.POINTER(_lib.Coo) * n)(*[t.c for t in trains])
        te = (C.POINTER(_lib.Coo) * n)(*[(t.c if t is not None else None) for t in tests]) if tests else None
        rk = (C.c_uint32 * n)(*[int(r) for r in ranks])
        self.ranks = [int(r) for r in ranks]
        self.dims = [tuple(max(a, b) for a, b in zip(trains[p].dims, tests[p].dims)) if tests and tests[p] is not None
                     else tuple(trains[p].dims) for p in range(n)]
        self.nproblems = n
        self._h = C.c_void_p()
        check(lib().ctc_batch_create(ctx._h, n, tr, te, rk, C.byref(self._h)))
        np_, nc, nnz = C.c_uint32(), C.c_uint32(), C.c_uint64()
        check(lib().ctc_batch_info(self._h, C.byref(np_), C.byref(nc), C.byref(nnz)))
        self.nclasses, self.train_nnz = nc.value, nnz.value
        _lib.release_at_exit(self)

    def _release(self):
        if getattr(self, "_h", None):
            lib().ctc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None and _lib.alive():
            self._release()

    def set_factor(self, p, mode, F):
        F = np.ascontiguousarray(F, dtype=np.float32)
        assert F.shape == (self.dims[p][mode], self.ranks[p]), F.shape
        check(lib().ctc_batch_set_factor(self._h, p, mode, F.ctypes.data_as(C.POINTER(C.c_float))))

    def factor(self, p, mode):
        out = np.empty((self.dims[p][mode], self.ranks[p]), np.float32)
        check(lib().ctc_batch_get_factor(self._h, p, mode, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def classes(self):
        out = []
        for k in range(self.nclasses):
            r, n = C.c_uint32(), C.c_uint32()
            check(lib().ctc_batch_class_rank(self._h, k, C.byref(r), C.byref(n)))
            out.append((r.value, n.value))
        return out

    def autotune(self, opt: Options):
        o = opt.to_c()
        res = (_lib.TuneResult * 4)()
        n = C.c_uint32()
        check(lib().ctc_batch_autotune(self._h, C.byref(o), res, C.byref(n)))
        return [{"tile_len": r.tile_len, "block": r.block, "ms": r.ms, "trials": r.trials} for r in res[:n.value]]

    def als(self, opt: Options, iters=10, graph=True):
        """Runs `iters` ALS iterations on every problem; returns device ms."""
        o = opt.to_c()
        ms = C.c_double()
        check(lib().ctc_batch_als(self._h, C.byref(o), int(iters), 1 if graph else 0, C.byref(ms)))
        return ms.value

    def rmse(self):
        tr = np.zeros(self.nproblems)
        te = np.zeros(self.nproblems)
        check(lib().ctc_batch_rmse(self._h, tr.ctypes.data_as(C.POINTER(C.c_double)),
                                   te.ctypes.data_as(C.POINTER(C.c_double))))
        return tr, te


# ---------------------------------------------------------------------------
# reference-shaped entry points
# ---------------------------------------------------------------------------
def _mats(arrs):
    ms = []
    for a in arrs:
        assert a.dtype == np.float64 and a.flags.c_contiguous
        m = _lib.Matrix()
        m.I, m.J = a.shape
        m.values = a.ctypes.data_as(C.POINTER(C.c_double))
        m.row_col_flag = 1
        ms.append(m)
    arr = (C.POINTER(_lib.Matrix) * len(ms))(*[C.pointer(m) for m in ms])
    return arr, ms


def tc_main_ciss(traina: Coo, trainb, trainc, validation, test, algorithm_index, opt: Options = None):
    """SG/completion.c:422-497: random init, solve, print progress."""
    o = opt.to_c() if opt else None
    check(lib().ctc_tc_main_ciss(traina.c, trainb.c if trainb else None, trainc.c if trainc else None,
                                 validation.c if validation else None, test.c if test else None,
                                 algorithm_index, C.byref(o) if o else None))


def _tc(fn, traina, trainb, trainc, validation, test, mats, best_mats, algorithm_index, lam, *extra,
        tolerance=1e-4, max_badepochs=3):
    marr, keep = _mats(mats)
    barr, keep2 = _mats(best_mats) if best_mats is not None else (None, None)
    best = C.c_double(1e7)
    tol = C.c_double(tolerance)
    nbad = C.c_uint64(0)
    bep = C.c_uint64(0)
    mbad = C.c_uint64(max_badepochs)
    check(fn(traina.c, trainb.c if trainb else None, trainc.c if trainc else None,
             validation.c if validation else None, test.c if test else None, marr, barr,
             algorithm_index, lam, *extra, C.byref(best), C.byref(tol), C.byref(nbad), C.byref(bep),
             C.byref(mbad)))
    return {"best_rmse": best.value, "best_epoch": bep.value}


def tc_als(traina, trainb, trainc, validation, test, mats, best_mats, algorithm_index=ALS,
           regularization_index=10.0, **kw):
    """SG/als.cuh:20-33; mats (float64 row-major arrays) are updated in place."""
    return _tc(lib().ctc_tc_als, traina, trainb, trainc, validation, test, mats, best_mats,
               algorithm_index, regularization_index, **kw)


def tc_sgd(traina, trainb, trainc, validation, test, mats, best_mats, algorithm_index=SGD,
           regularization_index=1e-3, learning_rate=1e-3, **kw):
    """SG/sgd.cuh:18-32."""
    return _tc(lib().ctc_tc_sgd, traina, trainb, trainc, validation, test, mats, best_mats,
               algorithm_index, regularization_index, learning_rate, **kw)


def tc_ccd(traina, trainb, trainc, validation, test, mats, best_mats, algorithm_index=CCD,
           regularization_index=100.0, **kw):
    """SG/ccd.cuh:20-33."""
    return _tc(lib().ctc_tc_ccd, traina, trainb, trainc, validation, test, mats, best_mats,
               algorithm_index, regularization_index, **kw)
