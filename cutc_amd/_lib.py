"""ctypes binding of libctc.so (include/ctc.h).  Loading fails loudly when the
library has not been built: there is no fallback path."""
from __future__ import annotations

import atexit
import weakref
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libctc.so")
MAX_MODES = 4
TILE_LEN = 64

ALS, SGD, CCD = 0, 1, 2
LR_BOLD_DRIVER, LR_EXP_DECAY, LR_CONSTANT = 0, 1, 2
PREC_FP32, PREC_MIXED = 0, 1
QUIRK_TILE_HEAD, QUIRK_LAST_LANE, QUIRK_TRUNC_VALUES, QUIRK_NO_ADDBACK, QUIRK_CCD_GRID = 1, 2, 4, 8, 16  # ctc_ref_quirk
ORDER_FIXED, ORDER_ROTATE = 0, 1
SGD_AUTO, SGD_ATOMIC, SGD_RACY, SGD_NEXT_RACY, SGD_HOT_ATOMIC, SGD_SYNC = 0, 1, 2, 3, 4, 5
ALS_AUTO, ALS_SPANS, ALS_ROWS = 0, 1, 2
VALUES_RAW, VALUES_AFFINE_1_5, VALUES_INT_1_5 = 0, 1, 2

STATUS = {
    0: "CTC_OK", 1: "CTC_ERR_INVALID_ARG", 2: "CTC_ERR_IO", 3: "CTC_ERR_PARSE",
    4: "CTC_ERR_INDEX_BASE", 5: "CTC_ERR_NOMEM", 6: "CTC_ERR_HIP", 7: "CTC_ERR_NOT_SPD",
    8: "CTC_ERR_UNSUPPORTED", 9: "CTC_ERR_NO_DEVICE",
}


class CtcError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Coo(C.Structure):
    _fields_ = [
        ("nmodes", C.c_uint32),
        ("nnz", C.c_uint64),
        ("dims", C.c_uint64 * MAX_MODES),
        ("ind", C.POINTER(C.c_uint32) * MAX_MODES),
        ("vals", C.POINTER(C.c_double)),
        ("base", C.c_uint32 * MAX_MODES),
    ]


class SynthSpec(C.Structure):
    _fields_ = [
        ("nmodes", C.c_uint32),
        ("dims", C.c_uint64 * MAX_MODES),
        ("nnz", C.c_uint64),
        ("zipf", C.c_double * MAX_MODES),
        ("true_rank", C.c_uint32),
        ("value_map", C.c_int32),
        ("noise_std", C.c_double),
        ("test_fraction", C.c_double),
        ("seed", C.c_uint64),
    ]


class CissView(C.Structure):
    _fields_ = [
        ("nmodes", C.c_uint32),
        ("mode", C.c_uint32),
        ("nnz", C.c_uint64),
        ("ntiles", C.c_uint64),
        ("dlength", C.c_uint64),
        ("dims", C.c_uint64 * MAX_MODES),
        ("directory", C.POINTER(C.c_uint32)),
        ("rowptr", C.POINTER(C.c_uint32)),
        ("oidx", C.POINTER(C.c_uint32) * (MAX_MODES - 1)),
        ("vals", C.POINTER(C.c_float)),
        ("tile_slot", C.POINTER(C.c_uint32)),
        ("tile_mask", C.POINTER(C.c_uint64)),
        ("perm", C.POINTER(C.c_uint32)),
    ]


class Options(C.Structure):
    _fields_ = [
        ("algorithm", C.c_int32),
        ("rank", C.c_uint32),
        ("lambda_", C.c_double),
        ("learning_rate", C.c_double),
        ("lr_schedule", C.c_int32),
        ("lr_decay", C.c_double),
        ("sgd_scale", C.c_double),
        ("max_epochs", C.c_uint32),
        ("tolerance", C.c_double),
        ("max_bad_epochs", C.c_uint32),
        ("tile_len", C.c_uint32),
        ("block", C.c_uint32),
        ("seed", C.c_uint64),
        ("mode_order", C.c_int32),
        ("precision", C.c_int32),
        ("inner_iters", C.c_uint32),
        ("refine", C.c_int32),
        ("eval_every_epoch", C.c_int32),
        ("verbose", C.c_int32),
        ("sgd_update", C.c_int32),
        ("spread", C.c_int32),
        ("sgd_hot_rows", C.c_uint32),
        ("als_kernel", C.c_int32),
        ("sgd_substep", C.c_uint32),
        ("ref_quirks", C.c_uint32),
    ]


class EpochStats(C.Structure):
    _fields_ = [
        ("loss", C.c_double),
        ("frob", C.c_double),
        ("rmse_tr", C.c_double),
        ("rmse_vl", C.c_double),
        ("rmse_te", C.c_double),
        ("epoch_ms", C.c_double),
        ("lr", C.c_double),
    ]


class Result(C.Structure):
    _fields_ = [
        ("epochs", C.c_uint32),
        ("best_epoch", C.c_uint32),
        ("best_rmse", C.c_double),
        ("ms_per_epoch", C.c_double),
        ("nnz_per_s", C.c_double),
        ("train_nnz", C.c_uint64),
        ("converged", C.c_int32),
        ("history", C.POINTER(EpochStats)),
        ("history_cap", C.c_uint32),
    ]


class Matrix(C.Structure):
    _fields_ = [
        ("I", C.c_uint64),
        ("J", C.c_uint64),
        ("values", C.POINTER(C.c_double)),
        ("row_col_flag", C.c_int32),
    ]


class TuneResult(C.Structure):
    _fields_ = [
        ("tile_len", C.c_uint32),
        ("block", C.c_uint32),
        ("ms", C.c_double),
        ("trials", C.c_uint32),
        ("sgd_substep", C.c_uint32),
    ]


P = C.POINTER
vp = C.c_void_p

# name -> (restype, argtypes); every entry point declared in include/ctc.h
SIGNATURES = {
    "ctc_status_string": (C.c_char_p, [C.c_int]),
    "ctc_last_error": (C.c_char_p, []),
    "ctc_version": (C.c_char_p, []),
    "ctc_coo_alloc": (C.c_int, [C.c_uint32, C.c_uint64, P(C.c_uint64), P(P(Coo))]),
    "ctc_coo_read": (C.c_int, [C.c_char_p, P(P(Coo))]),
    "ctc_coo_write_binary": (C.c_int, [P(Coo), C.c_char_p, C.c_uint32]),
    "ctc_coo_sort": (C.c_int, [P(Coo), P(C.c_uint32), C.c_uint32, P(P(Coo))]),
    "ctc_coo_split": (C.c_int, [P(Coo), C.c_uint64, P(P(Coo)), P(P(Coo)), P(P(Coo))]),
    "ctc_coo_write": (C.c_int, [P(Coo), C.c_char_p, C.c_uint32, C.c_int]),
    "ctc_coo_free": (None, [P(Coo)]),
    "ctc_synth_generate": (C.c_int, [P(SynthSpec), P(P(Coo)), P(P(Coo))]),
    "ctc_ciss_build": (C.c_int, [P(Coo), C.c_uint32, P(vp)]),
    "ctc_ciss_get_view": (C.c_int, [vp, P(CissView)]),
    "ctc_ciss_reference_layout": (C.c_int, [vp, P(Coo), C.c_uint32, C.c_int, P(C.c_uint64),
                                            P(C.c_uint64), P(C.c_double), P(C.c_uint64),
                                            P(C.c_uint64)]),
    "ctc_ciss_free": (None, [vp]),
    "ctc_context_create": (C.c_int, [C.c_int, P(vp)]),
    "ctc_context_destroy": (None, [vp]),
    "ctc_context_stream": (C.c_int, [vp, P(vp)]),
    "ctc_context_synchronize": (C.c_int, [vp]),
    "ctc_context_set_profiling": (C.c_int, [vp, C.c_int]),
    "ctc_context_kernel_stats": (C.c_int, [vp, C.c_char_p, P(C.c_uint64), P(C.c_double)]),
    "ctc_context_reset_stats": (C.c_int, [vp]),
    "ctc_context_copy_bandwidth": (C.c_int, [vp, C.c_uint64, P(C.c_double)]),
    "ctc_options_default": (C.c_int, [C.c_int32, P(Options)]),
    "ctc_model_create": (C.c_int, [vp, P(Coo), P(Coo), P(Coo), C.c_uint32, P(vp)]),
    "ctc_model_destroy": (None, [vp]),
    "ctc_model_init_random": (C.c_int, [vp, C.c_uint64]),
    "ctc_model_set_factor": (C.c_int, [vp, C.c_uint32, P(C.c_float)]),
    "ctc_model_get_factor": (C.c_int, [vp, C.c_uint32, C.c_int, P(C.c_float)]),
    "ctc_model_factor_device": (C.c_int, [vp, C.c_uint32, P(vp)]),
    "ctc_model_prepare": (C.c_int, [vp, P(Options)]),
    "ctc_solve": (C.c_int, [vp, P(Options), P(Result)]),
    "ctc_als_mode_update": (C.c_int, [vp, C.c_uint32, P(Options)]),
    "ctc_sgd_epoch": (C.c_int, [vp, P(Options), C.c_double]),
    "ctc_ccd_epoch": (C.c_int, [vp, P(Options)]),
    "ctc_mttkrp": (C.c_int, [vp, C.c_uint32, P(C.c_float)]),
    "ctc_evaluate": (C.c_int, [vp, C.c_double, P(C.c_double), P(C.c_double), P(C.c_double),
                               P(C.c_double), P(C.c_double)]),
    "ctc_grad": (C.c_int, [vp, P(vp), P(vp), C.c_double, C.c_int, P(C.c_double), vp]),
    "ctc_grad_noalloc": (C.c_int, [vp, P(vp), P(vp), C.c_double, C.c_int, vp, vp]),
    "ctc_ref_quirk_weights": (C.c_int, [vp, C.c_uint32, C.c_uint32, P(C.c_uint32)]),
    "ctc_tc_main_ciss": (C.c_int, [P(Coo), P(Coo), P(Coo), P(Coo), P(Coo), C.c_int, P(Options)]),
    "ctc_tc_als": (C.c_int, [P(Coo), P(Coo), P(Coo), P(Coo), P(Coo), P(P(Matrix)), P(P(Matrix)),
                             C.c_int, C.c_double, P(C.c_double), P(C.c_double), P(C.c_uint64),
                             P(C.c_uint64), P(C.c_uint64)]),
    "ctc_tc_sgd": (C.c_int, [P(Coo), P(Coo), P(Coo), P(Coo), P(Coo), P(P(Matrix)), P(P(Matrix)),
                             C.c_int, C.c_double, C.c_double, P(C.c_double), P(C.c_double),
                             P(C.c_uint64), P(C.c_uint64), P(C.c_uint64)]),
    "ctc_tc_ccd": (C.c_int, [P(Coo), P(Coo), P(Coo), P(Coo), P(Coo), P(P(Matrix)), P(P(Matrix)),
                             C.c_int, C.c_double, P(C.c_double), P(C.c_double), P(C.c_uint64),
                             P(C.c_uint64), P(C.c_uint64)]),
    "ctc_autotune": (C.c_int, [vp, P(Options), P(TuneResult)]),
    "ctc_batch_create": (C.c_int, [vp, C.c_uint32, P(P(Coo)), P(P(Coo)), P(C.c_uint32), P(vp)]),
    "ctc_batch_destroy": (None, [vp]),
    "ctc_batch_info": (C.c_int, [vp, P(C.c_uint32), P(C.c_uint32), P(C.c_uint64)]),
    "ctc_batch_class_rank": (C.c_int, [vp, C.c_uint32, P(C.c_uint32), P(C.c_uint32)]),
    "ctc_batch_set_factor": (C.c_int, [vp, C.c_uint32, C.c_uint32, P(C.c_float)]),
    "ctc_batch_get_factor": (C.c_int, [vp, C.c_uint32, C.c_uint32, P(C.c_float)]),
    "ctc_batch_autotune": (C.c_int, [vp, P(Options), P(TuneResult), P(C.c_uint32)]),
    "ctc_batch_als": (C.c_int, [vp, P(Options), C.c_uint32, C.c_int32, P(C.c_double)]),
    "ctc_batch_rmse": (C.c_int, [vp, P(C.c_double), P(C.c_double)]),
}

_lib = None
_shutting_down = False


_release_at_exit = weakref.WeakSet()


def release_at_exit(obj):
    """Objects holding HIP graphs are released before the interpreter exits:
    the HIP runtime aborts if it tears down with an instantiated graph alive."""
    _release_at_exit.add(obj)


def _on_exit():
    # Graph holders are released first, while the runtime is intact.  Other
    # objects that outlive interpreter shutdown must not call into the HIP
    # runtime while it is being torn down; the OS reclaims their memory.
    global _shutting_down
    for obj in list(_release_at_exit):
        try:
            obj._release()
        except Exception:  # noqa: BLE001
            pass
    _shutting_down = True


atexit.register(_on_exit)


def alive():
    """False once the interpreter has begun shutting down."""
    return not _shutting_down


def _share_hip_runtime_with_torch():
    """libctc and PyTorch-ROCm must share one HIP runtime per process.  Both
    runtimes carry the SONAME libamdhip64.so.7, but torch's libraries ask for
    the unversioned name, so if libctc loaded /opt/rocm's runtime first torch
    would load a second copy that cannot see the GPU.  Importing torch first
    (when installed) makes libctc bind to torch's already-loaded runtime.
    Set CTC_NO_TORCH_PRELOAD=1 to skip (torch must then not be used)."""
    if os.environ.get("CTC_NO_TORCH_PRELOAD") == "1":
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """The loaded libctc.  Raises if the library is missing (no fallback)."""
    global _lib
    if _lib is None:
        _share_hip_runtime_with_torch()
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libctc.so not built at {LIB_PATH}: run `python -m cutc_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != 0:
        msg = lib().ctc_last_error().decode(errors="replace")
        raise CtcError(status, msg)
    return status
