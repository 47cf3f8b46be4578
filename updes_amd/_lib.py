"""ctypes binding of ``librbfx.so`` (the C ABI declared in ``include/rbfx.h``).

Loading never falls back to anything: if the library is missing the import of
the compute functions raises, and every compute call checks the returned
status and raises ``RbfxError`` with the library's message.
"""
import atexit
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "librbfx.so")

RBFX_OK = 0
RBFX_ERR_ARG = -1
RBFX_ERR_TIMEOUT = -4
RBFX_ERR_STALL = -5
GETRS_NOALLOC_MAX_N = 65536  # RBFX_GETRS_NOALLOC_MAX_N of rbfx.h

POLYHARMONIC, GAUSSIAN, THIN_PLATE, MULTIQUADRIC, INV_MULTIQUADRIC = 0, 1, 2, 3, 4
VAL, DX, DY, LAP, DXX, DYY, DXY = range(7)
NOPS = 7
MAT_K, MAT_A = 0, 1
FLAG_LOCAL_SUPPORT = 1

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_d = ctypes.c_double
_f = ctypes.c_float
_sz = ctypes.c_size_t


class RbfxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("rbfx error %d: %s" % (status, msg))
        self.status = status


class Problem(ctypes.Structure):
    """Mirror of ``rbfx_problem`` (include/rbfx.h)."""
    _fields_ = [
        ("N", _i64), ("M", _i64), ("kernel", _i32), ("flags", _i32), ("param", _d),
        ("d_x", _p), ("d_y", _p),
        ("Ni", _i64), ("Nd", _i64), ("Nn", _i64), ("Nr", _i64), ("Npd", _i64),
        ("d_nx", _p), ("d_ny", _p), ("d_robin_normal_x", _p), ("d_robin_normal_y", _p),
        ("d_beta", _p), ("d_pd_i1", _p), ("d_pd_i2", _p),
        ("op_rbf_const", _d * NOPS), ("op_mon_const", _d * NOPS),
        ("d_op_rbf_row", _p * NOPS), ("d_op_mon_row", _p * NOPS),
        ("support_cols", _i64), ("d_support", _p),
    ]


class IrStats(ctypes.Structure):
    _fields_ = [("outer_iters", _i32), ("inner_iters", _i32), ("final_relres", _d),
                ("stalled", _i32), ("reserved", _i32)]


class Cardinal(ctypes.Structure):
    """Mirror of ``rbfx_cardinal``."""
    _fields_ = [("N", _i64), ("M", _i64), ("k", _i64), ("d_x", _p), ("d_y", _p), ("d_stencil", _p),
                ("d_beta", _p), ("d_alpha", _p), ("d_rs", _p), ("d_rev_ptr", _p), ("d_rev_idx", _p),
                ("d_ptp_inv", _p), ("d_union_ptr", _p), ("d_union_idx", _p), ("d_union_pos", _p)]


class RhsTerm(ctypes.Structure):
    """Mirror of ``rbfx_rhs_term``."""
    _fields_ = [("op", _i32), ("reserved", _i32), ("weight", _d), ("d_weight_row", _p), ("d_coeffs", _p)]


class ProfEntry(ctypes.Structure):
    """Mirror of ``rbfx_prof_entry``."""
    _fields_ = [("name", ctypes.c_char * 48), ("launches", _i64), ("total_ms", _d), ("flops", _d), ("bytes", _d)]


# name -> (restype, argtypes)
_SIGS = {
    "rbfx_abi_version": (ctypes.c_int, []),
    "rbfx_last_error": (ctypes.c_char_p, []),
    "rbfx_init": (ctypes.c_int, []),
    "rbfx_assemble": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.c_int, _i64, _i64, _p, _i64, _p, _i64, _p]),
    "rbfx_matvec_workspace": (_sz, [ctypes.POINTER(Problem), ctypes.c_int]),
    "rbfx_matvec": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.c_int, ctypes.c_int, _p, _p, _p, _sz, _p]),
    "rbfx_eval": (ctypes.c_int, [ctypes.c_int, _d, _i64, _i64, _p, _p, _p, _i64, _p, _p, ctypes.c_uint32, _p, _p]),
    "rbfx_eval_t": (ctypes.c_int, [ctypes.c_int, _d, _i64, _i64, _p, _p, _i64, _p, _p, ctypes.c_int, _p, _p, _p]),
    "rbfx_getrf_workspace": (_sz, [_i64, ctypes.c_int]),
    "rbfx_getrf_f64": (ctypes.c_int, [_i64, _p, _i64, _p, _p, _sz, _p, _p]),
    "rbfx_getrf_f32": (ctypes.c_int, [_i64, _p, _i64, _p, _p, _sz, _p, _p]),
    "rbfx_getrs_dinv_elems": (_i64, [_i64]),
    "rbfx_getrs_prepare_f64": (ctypes.c_int, [_i64, _p, _i64, _p, _p, _p, _p, _sz, _p]),
    "rbfx_getrs_prepare_f32": (ctypes.c_int, [_i64, _p, _i64, _p, _p, _p, _p, _sz, _p]),
    "rbfx_getrs_workspace": (_sz, [_i64]),
    "rbfx_getrs_f64": (ctypes.c_int, [ctypes.c_int, _i64, _p, _i64, _p, _p, _p, _p, _sz, _p]),
    "rbfx_getrs_f32": (ctypes.c_int, [ctypes.c_int, _i64, _p, _i64, _p, _p, _p, _p, _sz, _p]),
    "rbfx_getrs_status": (ctypes.c_int, [_i64, _p, _sz, _p]),
    "rbfx_getrs_noalloc_f64": (ctypes.c_int, [ctypes.c_int, _i64, _p, _i64, _p, _p, _p, _p]),
    "rbfx_set_wait_seconds": (ctypes.c_int, [ctypes.c_double]),
    "rbfx_rhs_workspace": (_sz, [ctypes.POINTER(Problem)]),
    "rbfx_rhs": (ctypes.c_int, [ctypes.POINTER(Problem), _d, _p, ctypes.c_int, ctypes.POINTER(RhsTerm), _p, _p, _p,
                                _sz, _p]),
    "rbfx_getrs_multi_workspace": (_sz, [_i64, _i64]),
    "rbfx_getrs_multi_f64": (ctypes.c_int, [_i64, _p, _i64, _p, _p, _i64, _p, _i64, _p, _sz, _p]),
    "rbfx_shutdown": (ctypes.c_int, []),
    "rbfx_gemv_workspace": (_sz, [ctypes.c_int, _i64, _i64]),
    "rbfx_gemv_f64": (ctypes.c_int, [ctypes.c_int, _i64, _i64, _d, _p, _i64, _p, _d, _p, _p, _sz, _p]),
    "rbfx_gmres_ir_workspace": (_sz, [_i64, _i32]),
    "rbfx_gmres_ir": (ctypes.c_int, [ctypes.c_int, _i64, ctypes.POINTER(Problem), _p, _i64, _p, _i64, _p, _p,
                                     _p, _p, _d, _i32, _i32, _p, _sz, ctypes.POINTER(IrStats), _p]),
    "rbfx_cardinal_coeffs": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.POINTER(Cardinal), _p, _sz, _p]),
    "rbfx_cardinal_form": (ctypes.c_int, [ctypes.POINTER(Cardinal), _p, _i64, _p, _i64, _p, _sz, _p]),
    "rbfx_gmres_ir_cardinal_workspace": (_sz, [_i64, _i64, _i32]),
    "rbfx_gmres_ir_cardinal": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Cardinal), _p, _i64, _d, _p, _i64, _p,
                                              _p, _p, _p, _d, _i32, _i32, _p, _sz, ctypes.POINTER(IrStats), _p]),
    "rbfx_frobenius_workspace": (_sz, [_i64]),
    "rbfx_frobenius_f64": (ctypes.c_int, [_i64, _i64, _p, _i64, _p, _i64, ctypes.POINTER(_d), _p, _sz, _p]),
    "rbfx_frobenius_pair_f64": (ctypes.c_int, [_i64, _i64, _p, _i64, _p, _i64, ctypes.POINTER(_d),
                                               ctypes.POINTER(_d), _p, _sz, _p]),
    "rbfx_profile_begin": (ctypes.c_int, []),
    "rbfx_profile_end": (ctypes.c_int, [ctypes.POINTER(ProfEntry), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "rbfx_gemm_f64": (ctypes.c_int, [_i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64, _p]),
    "rbfx_gemm_f32": (ctypes.c_int, [_i64, _i64, _i64, _f, _p, _i64, _p, _i64, _f, _p, _i64, _p]),
    "rbfx_set_panel_groups": (ctypes.c_int, [ctypes.c_int]),
    "rbfx_set_narrow_leaf_rows": (ctypes.c_int, [_i64]),
}

EXPORTED = sorted(_SIGS)

_lib = None


def load():
    """Load librbfx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("librbfx.so not built: run `python -m updes_amd.build` (%s missing)" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        atexit.register(_shutdown)
    return _lib


def _shutdown():
    # release the library's own streams and events while the HIP runtime is
    # still up (only if this process ever touched the GPU)
    if _lib is not None and torch.cuda.is_initialized():
        _lib.rbfx_shutdown()


def lib():
    return load()


def check(status):
    if status != RBFX_OK:
        msg = load().rbfx_last_error().decode(errors="replace")
        raise RbfxError(status, msg)
    return status


def require_gpu():
    """The compute path exists only on a HIP device; there is no CPU path."""
    if not torch.cuda.is_available():
        raise RuntimeError("updes_amd requires a ROCm GPU (gfx950); no CPU fallback exists")
    load()


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    """Device pointer of a tensor (or None)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def workspace(nbytes, device):
    nbytes = max(int(nbytes), 256)
    return torch.empty(nbytes, dtype=torch.uint8, device=device)
