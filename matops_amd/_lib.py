"""ctypes binding of libmops.so (the C ABI declared in include/mops.h).

The library is loaded from this package directory. There is no fallback:
if libmops.so is missing, importing a function that needs it raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmops.so")


class MData(ctypes.Structure):
    """mops_mdata_t / calgo mdata_t: pointer to element (0,0) and leading dim."""
    _fields_ = [("md", ctypes.c_void_p), ("step", ctypes.c_int)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_PM = ctypes.POINTER(MData)

class IrReport(ctypes.Structure):
    """mops_ir_report."""
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("backward_error", ctypes.c_double), ("first_error", ctypes.c_double),
                ("factor_ms", ctypes.c_double), ("refine_ms", ctypes.c_double)]


_PR = ctypes.POINTER(IrReport)

# name -> (restype, argtypes)
SIGNATURES = {
    "mops_last_error": (ctypes.c_char_p, []),
    "mops_device_count": (_I, []),
    "mops_thread_release": (None, []),
    "mops_d_status": (_I, [_P]),
    "mops_trace_enable": (None, [_I]),
    "mops_trace_collect": (_I, [_I, _P, _P, _P]),
    "mops_dmult_mm_blocked3": (_I, [_PM, _PM, _PM, _D, _D, _I, _I, _I, _I, _I, _I, _I, _I, _I]),
    "mops_d_dgemm": (_I, [_I, _I, _I, _I, _D, _P, _I, _P, _I, _D, _P, _I, _P]),
    "mops_d_gemm_bf16": (_I, [_I, _I, _I, ctypes.c_float, _P, _I, _P, _I, _P, _I, _P]),
    "mops_d_dgemm_noalloc": (_I, [_I, _I, _I, _I, _D, _P, _I, _P, _I, _D, _P, _I, _P]),
    "mops_dmmat_solve_blk": (_I, [_PM, _PM, _D, _I, _I, _I, _I, _I]),
    "mops_dmmat_solve_unb": (_I, [_PM, _PM, _D, _I, _I, _I, _I]),
    "mops_d_dtrsm": (_I, [_I, _I, _I, _D, _P, _I, _P, _I, _P]),
    "mops_dmmat_rank_blk": (_I, [_PM, _PM, _D, _D, _I, _I, _I, _I, _I, _I]),
    "mops_d_dsyrk": (_I, [_I, _I, _I, _D, _P, _I, _D, _P, _I, _P]),
    "mops_dmult_symm_blocked2": (_I, [_PM, _PM, _PM, _D, _D, _I, _I, _I, _I, _I, _I, _I, _I,
                                      _I]),
    "mops_d_dsymm": (_I, [_I, _I, _I, _D, _P, _I, _P, _I, _D, _P, _I, _P]),
    "mops_dmmat_trmm_blk": (_I, [_PM, _PM, _D, _I, _I, _I, _I, _I]),
    "mops_d_dtrmm": (_I, [_I, _I, _I, _D, _P, _I, _P, _I, _P]),
    "mops_dmmat_rank2_blk": (_I, [_PM, _PM, _PM, _D, _D, _I, _I, _I, _I, _I, _I]),
    "mops_d_dsyr2k": (_I, [_I, _I, _I, _D, _P, _I, _P, _I, _D, _P, _I, _P]),
    "mops_dmmat_trmupd_blk": (_I, [_PM, _PM, _PM, _D, _D, _I, _I, _I, _I, _I, _I]),
    "mops_d_dtrmupd": (_I, [_I, _I, _I, _D, _P, _I, _P, _I, _D, _P, _I, _P]),
    "mops_dgetrf_nopiv": (_I, [_PM, _I, _I, _I]),
    "mops_d_dgetrf_nopiv": (_I, [_I, _I, _I, _P, _I, _P]),
    "mops_dsytrf_nopiv": (_I, [_PM, _I, _I, _I]),
    "mops_d_dsytrf_nopiv": (_I, [_I, _I, _I, _P, _I, _P]),
    "mops_dsytrs_nopiv": (_I, [_PM, _PM, _I, _I, _I]),
    "mops_d_dsytrs_nopiv": (_I, [_I, _I, _I, _P, _I, _P, _I, _P]),
    "mops_dsytrf_ldl": (_I, [_PM, _P, _I, _I, _I, _I]),
    "mops_d_dsytrf_ldl": (_I, [_I, _I, _I, _P, _I, _P, _I, _P, _P]),
    "mops_dsytrs_ldl": (_I, [_PM, _PM, _P, _I, _I, _I, _I]),
    "mops_d_dsytrs_ldl": (_I, [_I, _I, _I, _P, _I, _P, _I, _P, _I, _P]),
    "mops_dsytrf_bk": (_I, [_PM, _P, _I, _I, _I, _I]),
    "mops_d_dsytrf_bk": (_I, [_I, _I, _I, _P, _I, _P, _I, _P]),
    "mops_dsytrs_bk": (_I, [_PM, _PM, _P, _I, _I, _I, _I]),
    "mops_d_dsytrs_bk": (_I, [_I, _I, _I, _P, _I, _P, _I, _P, _I, _P]),
    "mops_dinverse_trm": (_I, [_PM, _I, _I, _I]),
    "mops_d_dinverse_trm": (_I, [_I, _I, _I, _P, _I, _P]),
    "mops_dpotrf": (_I, [_PM, _I, _I, _I]),
    "mops_dpotrs": (_I, [_PM, _PM, _I, _I, _I]),
    "mops_d_dpotrf": (_I, [_I, _I, _I, _P, _I, _P, _P]),
    "mops_d_dpotrs": (_I, [_I, _I, _I, _P, _I, _P, _I, _P]),
    "mops_dpotrf_adjoint": (_I, [_PM, _PM, _I, _I]),
    "mops_dgesv_ir": (_I, [_PM, _PM, _PM, _I, _I, _I, _D, _I, _PR]),
    "mops_d_dgesv_ir": (_I, [_I, _I, _P, _I, _P, _I, _P, _I, _I, _D, _I, _PR, _P]),
    "mops_d_dpotrf_adjoint": (_I, [_I, _I, _P, _I, _P, _I, _P]),
    "mops_dgetrf": (_I, [_PM, _P, _I, _I, _I, _I]),
    "mops_dgetrs": (_I, [_PM, _PM, _P, _I, _I, _I, _I]),
    "mops_d_dgetrf": (_I, [_I, _I, _I, _P, _I, _P, _P, _P, _P]),
    "mops_d_dgetrs": (_I, [_I, _I, _I, _P, _I, _P, _I, _P, _P, _I, _P]),
}

_lib = None


def lib():
    """Loads libmops.so once and declares every signature.

    When torch is importable it is imported first: torch ships its own HIP
    runtime, and loading it before libmops makes both share that one runtime
    instead of two copies competing for the device."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                f"{LIBPATH} is not built; run `python -m matops_amd.build` first")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class MopsError(RuntimeError):
    def __init__(self, func, code, msg):
        super().__init__(f"{func} returned {code}: {msg}")
        self.code = code
        self.msg = msg


def check(func, code):
    if code != 0:
        msg = lib().mops_last_error()
        raise MopsError(func, code, msg.decode() if msg else "")
    return code


def last_error():
    msg = lib().mops_last_error()
    return msg.decode() if msg else ""
