"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker and the CPU baseline; the GPU product never
does. The oracle restates matops' algorithm (see oracle.h): the Go glue in
matops_port.c over either the restated calgo kernels (calgo_port.c,
"port") or the reference's own calgo C built into _ref/ ("reference").

Matrix views are column-major float64 numpy arrays (strides[0] == 8); the
leading dimension is strides[1] // 8, as in matops_amd.matops.
"""
import ctypes
import os
import platform
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF = os.path.join(HERE, "_ref", "libcalgo_ref.so")

TRANSA, TRANSB, LOWER, UPPER = 0x1, 0x2, 0x4, 0x8
LEFT, RIGHT, UNIT, TRANS = 0x10, 0x20, 0x40, 0x80


class OView(ctypes.Structure):
    _fields_ = [("p", ctypes.c_void_p), ("rows", ctypes.c_int), ("cols", ctypes.c_int),
                ("ld", ctypes.c_int)]


_lib = None
_kind = None


def build():
    """Builds liboracle.so (and _ref/ when the reference is present)."""
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        V = OView
        sig = {
            "o_use_kernels": (I, [I, ctypes.c_char_p]),
            "o_kernels_in_use": (I, []),
            "o_use_kernel_mask": (I, [I, ctypes.c_char_p]),
            "o_BlockingParams": (None, [I, I, I]),
            "o_NumWorkers": (I, [I]),
            "o_Mult": (I, [V, V, V, D, D, I]),
            "o_SolveTrm": (I, [V, V, D, I]),
            "o_RankUpdateSym": (I, [V, V, D, D, I]),
            "o_DecomposeCHOL": (I, [V, I, I, P, P]),
            "o_SolveCHOL": (None, [V, V, I]),
            "o_DecomposeLU": (I, [V, P, I, I]),
            "o_SolveLU": (I, [V, V, P, I, I]),
            "o_MultSym": (I, [V, V, V, D, D, I]),
            "o_MultTrm": (I, [V, V, D, I]),
            "o_RankUpdate2Sym": (I, [V, V, V, D, D, I]),
            "o_UpdateTrm": (I, [V, V, V, D, D, I]),
            "o_DecomposeLUnoPiv": (I, [V, I]),
            "o_DecomposeLDLnoPiv": (I, [V, I, I]),
            "o_SolveLDLnoPiv": (None, [V, V, I]),
            "o_InverseTrm": (I, [V, I, I]),
            "o_chol_adjoint": (None, [I, P, I, P, I, P, I]),
            "o_now": (D, []),
            "o_kern_gemm": (None, [P, I, P, I, P, I, D, D, I, I, I, I, I, I]),
            "o_kern_trsm": (None, [P, I, P, I, D, I, I, I, I]),
            "o_kern_trsm_unb": (None, [P, I, P, I, D, I, I, I, I]),
            "o_kern_syrk": (None, [P, I, P, I, D, D, I, I, I, I]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        use("port")
    return _lib


def use(kind):
    """Selects the kernels: "port" (restatement) or "reference" (calgo C
    built from the reference into _ref/). Returns the kind now in use."""
    global _kind
    L = lib() if _lib is None else _lib
    if kind == "reference":
        if not os.path.exists(REF) or L.o_use_kernels(1, REF.encode()) != 0:
            raise RuntimeError(f"reference calgo not available at {REF}")
    else:
        L.o_use_kernels(0, None)
    _kind = kind
    return _kind


def reference_available():
    return os.path.exists(REF)


def _v(X):
    if X.dtype != np.float64 or X.ndim != 2:
        raise TypeError("float64 2-D view expected")
    r, c = X.shape
    if r > 1 and X.strides[0] != 8:
        raise ValueError("rows must be contiguous")
    ld = X.strides[1] // 8 if c > 1 else max(r, 1)
    return OView(X.ctypes.data, r, c, max(ld, 1))


def BlockingParams(mb, nb, kb):
    lib().o_BlockingParams(mb, nb, kb)


def NumWorkers(w):
    return lib().o_NumWorkers(w)


def Mult(C, A, B, alpha, beta, flags):
    return lib().o_Mult(_v(C), _v(A), _v(B), alpha, beta, flags)


def SolveTrm(B, A, alpha, flags):
    return lib().o_SolveTrm(_v(B), _v(A), alpha, flags)


def RankUpdateSym(C, A, alpha, beta, flags):
    return lib().o_RankUpdateSym(_v(C), _v(A), alpha, beta, flags)


def MultSym(C, A, B, alpha, beta, flags):
    return lib().o_MultSym(_v(C), _v(A), _v(B), alpha, beta, flags)


def MultTrm(B, A, alpha, flags):
    return lib().o_MultTrm(_v(B), _v(A), alpha, flags)


def RankUpdate2Sym(C, A, B, alpha, beta, flags):
    return lib().o_RankUpdate2Sym(_v(C), _v(A), _v(B), alpha, beta, flags)


def UpdateTrm(C, A, B, alpha, beta, flags):
    return lib().o_UpdateTrm(_v(C), _v(A), _v(B), alpha, beta, flags)


def DecomposeLUnoPiv(A, nb):
    return lib().o_DecomposeLUnoPiv(_v(A), nb)


def DecomposeLDLnoPiv(A, flags, nb):
    return lib().o_DecomposeLDLnoPiv(_v(A), flags, nb)


def SolveLDLnoPiv(B, A, flags):
    lib().o_SolveLDLnoPiv(_v(B), _v(A), flags)


def InverseTrm(A, flags, nb):
    return lib().o_InverseTrm(_v(A), flags, nb)


def DecomposeCHOL(A, flags, nb):
    """Returns (code, panic_index, panic_value); code 2 = the reference's
    panic("illegal value at %d: %e")."""
    idx = ctypes.c_int(-1)
    val = ctypes.c_double(0.0)
    rc = lib().o_DecomposeCHOL(_v(A), flags, nb, ctypes.addressof(idx), ctypes.addressof(val))
    return rc, idx.value, val.value


def SolveCHOL(B, A, flags):
    lib().o_SolveCHOL(_v(B), _v(A), flags)


def DecomposeLU(A, pivots, nb):
    assert pivots.dtype == np.int64 and pivots.flags.c_contiguous
    return lib().o_DecomposeLU(_v(A), pivots.ctypes.data, len(pivots), nb)


def SolveLU(B, A, pivots, flags):
    assert pivots.dtype == np.int64
    return lib().o_SolveLU(_v(B), _v(A), pivots.ctypes.data, len(pivots), flags)


def _ld(X):
    if X.dtype != np.float64 or X.ndim != 2 or (X.shape[0] > 1 and X.strides[0] != 8):
        raise TypeError("column-major float64 matrix expected")
    return max(X.strides[1] // 8 if X.shape[1] > 1 else X.shape[0], 1)


def kern_gemm(C, A, B, alpha, beta, flags, P, S, L, R, E):
    """dmult_mm_blocked3 on C[R:E, S:L] with the selected kernels
    (calgo/cmops.h:133-138). Arrays are the parents; indices as calgo's."""
    lib().o_kern_gemm(C.ctypes.data, _ld(C), A.ctypes.data, _ld(A), B.ctypes.data, _ld(B),
                      alpha, beta, flags, P, S, L, R, E)


def kern_trsm(B, A, alpha, flags, N, S, E, unblocked=False):
    """dmmat_solve_blk (or dmmat_solve_unb) on columns (LEFT) or rows
    (RIGHT) S:E of B (calgo/cmops.h:203-209)."""
    f = lib().o_kern_trsm_unb if unblocked else lib().o_kern_trsm
    f(B.ctypes.data, _ld(B), A.ctypes.data, _ld(A), alpha, flags, N, S, E)


def kern_syrk(C, A, alpha, beta, flags, P, S, E):
    """dmmat_rank_blk on the block S:E (calgo/cmops.h:212-214)."""
    lib().o_kern_syrk(C.ctypes.data, _ld(C), A.ctypes.data, _ld(A), alpha, beta, flags, P, S, E)


def chol_adjoint(L, Lbar):
    n = L.shape[0]
    out = np.zeros((n, n), order="F")
    Lc = np.asfortranarray(L)
    Bc = np.asfortranarray(Lbar)
    lib().o_chol_adjoint(n, Lc.ctypes.data, n, Bc.ctypes.data, n, out.ctypes.data, n)
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def bench_cholesky_sample(nb=256, n=5120, nrhs=64, seed=3):
    """Times DecomposeCHOL(A, LOWER, nb) + SolveCHOL (nrhs RHS) once on one
    host core, on an n x n sample of the same kind of input as the GPU
    workload (A = G G^T / n + I). Uses the reference's calgo kernels when
    they were built, else the restated ones. Returns the cpu_baseline
    record (TFLOP/s with the same LAPACK flop count as the GPU figure)."""
    kind = "reference" if reference_available() else "port"
    use(kind)
    NumWorkers(1)
    g = np.random.Generator(np.random.Philox(seed))
    G = g.standard_normal((n, n))
    A0 = np.asfortranarray(G @ G.T / n + np.eye(n))
    A0 = np.asfortranarray(np.tril(A0) + np.tril(A0, -1).T)
    B0 = np.asfortranarray(g.standard_normal((n, nrhs)))
    A = A0.copy(order="F")
    B = B0.copy(order="F")
    t0 = lib().o_now()
    rc, _, _ = DecomposeCHOL(A, LOWER, nb)
    SolveCHOL(B, A, LOWER)
    t1 = lib().o_now()
    if rc != 0:
        raise RuntimeError("oracle Cholesky failed on the sample")
    flops = n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0 + 2.0 * n * n * nrhs
    return {
        "value": round(flops / (t1 - t0) / 1e12, 6),
        "unit": "TFLOP/s",
        "cores": 1,
        "kind": kind,
        "sample": f"DecomposeCHOL(A, LOWER, nb={nb}) + SolveCHOL, n={n}, {nrhs} RHS, "
                  f"one run, {t1 - t0:.2f} s on {_cpu_model()}",
        "seconds": round(t1 - t0, 3),
    }
