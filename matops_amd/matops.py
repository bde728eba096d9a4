"""The matops API of the dense FP64 path, on top of the C ABI.

Mirrors the Go functions the path consists of -- names, argument meaning,
validation and error behaviour -- with column-major numpy views standing in
for ``matrix.FloatMatrix``: any 2-D float64 array whose rows are contiguous
(``strides[0] == 8``) is a view; its leading dimension is
``strides[1] // 8``, exactly the (pointer at (0,0), step) pair calgo gets
(calgo/cmops.go:68-91). Sub-views of a Fortran-ordered array
(``A[r0:r1, c0:c1]``) therefore share memory with their parent the way
FloatMatrix.SubMatrix does.

Everything here computes on the GPU through libmops.so; there is no CPU
fallback.
"""
import ctypes
import math
import threading

import numpy as np

from . import _lib
from ._lib import MData, check

TRANSA = 0x1
TRANSB = 0x2
LOWER = 0x4
UPPER = 0x8
LEFT = 0x10
RIGHT = 0x20
UNIT = 0x40
TRANS = 0x80
NOTRANS = 0
NONE = 0

# CPU blocking hints the reference passes down (mmult.go:33-35). The GPU
# ignores them; they are kept so BlockingParams keeps its meaning.
_vpLen, _nB, _mB = 196, 68, 68
_nWorker = 1
_limitOne = 200 * 200 * 200  # problems below this run in one call (mmult.go:39)
_panicOnError = False


class MatopsError(Exception):
    """The error value a Go call would return."""


def SetPanicOnError(newval):
    """mmult.go:55-57. With panic-on-error, errors raise RuntimeError
    (Go panic) instead of being returned."""
    global _panicOnError
    _panicOnError = bool(newval)


def BlockingParams(mb, nb, kb):
    """mmult.go:63-67 (hints only on the GPU)."""
    global _vpLen, _nB, _mB
    _vpLen, _nB, _mB = kb, nb, mb


def NumWorkers(newWorkers):
    """mmult.go:69-73. Returns the previous value. With more than one
    worker, Mult splits C into blocks and calls the kernel-level entry point
    from that many host threads at once, as the Go code starts goroutines
    (mmult.go:246-254); each thread has its own HIP stream and workspace."""
    global _nWorker
    old = _nWorker
    _nWorker = newWorkers
    return old


def blockIndex4(i, r, sz):
    """mmult.go:105-110: start of block i of r over sz, rounded down to a
    multiple of 4 (the last block ends at sz)."""
    if i == r:
        return sz
    v = i * sz // r
    return v - (v & 0x3)


def divideWork(rows, cols, workers):
    """mmult.go:137-167: (column workers, row workers) for a C of rows x
    cols."""
    nwsqrt = int(math.sqrt(float(workers)))
    issquare = nwsqrt * nwsqrt == workers
    if workers == 2 or (workers & 0x1) != 0:
        return (workers, 1) if cols > rows else (1, workers)
    if issquare:
        return nwsqrt, nwsqrt
    return (workers // 2, 2) if cols > rows else (2, workers // 2)


def scheduleWork(colworks, rowworks, cols, rows, worker):
    """mmult.go:171-189: runs worker(colstart, colend, rowstart, rowend) for
    every block at once, each on its own thread, and waits for all. The
    first failure is re-raised."""
    errs = []

    def run(*blk):
        try:
            worker(*blk)
        except Exception as e:  # noqa: BLE001 - re-raised below
            errs.append(e)

    ths = []
    for k in range(colworks):
        cs, ce = blockIndex4(k, colworks, cols), blockIndex4(k + 1, colworks, cols)
        for l in range(rowworks):
            rs, re_ = blockIndex4(l, rowworks, rows), blockIndex4(l + 1, rowworks, rows)
            t = threading.Thread(target=run, args=(cs, ce, rs, re_))
            t.start()
            ths.append(t)
    for t in ths:
        t.join()
    if errs:
        raise errs[0]


def _onError(msg):
    # mmult.go:46-51
    if _panicOnError:
        raise RuntimeError(msg)
    return MatopsError(msg)


def _view(X):
    """(MData, rows, cols) of a column-major float64 view."""
    if not isinstance(X, np.ndarray) or X.dtype != np.float64 or X.ndim != 2:
        raise TypeError("matrix views are 2-D float64 numpy arrays")
    rows, cols = X.shape
    if rows > 1 and X.strides[0] != 8:
        raise ValueError("view rows must be contiguous (column-major storage)")
    ld = X.strides[1] // 8 if cols > 1 else max(rows, 1)
    if cols > 1 and (X.strides[1] % 8 or ld < rows):
        raise ValueError("bad column stride for a column-major view")
    return MData(ctypes.c_void_p(X.ctypes.data), max(ld, 1)), rows, cols


def Mult(C, A, B, alpha, beta, flags):
    """C = beta*C + alpha*op(A)*op(B)  (mmult.go:198-256)."""
    ar, ac = A.shape
    br, bc = B.shape
    cr, cc = C.shape
    tt = flags & (TRANSA | TRANSB)
    if tt == TRANSA | TRANSB:
        empty = ac == 0 or br == 0
        ok = cr == ac and cc == br and ar == bc
    elif tt == TRANSA:
        empty = ac == 0 or bc == 0
        ok = cr == ac and cc == bc and ar == br
    elif tt == TRANSB:
        empty = ar == 0 or br == 0
        ok = cr == ar and cc == br and ac == bc
    else:
        empty = ar == 0 or bc == 0
        ok = cr == ar and cc == bc and ac == br
    if empty:
        return None
    if not ok:
        return MatopsError("Mult: size mismatch")
    P = A.shape[0] if flags & TRANSA else A.shape[1]
    psize = C.size * A.shape[1]  # as mmult.go:227 counts it
    Cm, _, _ = _view(C)
    Am, _, _ = _view(A)
    Bm, _, _ = _view(B)
    L = _lib.lib()

    def worker(cs, ce, rs, re_):
        check("mops_dmult_mm_blocked3", L.mops_dmult_mm_blocked3(
            ctypes.byref(Cm), ctypes.byref(Am), ctypes.byref(Bm), alpha, beta, flags,
            P, cs, ce, rs, re_, _vpLen, _nB, _mB))

    if _nWorker <= 1 or psize <= _limitOne:
        worker(0, cc, 0, cr)
        return None
    colworks, rowworks = divideWork(cr, cc, _nWorker)
    scheduleWork(colworks, rowworks, cc, cr, worker)
    return None


def MultSym(C, A, B, alpha, beta, flags):
    """C = beta*C + alpha*A*B (LEFT) or alpha*B*A (RIGHT) with A symmetric,
    one triangle stored (mmult.go:266-326). Validation as the reference's
    (it checks the shapes of C = A*B for every flag combination)."""
    ar, ac = A.shape
    br, bc = B.shape
    cr, cc = C.shape
    tt = flags & (TRANSA | TRANSB)
    if tt == TRANSA | TRANSB:
        empty = ac == 0 or br == 0
        ok = ar == ac and cr == ac and cc == br and ar == bc
    elif tt == TRANSA:
        empty = ac == 0 or bc == 0
        ok = ar == ac and cr == ac and cc == bc and ar == br
    elif tt == TRANSB:
        empty = ar == 0 or br == 0
        ok = ar == ac and cr == ar and cc == br and ac == bc
    else:
        empty = ar == 0 or bc == 0
        ok = ar == ac and cr == ar and cc == bc and ac == br
    if empty:
        return None
    if not ok:
        return MatopsError("MultSym: size mismatch")
    psize = C.size * A.shape[1]
    Cm, _, _ = _view(C)
    Am, _, _ = _view(A)
    Bm, _, _ = _view(B)
    L = _lib.lib()

    def worker(cs, ce, rs, re_):
        check("mops_dmult_symm_blocked2", L.mops_dmult_symm_blocked2(
            ctypes.byref(Cm), ctypes.byref(Am), ctypes.byref(Bm), alpha, beta, flags,
            ac, cs, ce, rs, re_, _vpLen, _nB, _mB))

    if _nWorker <= 1 or psize <= _limitOne:
        worker(0, cc, 0, cr)
        return None
    colworks, rowworks = divideWork(cr, cc, _nWorker)
    scheduleWork(colworks, rowworks, cc, cr, worker)
    return None


def MultTrm(B, A, alpha, flags):
    """B = alpha*op(A)*B (LEFT) or alpha*B*op(A) (RIGHT), A triangular
    (mmult.go:338-372)."""
    br, bc = B.shape
    ar, ac = A.shape
    empty, ok = False, True
    side = flags & (LEFT | RIGHT)
    if side == LEFT:
        empty = br == 0 or bc == 0
        ok = br == ac and ac == ar
    elif side == RIGHT:
        empty = bc == 0 or br == 0
        ok = bc == ar and ac == ar
    if empty:
        return None
    if not ok:
        return _onError("A, B size mismatch")
    E = br if flags & RIGHT else bc
    if ac == 0 or E <= 0:  # calgo/dmmat_trid.c:790-791
        return None
    Bm, _, _ = _view(B)
    Am, _, _ = _view(A)
    check("mops_dmmat_trmm_blk", _lib.lib().mops_dmmat_trmm_blk(
        ctypes.byref(Bm), ctypes.byref(Am), alpha, flags, ac, 0, E, _nB))
    return None


def RankUpdate2Sym(C, A, B, alpha, beta, flags):
    """C = beta*C + alpha*A*B^T + alpha*B*A^T (alpha*A^T*B + alpha*B^T*A
    with TRANSA) on one triangle (mmult.go:453-476)."""
    if C.shape[0] != C.shape[1]:
        return _onError("C not a square matrix")
    P = A.shape[0] if flags & TRANSA else A.shape[1]
    E = C.shape[0]
    if P == 0 or E == 0:
        return None
    Cm, _, _ = _view(C)
    Am, _, _ = _view(A)
    Bm, _, _ = _view(B)
    check("mops_dmmat_rank2_blk", _lib.lib().mops_dmmat_rank2_blk(
        ctypes.byref(Cm), ctypes.byref(Am), ctypes.byref(Bm), alpha, beta, flags, P, 0, E,
        _vpLen, _nB))
    return None


def UpdateTrm(C, A, B, alpha, beta, flags):
    """C = beta*C + alpha*op(A)*op(B) on the LOWER or UPPER triangle of C
    (mmult.go:501-520)."""
    if C.shape[0] != C.shape[1]:
        return _onError("C not a square matrix")
    P = A.shape[0] if flags & TRANSA else A.shape[1]
    E = C.shape[0]
    if P == 0 or E == 0:
        return None
    Cm, _, _ = _view(C)
    Am, _, _ = _view(A)
    Bm, _, _ = _view(B)
    check("mops_dmmat_trmupd_blk", _lib.lib().mops_dmmat_trmupd_blk(
        ctypes.byref(Cm), ctypes.byref(Am), ctypes.byref(Bm), alpha, beta, flags, P, 0, E,
        _vpLen, _nB))
    return None


def SolveTrm(B, A, alpha, flags):
    """B = alpha*op(A)^-1*B or alpha*B*op(A)^-1  (mmult.go:388-422)."""
    br, bc = B.shape
    ar, ac = A.shape
    empty, ok = False, True
    side = flags & (LEFT | RIGHT)
    if side == LEFT:
        empty = br == 0
        ok = br == ac and ac == ar
    elif side == RIGHT:
        empty = bc == 0
        ok = bc == ar and ac == ar
    if empty:
        return None
    if not ok:
        return _onError("A, B size mismatch")
    E = br if flags & RIGHT else bc
    # calgo.DSolveBlk returns early when N == 0 or E-S <= 0 (cmops.go:203)
    if ac == 0 or E <= 0:
        return None
    Bm, _, _ = _view(B)
    Am, _, _ = _view(A)
    check("mops_dmmat_solve_blk", _lib.lib().mops_dmmat_solve_blk(
        ctypes.byref(Bm), ctypes.byref(Am), alpha, flags, ac, 0, E, _nB))
    return None


def RankUpdateSym(C, A, alpha, beta, flags):
    """C = beta*C + alpha*A*A^T (alpha*A^T*A with TRANSA) on one triangle
    (mmult.go:426-445)."""
    if C.shape[0] != C.shape[1]:
        return _onError("C not a square matrix")
    P = A.shape[0] if flags & TRANSA else A.shape[1]
    E = C.shape[0]
    if P == 0 or E == 0:
        return None
    Cm, _, _ = _view(C)
    Am, _, _ = _view(A)
    check("mops_dmmat_rank_blk", _lib.lib().mops_dmmat_rank_blk(
        ctypes.byref(Cm), ctypes.byref(Am), alpha, beta, flags, P, 0, E, _vpLen, _nB))
    return None


def DecomposeCHOL(A, flags, nb):
    """Cholesky in place (chol.go:116-127). Returns (A, err). A NaN square
    root panics with the reference's message (chol.go:33-36)."""
    if A.shape[0] != A.shape[1]:
        return A, MatopsError("A not a square matrix")
    n = A.shape[0]
    if n == 0:
        return A, None
    Am, _, _ = _view(A)
    rc = _lib.lib().mops_dpotrf(ctypes.byref(Am), flags, n, nb)
    if rc > 0:
        raise RuntimeError(_lib.last_error())  # Go panic
    check("mops_dpotrf", rc)
    return A, None


def SolveCHOL(B, A, flags):
    """B = A^-1 B from the Cholesky factor (chol.go:147-158)."""
    n = A.shape[0]
    if n == 0 or B.shape[1] == 0:
        return
    Bm, _, _ = _view(B)
    Am, _, _ = _view(A)
    check("mops_dpotrs", _lib.lib().mops_dpotrs(
        ctypes.byref(Bm), ctypes.byref(Am), flags, n, B.shape[1]))


def CholeskyAdjoint(L, Lbar, nb=0):
    """Reverse-mode derivative of DecomposeCHOL(A, LOWER, nb) (no counterpart
    in the reference). L: lower Cholesky factor; Lbar: d loss / dL (lower
    triangle read). Lbar is overwritten with the symmetric d loss / dA and
    returned."""
    n = L.shape[0]
    if L.shape != (n, n) or Lbar.shape != (n, n):
        return Lbar, MatopsError("L, Lbar size mismatch")
    if n == 0:
        return Lbar, None
    Lm, _, _ = _view(L)
    Bm, _, _ = _view(Lbar)
    check("mops_dpotrf_adjoint", _lib.lib().mops_dpotrf_adjoint(
        ctypes.byref(Lm), ctypes.byref(Bm), n, nb))
    return Lbar, None


PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16 = 2


def SolveIR(X, A, B, prec=PREC_FP32, tol=1e-12, maxit=10):
    """X = A^-1 B by LU in precision prec plus FP64 iterative refinement
    (no counterpart in the reference). A and B are not modified. Returns the
    report as a dict (iterations, converged, backward_error, first_error,
    factor_ms, refine_ms)."""
    n = A.shape[0]
    if A.shape != (n, n) or B.shape[0] != n or X.shape != B.shape:
        raise ValueError("SolveIR: size mismatch")
    rep = _lib.IrReport()
    if n == 0 or B.shape[1] == 0:
        return {"iterations": 0, "converged": 1, "backward_error": 0.0, "first_error": 0.0,
                "factor_ms": 0.0, "refine_ms": 0.0}
    Am, _, _ = _view(A)
    Bm, _, _ = _view(B)
    Xm, _, _ = _view(X)
    check("mops_dgesv_ir", _lib.lib().mops_dgesv_ir(
        ctypes.byref(Am), ctypes.byref(Bm), ctypes.byref(Xm), n, B.shape[1], prec, tol, maxit,
        ctypes.byref(rep)))
    return {f: getattr(rep, f) for f, _ in rep._fields_}


def DecomposeLUnoPiv(A, nb):
    """LU without pivoting in place (lu.go:291-299). Returns (A, err)."""
    if min(A.shape) == 0:
        return A, None
    Am, _, _ = _view(A)
    check("mops_dgetrf_nopiv", _lib.lib().mops_dgetrf_nopiv(
        ctypes.byref(Am), A.shape[0], A.shape[1], nb))
    return A, None


def DecomposeLDLnoPiv(A, W, flags, nb):
    """A = L D L^T (LOWER) or U D U^T without pivoting, in place
    (ldlnp.go:211-241). W: the workspace the reference's blocked form needs
    (checked as the reference does; the GPU keeps its own). Returns (A, err)
    with A None on error, as the reference."""
    if A.shape[0] != A.shape[1]:
        return None, MatopsError("A not a square matrix")
    n = A.shape[0]
    if not (n < nb or nb == 0) and W is None:
        return None, MatopsError("No workspace for blocking invocation")
    if n == 0:
        return A, None
    Am, _, _ = _view(A)
    check("mops_dsytrf_nopiv", _lib.lib().mops_dsytrf_nopiv(ctypes.byref(Am), flags, n, nb))
    return A, None


def SolveLDLnoPiv(B, A, flags):
    """B = A^-1 B from DecomposeLDLnoPiv's factors (ldlnp.go:251-264)."""
    n = A.shape[0]
    if n == 0 or B.shape[1] == 0:
        return
    Bm, _, _ = _view(B)
    Am, _, _ = _view(A)
    check("mops_dsytrs_nopiv", _lib.lib().mops_dsytrs_nopiv(
        ctypes.byref(Bm), ctypes.byref(Am), flags, n, B.shape[1]))


def _pivot_array(ipiv):
    if not isinstance(ipiv, np.ndarray) or ipiv.dtype != np.int64 \
            or not ipiv.flags.c_contiguous or ipiv.ndim != 1:
        raise TypeError("ipiv must be a contiguous 1-D int64 array (Go []int)")
    return ctypes.c_void_p(ipiv.ctypes.data) if len(ipiv) else None


def DecomposeLDL(A, W, ipiv, flags, nb):
    """A = P^T L D L^T P (LOWER) or P^T U D U^T P with 1x1 pivots, in place
    (ldl.go:439-460). W: the reference's workspace for the blocked forms
    (not read; the GPU keeps its own). ipiv: int64 array of N entries,
    1-based, 0 = no interchange. Returns (A, err); a zero pivot gives the
    reference's error, or raises with SetPanicOnError (ldl.go:60, 335)."""
    if A.shape[0] != A.shape[1]:
        return None, MatopsError("A not a square matrix")
    p = _pivot_array(ipiv)
    ipiv[:] = 0
    n = A.shape[0]
    if n == 0:
        return A, None
    if len(ipiv) < n:
        raise IndexError("index out of range")  # the reference's slice access
    Am, _, _ = _view(A)
    rc = _lib.lib().mops_dsytrf_ldl(ctypes.byref(Am), p, len(ipiv), flags, n, nb)
    if rc > 0:
        return A, _onError(_lib.last_error())
    check("mops_dsytrf_ldl", rc)
    return A, None


def SolveLDL(B, A, ipiv, flags):
    """B = A^-1 B from DecomposeLDL's factors and pivots (ldl.go:486-511);
    UPPER is tested first, neither bit leaves B as is."""
    n = A.shape[0]
    if n == 0 or B.shape[1] == 0:
        return
    piv = np.ascontiguousarray(ipiv, dtype=np.int64)
    Bm, _, _ = _view(B)
    Am, _, _ = _view(A)
    check("mops_dsytrs_ldl", _lib.lib().mops_dsytrs_ldl(
        ctypes.byref(Bm), ctypes.byref(Am), _pivot_array(piv), len(piv), flags, n,
        B.shape[1]))


def DecomposeBK(A, W, ipiv, flags, nb):
    """A = L D L^T (LOWER) or U D U^T (UPPER) with Bunch-Kaufman 1x1 / 2x2
    pivots, in place (ldlbk.go:1114-1140). W: the reference's workspace,
    checked as it does (N x 2 unblocked, N x (nb + 1) blocked; not read).
    ipiv: int64 array, LAPACK's format (1-based; negative for 2x2 blocks).
    Returns (A, err)."""
    p = _pivot_array(ipiv)
    ipiv[:] = 0
    n = A.shape[0]
    wr, wc = (0, 0) if W is None else W.shape
    need = 2 if (n < nb or nb == 0) else nb + 1
    if wc < need or wr < n:
        return None, MatopsError("Workspace too small")
    if n == 0 or not flags & (LOWER | UPPER):
        return A, None
    if A.shape[1] != n:
        raise ValueError("A must be square")
    if len(ipiv) < n:
        raise IndexError("index out of range")
    Am, _, _ = _view(A)
    check("mops_dsytrf_bk", _lib.lib().mops_dsytrf_bk(
        ctypes.byref(Am), p, len(ipiv), flags, n, nb))
    return A, None


def SolveBK(B, A, ipiv, flags):
    """B = A^-1 B from DecomposeBK's factors and pivots (ldlbks.go:303-316);
    LOWER is tested first. UPPER follows LAPACK's dsytrs (the reference's
    UPPER branch does not solve)."""
    n = A.shape[0]
    if n == 0 or B.shape[1] == 0:
        return
    piv = np.ascontiguousarray(ipiv, dtype=np.int64)
    Bm, _, _ = _view(B)
    Am, _, _ = _view(A)
    check("mops_dsytrs_bk", _lib.lib().mops_dsytrs_bk(
        ctypes.byref(Bm), ctypes.byref(Am), _pivot_array(piv), len(piv), flags, n,
        B.shape[1]))


def InverseTrm(A, flags, nb):
    """Inverse of the LOWER (else upper) triangle of A in place, UNIT for a
    unit diagonal (trinv.go:219-245). Returns (A, err)."""
    n = A.shape[0]
    if n == 0:
        return A, None
    Am, _, _ = _view(A)
    check("mops_dinverse_trm", _lib.lib().mops_dinverse_trm(ctypes.byref(Am), flags, n, nb))
    return A, None


def DecomposeLU(A, pivots, nb):
    """LU with partial pivoting in place (lu.go:257-273). pivots: int64
    numpy array (Go []int), at least min(rows, cols) long. Returns (A, err)."""
    mlen = min(A.shape)
    if len(pivots) < mlen:
        return A, MatopsError("pivot array < min(A.Rows(),A.Cols())")
    if pivots.dtype != np.int64 or not pivots.flags.c_contiguous:
        raise TypeError("pivots must be a contiguous int64 array")
    pivots[:] = 0
    if mlen == 0:
        return A, None
    Am, _, _ = _view(A)
    check("mops_dgetrf", _lib.lib().mops_dgetrf(
        ctypes.byref(Am), ctypes.c_void_p(pivots.ctypes.data), len(pivots),
        A.shape[0], A.shape[1], nb))
    return A, None


def SolveLU(B, A, pivots, flags):
    """B = A^-1 B (or A^-T B with TRANSA) from DecomposeLU (lu.go:320-334)."""
    n = A.shape[0]
    if n == 0 or B.shape[1] == 0:
        return None
    piv = np.ascontiguousarray(pivots, dtype=np.int64)
    Bm, _, _ = _view(B)
    Am, _, _ = _view(A)
    check("mops_dgetrs", _lib.lib().mops_dgetrs(
        ctypes.byref(Bm), ctypes.byref(Am), ctypes.c_void_p(piv.ctypes.data),
        len(piv), flags, n, B.shape[1]))
    return None
