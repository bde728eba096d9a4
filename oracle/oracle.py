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
            "o_DecomposeLDL": (I, [V, V, P, I, I, I, P]),
            "o_SolveLDL": (None, [V, V, P, I, I]),
            "o_DecomposeBK": (I, [V, V, P, I, I, I, P]),
            "o_SolveBK": (None, [V, V, P, I, I]),
            "o_iamax": (I, [P, I, I]),
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


def _w(W, n):
    return W if W is not None else np.zeros((max(n, 1), 2), order="F")


def DecomposeLDL(A, W, ipiv, flags, nb):
    """ldl.go:439-484. Returns None or the error text."""
    assert ipiv.dtype == np.int64 and ipiv.flags.c_contiguous
    msg = ctypes.c_char_p()
    rc = lib().o_DecomposeLDL(_v(A), _v(_w(W, A.shape[0])), ipiv.ctypes.data, len(ipiv), flags,
                              nb, ctypes.addressof(msg))
    return msg.value.decode() if rc else None


def SolveLDL(B, A, ipiv, flags):
    assert ipiv.dtype == np.int64
    lib().o_SolveLDL(_v(B), _v(A), ipiv.ctypes.data, len(ipiv), flags)


def DecomposeBK(A, W, ipiv, flags, nb):
    """ldlbk.go:1114-1140 (the blocked branch runs the unblocked algorithm).
    Returns None or the error text."""
    assert ipiv.dtype == np.int64 and ipiv.flags.c_contiguous
    msg = ctypes.c_char_p()
    rc = lib().o_DecomposeBK(_v(A), _v(W), ipiv.ctypes.data, len(ipiv), flags, nb,
                             ctypes.addressof(msg))
    return msg.value.decode() if rc else None


def SolveBK(B, A, ipiv, flags):
    """ldlbks.go:303-316, LOWER only (the reference's UPPER branch fails)."""
    assert ipiv.dtype == np.int64
    lib().o_SolveBK(_v(B), _v(A), ipiv.ctypes.data, len(ipiv), flags)


def iamax(x):
    """IAMax of the selected kernels (dvec_iamax) on a contiguous vector."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().o_iamax(x.ctypes.data, 1, len(x))


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


def _flush_cache():
    """Touches 256 MiB so the next timed run starts with cold caches (the
    reference's perf mains call mperf.FlushCache() before each run,
    test/perf_lu.go:188-201)."""
    buf = np.ones(32 << 20)
    buf *= 1.0001
    return float(buf[::4096].sum())


def host_cores():
    """Cores the host offers this process (affinity), and the CPU model."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return n, _cpu_model()


def _timed(fn, runs=1):
    """min over `runs` of fn()'s wall time (CLOCK_MONOTONIC), cache flushed
    before each run."""
    best = None
    for _ in range(runs):
        _flush_cache()
        t0 = lib().o_now()
        fn()
        t = lib().o_now() - t0
        best = t if best is None else min(best, t)
    return best


def bench_cholesky_sample(nb=256, n=8192, nrhs=64, seed=3):
    """Times DecomposeCHOL(A, LOWER, nb) + SolveCHOL (nrhs RHS) once, on one
    host core (the reference's Cholesky runs its kernels serially: only Mult
    splits over workers, and blockedCHOL calls SolveTrm / RankUpdateSym), on
    an n x n input of the GPU workload's kind (A = G G^T / n + I). Uses the
    reference's calgo kernels when they were built, else the restated ones.
    Returns the cpu_baseline record (TFLOP/s with the same LAPACK flop count
    as the GPU figure)."""
    kind = "reference" if reference_available() else "port"
    use(kind)
    NumWorkers(1)
    g = np.random.Generator(np.random.Philox(seed))
    G = g.standard_normal((n, n))
    A0 = np.asfortranarray(G @ G.T / n + np.eye(n))
    del G
    A0 = np.asfortranarray(np.tril(A0) + np.tril(A0, -1).T)
    B0 = np.asfortranarray(g.standard_normal((n, nrhs)))
    A = A0.copy(order="F")
    B = B0.copy(order="F")
    rcs = []

    def run():
        rcs.append(DecomposeCHOL(A, LOWER, nb)[0])
        SolveCHOL(B, A, LOWER)

    t = _timed(run)
    if rcs != [0]:
        raise RuntimeError("oracle Cholesky failed on the sample")
    flops = n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0 + 2.0 * n * n * nrhs
    ncores, model = host_cores()
    return {
        "value": round(flops / t / 1e12, 6),
        "unit": "TFLOP/s",
        "cores": 1,
        "kind": kind,
        "sample": f"DecomposeCHOL(A, LOWER, nb={nb}) + SolveCHOL, n={n}, {nrhs} RHS (the full "
                  f"workload), one run after a cache flush, {t:.2f} s on {model} "
                  f"({ncores} cores available)",
        "seconds": round(t, 3),
        "nproc": ncores,
        "cpu_model": model,
    }


def bench_sym_piv_sample(n=2048, nrhs=16, seed=9):
    """One-core times of DecomposeLDL(A, W, ipiv, LOWER, 0) + SolveLDL and of
    DecomposeBK(A, W, ipiv, LOWER, 0) + SolveBK (nrhs RHS) at order n, A =
    G + G^T with G of N(0,1) entries (its diagonal scaled by 0.05 for
    Bunch-Kaufman, which then takes 2x2 pivots), one run each after a cache
    flush. TFLOP/s with n^3/3 + 2 n^2 nrhs flops."""
    kind = "reference" if reference_available() else "port"
    use(kind)
    g = np.random.Generator(np.random.Philox(seed))
    G = g.standard_normal((n, n))
    S = G + G.T
    B0 = np.asfortranarray(g.standard_normal((n, nrhs)))
    flops = n ** 3 / 3.0 + 2.0 * n * n * nrhs
    out = {"n": n, "nrhs": nrhs, "kind": kind, "cores": 1}
    for name, scale, fac, sol in (("ldl", 1.0, DecomposeLDL, SolveLDL),
                                  ("bk", 0.05, DecomposeBK, SolveBK)):
        A = np.asfortranarray(S.copy())
        A[np.diag_indices(n)] *= scale
        X = B0.copy(order="F")
        ipiv = np.zeros(n, np.int64)
        W = np.zeros((n, 2), order="F")
        msg = []

        def run():
            msg.append(fac(A, W, ipiv, LOWER, 0))
            sol(X, A, ipiv, LOWER)

        t = _timed(run)
        if msg != [None]:
            raise RuntimeError(f"oracle {name} failed on the sample: {msg}")
        out[name] = {"seconds": round(t, 3), "tflops": round(flops / t / 1e12, 6)}
    return out


def bench_mult_and_lu(workers, n_mult=1024, n_lu=4096, nb=256, nrhs=16, mult_runs=5):
    """CPU legs of the other workloads with `workers` Mult workers
    (NumWorkers, mmult.go:69-73; the reference splits only Mult over them):
    DGEMM Mult(C, A, B, 1, 0) at n_mult^3 (min of mult_runs, cache flushed),
    and DecomposeLU(A, piv, nb) + SolveLU with nrhs RHS at n_lu (one run), A
    of N(0,1) entries. Returns {"dgemm": {...}, "lu": {...}} in TFLOP/s."""
    kind = "reference" if reference_available() else "port"
    use(kind)
    old = NumWorkers(workers)
    try:
        g = np.random.Generator(np.random.Philox(1))
        A = np.asfortranarray(g.uniform(-1, 1, (n_mult, n_mult)))
        B = np.asfortranarray(g.uniform(-1, 1, (n_mult, n_mult)))
        C = np.zeros((n_mult, n_mult), order="F")
        tm = _timed(lambda: Mult(C, A, B, 1.0, 0.0, 0), runs=mult_runs)
        g = np.random.Generator(np.random.Philox(5))
        L0 = np.asfortranarray(g.standard_normal((n_lu, n_lu)))
        X0 = np.asfortranarray(g.standard_normal((n_lu, nrhs)))
        F = L0.copy(order="F")
        X = X0.copy(order="F")
        piv = np.zeros(n_lu, dtype=np.int64)

        def lu():
            DecomposeLU(F, piv, nb)
            SolveLU(X, F, piv, 0)

        tl = _timed(lu)
    finally:
        NumWorkers(old)
    lu_flops = 2 * n_lu ** 3 / 3 - n_lu ** 2 / 2 + 5 * n_lu / 6 + (2 * n_lu ** 2 - n_lu) * nrhs
    return {
        "dgemm": {"n": n_mult, "seconds": round(tm, 4),
                  "tflops": round(2 * n_mult ** 3 / tm / 1e12, 6), "runs": mult_runs},
        "lu": {"n": n_lu, "nb": nb, "nrhs": nrhs, "seconds": round(tl, 3),
               "tflops": round(lu_flops / tl / 1e12, 6)},
        "workers": workers,
        "kind": kind,
    }
