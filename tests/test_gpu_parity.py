"""GPU parity of the BLAS-3 drop-ins and the Cholesky path against the
oracle: the golden fixtures (outputs of the oracle with the reference's own
calgo kernels, tests/golden/golden.npz), the oracle run live on the same
inputs (DGEMM 1024^3, DPOTRF), and the kernel-level entry points with
non-zero block indices, called from several host threads at once the way
Mult's goroutines call DMult (mmult.go:246-254).

Bounds: GEMM and SYRK componentwise, |C - C_ref| <= 2 k eps |op(A)||op(B)|
plus the beta term (both results within k eps of the exact product); the
triangular solves normwise, ||X - X_ref|| <= 40 n eps cond(op(A)) ||X_ref||;
Cholesky factors and solves within the same kind of bound with the
condition number of A.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import matops_amd as mo
from matops_amd import _lib
from helpers import EPS, embed, fmat, spd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import golden_inputs as gi  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = os.path.join(ROOT, "tests", "golden", "golden.npz")
ORACLE_KIND = "reference" if O.reference_available() else "port"


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _op(X, t):
    return X.T if t else X


# ----------------------------------------------------------- golden: GEMM


@pytest.mark.parametrize("flags", gi.GEMM_FLAGS)
def test_golden_gemm(gold, flags):
    A, B, C0 = gi.gemm_case(flags)
    assert gi.digest(np.concatenate([A.ravel("F"), B.ravel("F")])) == str(gold[f"gemm_{flags}_sha"])
    ta, tb = bool(flags & 1), bool(flags & 2)
    k = A.shape[0] if ta else A.shape[1]
    absprod = np.abs(_op(A, ta)) @ np.abs(_op(B, tb))
    for ab, (al, be) in enumerate(gi.GEMM_AB):
        C = C0.copy(order="F")
        assert mo.Mult(C, A, B, al, be, flags) is None
        ref = gold[f"gemm_{flags}_{ab}_C"]
        bound = 2 * k * EPS * abs(al) * absprod + 2 * EPS * np.abs(be * C0) + 1e-300
        assert np.all(np.abs(C - ref) <= bound), (flags, al, be)


# ----------------------------------------------------------- golden: TRSM


@pytest.mark.parametrize("flags", gi.TRSM_FLAGS)
def test_golden_trsm(gold, flags):
    T, B0 = gi.trsm_case(flags)
    B = B0.copy(order="F")
    assert mo.SolveTrm(B, T, gi.TRSM_ALPHA, flags) is None
    ref = gold[f"trsm_{flags}_X"]
    Te = T.copy()
    if flags & mo.UNIT:
        Te[np.diag_indices_from(Te)] = 1.0
    n = T.shape[0]
    cond = np.linalg.cond(Te, np.inf)
    err = np.linalg.norm(B - ref, np.inf) / np.linalg.norm(ref, np.inf)
    assert err <= 40 * n * EPS * cond, (flags, err, cond)


# ----------------------------------------------------------- golden: SYRK


@pytest.mark.parametrize("flags", gi.SYRK_FLAGS)
def test_golden_syrk(gold, flags):
    A, C0 = gi.syrk_case(flags)
    C = C0.copy(order="F")
    assert mo.RankUpdateSym(C, A, -1.0, 1.0, flags) is None
    ref = gold[f"syrk_{flags}_C"]
    ta = bool(flags & mo.TRANSA)
    k = A.shape[0] if ta else A.shape[1]
    absprod = np.abs(A.T) @ np.abs(A) if ta else np.abs(A) @ np.abs(A.T)
    bound = 2 * k * EPS * absprod + 2 * EPS * np.abs(C0)
    n = C0.shape[0]
    tri = np.tril_indices(n) if flags & mo.LOWER else np.triu_indices(n)
    other = np.triu_indices(n, 1) if flags & mo.LOWER else np.tril_indices(n, -1)
    assert np.all((np.abs(C - ref) <= bound)[tri])
    np.testing.assert_array_equal(C[other], ref[other])


# ------------------------------------------------------- golden: Cholesky


@pytest.mark.parametrize("n,nb", gi.CHOL_CASES)
@pytest.mark.parametrize("uplo", [mo.LOWER, mo.UPPER])
def test_golden_cholesky(gold, n, nb, uplo):
    S, B = gi.chol_case(n)
    key = f"chol_{n}_{nb}_{uplo}"
    assert gi.digest(S) == str(gold[key + "_Asha"])
    F = S.copy(order="F")
    _, err = mo.DecomposeCHOL(F, uplo, nb)
    assert err is None
    cond = np.linalg.cond(S)
    if key + "_F" in gold.files:
        Fg = gold[key + "_F"]
        tri = np.tril_indices(n) if uplo == mo.LOWER else np.triu_indices(n)
        d = np.abs(F - Fg)[tri].max() / np.abs(Fg[tri]).max()
        assert d <= 10 * n * EPS * cond
        other = np.triu_indices(n, 1) if uplo == mo.LOWER else np.tril_indices(n, -1)
        np.testing.assert_array_equal(F[other], S[other])  # untouched triangle
    X = B.copy(order="F")
    mo.SolveCHOL(X, F, uplo)
    Xg = gold[key + "_X"]
    assert np.abs(X - Xg).max() <= 20 * n * EPS * cond * np.abs(Xg).max()


# --------------------------------------------------------- live oracle


def test_dgemm_1024_against_oracle_mult():
    """C1: Mult(C, A, B, 1, 0, NOTRANS) at 1024^3, A and B U[-1,1], C filled
    with NaN first (beta = 0 clears it), against the oracle's Mult."""
    n = 1024
    A, B = fmat(n, n, 1), fmat(n, n, 2)
    C = np.full((n, n), np.nan, order="F")
    assert mo.Mult(C, A, B, 1.0, 0.0, mo.NOTRANS) is None
    O.use(ORACLE_KIND)
    Co = np.full((n, n), np.nan, order="F")
    assert O.Mult(Co, A, B, 1.0, 0.0, 0) == 0
    O.use("port")
    assert np.all(np.isfinite(C))
    assert np.all(np.abs(C - Co) <= 2 * n * EPS * (np.abs(A) @ np.abs(B)))


@pytest.mark.parametrize("uplo", [mo.LOWER, mo.UPPER])
def test_dpotrf_against_oracle(uplo):
    """DecomposeCHOL(A, uplo, 256) + SolveCHOL at n = 1500 against the
    oracle's DecomposeCHOL / SolveCHOL on the same input."""
    n, nb, nrhs = 1500, 256, 7
    A0 = spd(n, 17)
    B0 = fmat(n, nrhs, 18, "normal")
    F = A0.copy(order="F")
    _, err = mo.DecomposeCHOL(F, uplo, nb)
    assert err is None
    X = B0.copy(order="F")
    mo.SolveCHOL(X, F, uplo)
    O.use(ORACLE_KIND)
    Fo = A0.copy(order="F")
    assert O.DecomposeCHOL(Fo, uplo, nb)[0] == 0
    Xo = B0.copy(order="F")
    O.SolveCHOL(Xo, Fo, uplo)
    O.use("port")
    tri = np.tril_indices(n) if uplo == mo.LOWER else np.triu_indices(n)
    cond = np.linalg.cond(A0)
    assert np.abs(F - Fo)[tri].max() <= 10 * n * EPS * cond * np.abs(Fo[tri]).max()
    assert np.abs(X - Xo).max() <= 20 * n * EPS * cond * np.abs(Xo).max()


# ------------------------------------- kernel-level boundary, block indices


def _md(X):
    return _lib.MData(ctypes.c_void_p(X.ctypes.data), X.strides[1] // 8)


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_kernel_gemm_block_indices_against_oracle(flags):
    """mops_dmult_mm_blocked3 on C[R:E, S:L] against the oracle's kernel on
    the same block; everything outside the block untouched."""
    m, n, p = 300, 260, 190
    R, E, S, L = 13, 277, 9, 241
    ta, tb = bool(flags & 1), bool(flags & 2)
    A = fmat(*((p, m) if ta else (m, p)), seed=91)
    B = fmat(*((n, p) if tb else (p, n)), seed=92)
    C0 = fmat(m, n, seed=93)
    C = C0.copy(order="F")
    rc = _lib.lib().mops_dmult_mm_blocked3(ctypes.byref(_md(C)), ctypes.byref(_md(A)),
                                           ctypes.byref(_md(B)), -1.5, 0.5, flags, p, S, L, R,
                                           E, 196, 68, 68)
    assert rc == 0, _lib.last_error()
    O.use(ORACLE_KIND)
    Co = C0.copy(order="F")
    O.kern_gemm(Co, A, B, -1.5, 0.5, flags, p, S, L, R, E)
    O.use("port")
    opA, opB = _op(A, ta), _op(B, tb)
    bound = 3.0 * p * EPS * (np.abs(opA[R:E]) @ np.abs(opB[:, S:L])) \
        + 2 * EPS * np.abs(C0[R:E, S:L])
    assert np.all(np.abs(C[R:E, S:L] - Co[R:E, S:L]) <= bound)
    mask = np.ones((m, n), bool)
    mask[R:E, S:L] = False
    np.testing.assert_array_equal(C[mask], C0[mask])


@pytest.mark.parametrize("flags", [mo.LEFT | mo.LOWER, mo.LEFT | mo.UPPER | mo.TRANSA,
                                   mo.RIGHT | mo.LOWER | mo.TRANSA, mo.RIGHT | mo.UPPER,
                                   mo.LEFT | mo.LOWER | mo.UNIT])
@pytest.mark.parametrize("unb", [False, True])
def test_kernel_trsm_block_indices_against_oracle(flags, unb):
    """mops_dmmat_solve_blk / mops_dmmat_solve_unb on columns (LEFT) or rows
    (RIGHT) S:E of B against the oracle's kernel on the same block."""
    n, nrhs, S, E = 200, 90, 17, 71
    T = fmat(n, n, 101, "normal") * 0.3
    T[np.diag_indices(n)] = 2.0 + np.abs(fmat(n, 1, 102)[:, 0])
    T = np.asfortranarray(np.triu(T) if flags & mo.UPPER else np.tril(T))
    right = bool(flags & mo.RIGHT)
    B0 = fmat(*((nrhs, n) if right else (n, nrhs)), seed=103)
    B = B0.copy(order="F")
    Lb = _lib.lib()
    if unb:
        rc = Lb.mops_dmmat_solve_unb(ctypes.byref(_md(B)), ctypes.byref(_md(T)), 2.0, flags, n,
                                     S, E)
    else:
        rc = Lb.mops_dmmat_solve_blk(ctypes.byref(_md(B)), ctypes.byref(_md(T)), 2.0, flags, n,
                                     S, E, 68)
    assert rc == 0, _lib.last_error()
    O.use(ORACLE_KIND)
    Bo = B0.copy(order="F")
    O.kern_trsm(Bo, T, 2.0, flags, n, S, E, unblocked=unb)
    O.use("port")
    Te = T.copy()
    if flags & mo.UNIT:
        Te[np.diag_indices(n)] = 1.0
    cond = np.linalg.cond(Te, np.inf)
    blk = (slice(S, E), slice(None)) if right else (slice(None), slice(S, E))
    err = np.abs(B[blk] - Bo[blk]).max() / np.abs(Bo[blk]).max()
    assert err <= 40 * n * EPS * cond
    mask = np.ones(B.shape, bool)
    mask[blk] = False
    np.testing.assert_array_equal(B[mask], B0[mask])


@pytest.mark.parametrize("flags", [mo.LOWER, mo.UPPER, mo.LOWER | mo.TRANSA,
                                   mo.UPPER | mo.TRANSA])
def test_kernel_syrk_block_indices(flags):
    """mops_dmmat_rank_blk on the block S:E: UPPER against the oracle's
    kernel (calgo/dmmat_rank.c:62-126); LOWER with S > 0 updates the lower
    triangle of C[S:E, S:E] (the reference's rectangle there mixes two index
    bases, so it is compared with numpy on that block instead); S = 0 with
    the oracle for both."""
    n, k, S, E = 260, 70, 30, 230
    ta = bool(flags & mo.TRANSA)
    A = fmat(*((k, n) if ta else (n, k)), seed=121)
    R = A.T if ta else A
    C0 = fmat(n, n, seed=122)
    Lb = _lib.lib()
    for s0 in (0, S):
        C = C0.copy(order="F")
        rc = Lb.mops_dmmat_rank_blk(ctypes.byref(_md(C)), ctypes.byref(_md(A)), -1.0, 1.0, flags,
                                    k, s0, E, 196, 68)
        assert rc == 0, _lib.last_error()
        blk = C0[s0:E, s0:E] - R[s0:E] @ R[s0:E].T
        tri = np.tril_indices(E - s0) if flags & mo.LOWER else np.triu_indices(E - s0)
        absb = np.abs(R[s0:E]) @ np.abs(R[s0:E]).T + np.abs(C0[s0:E, s0:E])
        assert np.all(np.abs(C[s0:E, s0:E] - blk)[tri] <= 2 * k * EPS * absb[tri])
        mask = np.ones((n, n), bool)
        mask[s0:E, s0:E][tri] = False
        np.testing.assert_array_equal(C[mask], C0[mask])
        if s0 == 0 or not flags & mo.LOWER:
            O.use(ORACLE_KIND)
            Co = C0.copy(order="F")
            O.kern_syrk(Co, A, -1.0, 1.0, flags, k, s0, E)
            O.use("port")
            assert np.all(np.abs(C - Co)[s0:E, s0:E][tri] <= 3 * k * EPS * absb[tri])


# --------------------------------------------- concurrent host threads


@pytest.mark.parametrize("workers", [2, 4, 6])
@pytest.mark.parametrize("flags", [0, 3])
def test_mult_numworkers_threads(workers, flags):
    """Mult with NumWorkers(W) calls mops_dmult_mm_blocked3 from W host
    threads at once on the divideWork/blockIndex4 blocks (mmult.go:137-188,
    246-254). The result equals, bit for bit, the same block calls made one
    after the other on one thread, and agrees with the oracle's Mult at
    NumWorkers(W) and with numpy within the bound."""
    m, n, p = 700, 520, 300
    ta, tb = bool(flags & 1), bool(flags & 2)
    A = fmat(*((p, m) if ta else (m, p)), seed=131)
    B = fmat(*((n, p) if tb else (p, n)), seed=132)
    C0 = fmat(m, n, seed=133)
    old = mo.NumWorkers(workers)
    try:
        C = C0.copy(order="F")
        assert mo.Mult(C, A, B, -1.0, 0.5, flags) is None
    finally:
        mo.NumWorkers(old)
    # the same blocks, sequentially
    Cs = C0.copy(order="F")
    cw, rw = mo.divideWork(m, n, workers)
    Lb = _lib.lib()
    for kk in range(cw):
        cs, ce = mo.blockIndex4(kk, cw, n), mo.blockIndex4(kk + 1, cw, n)
        for ll in range(rw):
            rs, re_ = mo.blockIndex4(ll, rw, m), mo.blockIndex4(ll + 1, rw, m)
            assert Lb.mops_dmult_mm_blocked3(ctypes.byref(_md(Cs)), ctypes.byref(_md(A)),
                                             ctypes.byref(_md(B)), -1.0, 0.5, flags, p, cs, ce,
                                             rs, re_, 196, 68, 68) == 0
    np.testing.assert_array_equal(C, Cs)
    O.use(ORACLE_KIND)
    oldo = O.NumWorkers(workers)
    Co = C0.copy(order="F")
    O.Mult(Co, A, B, -1.0, 0.5, flags)
    O.NumWorkers(oldo)
    O.use("port")
    opA, opB = _op(A, ta), _op(B, tb)
    bound = 2 * p * EPS * (np.abs(opA) @ np.abs(opB)) + 2 * EPS * np.abs(0.5 * C0)
    assert np.all(np.abs(C - Co) <= bound)
    assert np.all(np.abs(C - (0.5 * C0 - opA @ opB)) <= bound)


def test_threads_on_views_disjoint_blocks():
    """Concurrent calls on disjoint blocks of one parent with ld > rows leave
    everything outside the view untouched."""
    m, n, p = 520, 480, 260
    A, B = fmat(m, p, 141), fmat(p, n, 142)
    Cp, Cv = embed(np.full((m, n), np.nan), 11, 5, 3)
    old = mo.NumWorkers(4)
    try:
        mo.Mult(Cv, A, B, 1.0, 0.0, mo.NOTRANS)
    finally:
        mo.NumWorkers(old)
    assert np.all(np.abs(Cv - A @ B) <= 2 * p * EPS * (np.abs(A) @ np.abs(B)))
    mask = np.ones(Cp.shape, bool)
    mask[5:5 + m, 3:3 + n] = False
    assert np.all(np.isnan(Cp[mask]))
