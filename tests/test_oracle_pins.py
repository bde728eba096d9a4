"""CPU pins of oracle functions that the golden fixtures alone do not pin:
SolveLU (both branches), RankUpdateSym (all four flag sets, and the
reference's handling of beta), and the kernel-level boundary with non-zero
block indices (dmult_mm_blocked3 [R:E, S:L], dmmat_solve_blk /
dmmat_solve_unb [S:E], dmmat_rank_blk [S:E]).

Every check compares with an independent computation: LAPACK through scipy
(lu_factor / lu_solve, solve_triangular), numpy products, or the closed
form of what a block index selects. Each runs on the restated kernels
("port") and on the reference's calgo C built from source ("reference")
when that was built.
"""
import os
import sys

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import oracle as O
from helpers import EPS, fmat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

KINDS = ["port"] + (["reference"] if O.reference_available() else [])
LOWER, UPPER, TRANSA, TRANSB = O.LOWER, O.UPPER, O.TRANSA, O.TRANSB
LEFT, RIGHT, UNIT = O.LEFT, O.RIGHT, O.UNIT


@pytest.fixture(params=KINDS)
def kind(request):
    O.use(request.param)
    yield request.param
    O.use("port")


def _perm(piv, m):
    """Row permutation applied by the reference's forward pivots."""
    p = np.arange(m)
    for k, q in enumerate(piv):
        p[[k, q]] = p[[q, k]]
    return p


# ------------------------------------------------------------ SolveLU


@pytest.mark.parametrize("n,nb,nrhs", [(7, 0, 3), (64, 12, 5), (150, 32, 9), (300, 68, 4)])
def test_solvelu_matches_lapack_lu_solve(kind, n, nb, nrhs):
    """SolveLU (lu.go:320-334, non-TRANSA branch) with the oracle's own
    factors solves A X = B as LAPACK's getrs does with LAPACK's factors."""
    A = fmat(n, n, 500 + n, "normal")
    B = fmat(n, nrhs, 600 + n, "normal")
    F = A.copy(order="F")
    piv = np.zeros(n, dtype=np.int64)
    O.DecomposeLU(F, piv, nb)
    X = B.copy(order="F")
    O.SolveLU(X, F, piv, 0)
    lu, lpiv = sla.lu_factor(A)
    assert np.array_equal(piv, lpiv)  # the same pivots as LAPACK (ipiv - 1)
    Xl = sla.lu_solve((lu, lpiv), B)
    cond = np.linalg.cond(A, 1)
    assert np.max(np.abs(X - Xl)) <= 50 * n * EPS * cond * np.max(np.abs(Xl))
    eta = np.linalg.norm(A @ X - B, 1) / (np.linalg.norm(A, 1) * np.linalg.norm(X, 1)
                                          + np.linalg.norm(B, 1))
    assert eta <= 10 * n * EPS


def test_solvelu_transa_branch_is_the_documented_order(kind):
    """The TRANSA branch of SolveLU applies the pivots first, then L^T and
    U^T (lu.go:322-326): X = U^-T L^-T P B. Pinned against that order
    computed with scipy's triangular solves from the factors, and shown not
    to solve A^T X = B for a matrix whose pivots move rows."""
    n, nrhs = 40, 3
    A = fmat(n, n, 77, "normal")
    B = fmat(n, nrhs, 78, "normal")
    F = A.copy(order="F")
    piv = np.zeros(n, dtype=np.int64)
    O.DecomposeLU(F, piv, 0)
    assert np.any(piv != np.arange(n))
    X = B.copy(order="F")
    O.SolveLU(X, F, piv, TRANSA)
    L = np.tril(F, -1) + np.eye(n)
    U = np.triu(F)
    PB = B[_perm(piv, n)]
    Y = sla.solve_triangular(L, PB, trans="T", lower=True, unit_diagonal=True)
    Xd = sla.solve_triangular(U, Y, trans="T", lower=False)
    np.testing.assert_allclose(X, Xd, rtol=1e-10, atol=1e-12)
    Xt = np.linalg.solve(A.T, B)
    assert np.max(np.abs(X - Xt)) > 1e-3 * np.max(np.abs(Xt))


# ------------------------------------------------------ RankUpdateSym


@pytest.mark.parametrize("flags", [LOWER, UPPER, LOWER | TRANSA, UPPER | TRANSA])
@pytest.mark.parametrize("n,k", [(1, 1), (60, 40), (150, 70)])
def test_rankupdatesym_all_flag_sets(kind, flags, n, k):
    """RankUpdateSym (mmult.go:426-445) with beta = 1: the requested
    triangle of C - A A^T (or C - A^T A) within k eps |A||A|^T, the other
    strict triangle untouched."""
    ta = bool(flags & TRANSA)
    A = fmat(*((k, n) if ta else (n, k)), seed=70 + n)
    C0 = fmat(n, n, seed=71 + n)
    C = C0.copy(order="F")
    assert O.RankUpdateSym(C, A, -1.0, 1.0, flags) == 0
    AA = A.T @ A if ta else A @ A.T
    absAA = np.abs(A.T) @ np.abs(A) if ta else np.abs(A) @ np.abs(A.T)
    ref = C0 - AA
    tri = np.tril_indices(n) if flags & LOWER else np.triu_indices(n)
    other = np.triu_indices(n, 1) if flags & LOWER else np.tril_indices(n, -1)
    assert np.all(np.abs(C - ref)[tri] <= (k + 2) * EPS * (absAA + np.abs(C0))[tri])
    np.testing.assert_array_equal(C[other], C0[other])


@pytest.mark.parametrize("flags", [LOWER, UPPER])
def test_rankupdatesym_beta_scales_only_diagonal_blocks(kind, flags):
    """The reference applies beta only to its NB x NB diagonal blocks
    (calgo/dmmat_rank.c:99-105, 117-123; NB = 68 from mmult.go:34):
    pinned with beta = 0.5 at n = 150 (blocks [0,68), [68,136), [136,150))."""
    n, k, nb = 150, 30, 68
    A = fmat(n, k, seed=81)
    C0 = fmat(n, n, seed=82)
    C = C0.copy(order="F")
    O.RankUpdateSym(C, A, 1.0, 0.5, flags)
    AA = A @ A.T
    blk = np.arange(n) // nb
    diag_blk = blk[:, None] == blk[None, :]
    ref = np.where(diag_blk, 0.5 * C0, C0) + AA
    tri = np.tril(np.ones((n, n), bool)) if flags & LOWER else np.triu(np.ones((n, n), bool))
    err = np.abs(C - ref)[tri]
    assert np.all(err <= (k + 2) * EPS * (np.abs(A) @ np.abs(A.T) + np.abs(C0))[tri])


# ------------------------------------- kernel-level boundary, block indices


@pytest.mark.parametrize("flags", [0, TRANSA, TRANSB, TRANSA | TRANSB])
def test_kernel_gemm_block_indices(kind, flags):
    """dmult_mm_blocked3 updates exactly C[R:E, S:L] with rows R:E of op(A)
    and columns S:L of op(B) (calgo/dmmat_gemm.c:364-416)."""
    m, n, p = 90, 70, 50
    R, E, S, L = 13, 77, 9, 61
    ta, tb = bool(flags & TRANSA), bool(flags & TRANSB)
    A = fmat(*((p, m) if ta else (m, p)), seed=91)
    B = fmat(*((n, p) if tb else (p, n)), seed=92)
    C0 = fmat(m, n, seed=93)
    C = C0.copy(order="F")
    O.kern_gemm(C, A, B, -1.5, 0.5, flags, p, S, L, R, E)
    opA = A.T if ta else A
    opB = B.T if tb else B
    ref = C0.copy()
    ref[R:E, S:L] = 0.5 * C0[R:E, S:L] - 1.5 * opA[R:E] @ opB[:, S:L]
    bound = 1.5 * p * EPS * (np.abs(opA[R:E]) @ np.abs(opB[:, S:L])) + EPS * np.abs(C0[R:E, S:L])
    assert np.all(np.abs(C[R:E, S:L] - ref[R:E, S:L]) <= bound + 1e-300)
    mask = np.ones((m, n), bool)
    mask[R:E, S:L] = False
    np.testing.assert_array_equal(C[mask], C0[mask])


def _tri(n, upper, seed):
    T = fmat(n, n, seed, "normal") * 0.3
    T[np.diag_indices(n)] = 2.0 + np.abs(fmat(n, 1, seed + 1)[:, 0])
    return np.asfortranarray(np.triu(T) if upper else np.tril(T))


@pytest.mark.parametrize("side", [LEFT, RIGHT])
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("unb", [False, True])
def test_kernel_trsm_block_indices(kind, side, uplo, unb):
    """dmmat_solve_blk / dmmat_solve_unb solve only columns (LEFT) or rows
    (RIGHT) S:E of B (calgo/dmmat_solve.c:405, 425, 864-909)."""
    n, nrhs, S, E = 50, 30, 7, 23
    T = _tri(n, uplo == UPPER, 101)
    right = side == RIGHT
    B0 = fmat(*((nrhs, n) if right else (n, nrhs)), seed=102)
    B = B0.copy(order="F")
    O.kern_trsm(B, T, 2.0, side | uplo, n, S, E, unblocked=unb)
    ref = B0.copy()
    if right:
        ref[S:E] = 2.0 * sla.solve(T.T, B0[S:E].T).T
    else:
        ref[:, S:E] = 2.0 * sla.solve_triangular(T, B0[:, S:E], lower=uplo == LOWER)
    np.testing.assert_allclose(B, ref, rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("unb", [False, True])
def test_kernel_trsm_triangle_choice_without_uplo_bits(kind, unb):
    """With neither LOWER nor UPPER set, dmmat_solve_unb solves with the
    upper triangle (it tests LOWER first, calgo/dmmat_solve.c:423).
    dmmat_solve_blk tests UPPER first (:879) and so walks the lower
    blocked variant, but that variant solves its diagonal blocks with
    dmmat_solve_unb and the same flags (:587), i.e. with their upper
    triangles: for N <= NB = 68 (one block) the reference's blocked solve is
    the upper solve. The restatement ("port") takes the lower triangle
    throughout instead, as the GPU path does; both are pinned here."""
    n, nrhs = 20, 4
    G = fmat(n, n, 111, "normal") * 0.3
    G[np.diag_indices(n)] = 3.0
    G = np.asfortranarray(G)
    B0 = fmat(n, nrhs, 112)
    B = B0.copy(order="F")
    O.kern_trsm(B, G, 1.0, LEFT, n, 0, nrhs, unblocked=unb)
    upper = unb or kind == "reference"
    tri = np.triu(G) if upper else np.tril(G)
    np.testing.assert_allclose(B, np.linalg.solve(tri, B0), rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("ta", [False, True])
def test_kernel_syrk_block_indices(kind, ta):
    """dmmat_rank_blk with S > 0 (calgo/dmmat_rank.c:62-126). UPPER updates
    the upper triangle of C[S:E, S:E] from rows (columns with TRANSA) S:E of
    A. LOWER updates, for each NB-row block i of [S, E), the diagonal
    triangle from rows i.. of A and the rectangle C[i:i+nI, 0:i] from rows
    S..S+i of A (:116-123): two index bases that agree only for S = 0.
    Both pinned against numpy with NB = 68 (a single block for E - S <= 68)."""
    n, k, S, E = 120, 25, 10, 60
    A = fmat(*((k, n) if ta else (n, k)), seed=121)
    R = A.T if ta else A  # rows of R are the vectors
    flags_t = TRANSA if ta else 0
    C0 = fmat(n, n, seed=122)
    # UPPER: the triangle of the block
    C = C0.copy(order="F")
    O.kern_syrk(C, A, 1.0, 1.0, UPPER | flags_t, k, S, E)
    ref = C0.copy()
    iu = np.triu_indices(E - S)
    blk = C0[S:E, S:E] + R[S:E] @ R[S:E].T
    ref[S:E, S:E][iu] = blk[iu]
    np.testing.assert_allclose(C, ref, rtol=0, atol=(k + 2) * EPS * 4 * k)
    # LOWER: diagonal triangle + the reference's rectangle left of it
    C = C0.copy(order="F")
    O.kern_syrk(C, A, 1.0, 1.0, LOWER | flags_t, k, S, E)
    ref = C0.copy()
    il = np.tril_indices(E - S)
    blk = C0[S:E, S:E] + R[S:E] @ R[S:E].T
    ref[S:E, S:E][il] = blk[il]
    ref[S:E, 0:S] += R[S:E] @ R[S:2 * S].T
    np.testing.assert_allclose(C, ref, rtol=0, atol=(k + 2) * EPS * 4 * k)
