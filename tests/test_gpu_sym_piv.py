"""GPU parity of DecomposeLDL / SolveLDL (1x1 diagonal pivots, ldl.go) and
DecomposeBK / SolveBK (Bunch-Kaufman, ldlbk.go, ldlbks.go) through the C ABI.

Checks: the pivot vectors equal the oracle's (and LAPACK dsytrf's, for
Bunch-Kaufman) exactly; the factors agree with the oracle's to rounding
(c n eps max|A| times the growth of the factors); the backward error of
P A P^T = T D T^T from the GPU's own factors, a few n eps;
exact results on inputs with exact arithmetic (the diagonal case, a zero
pivot in the blocked LOWER form's unblocked tail, with the state the
reference leaves); the stored triangle only is written; the solves against
LAPACK dsytrs and numpy; the device twins equal the host entry points
bit for bit. The oracle's blocked LOWER LDL runs with a workspace of
nb + 1 columns (with nb it is defective, as the reference's is), and on the
restated kernels for an odd order (the reference's GEMV writes out of bounds
there: calgo/dmvec_gemv.c:241 offsets Y by X's increment).
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

import matops_amd as mo
from matops_amd import _lib
from helpers import EPS
from oracle import oracle as O

pytestmark = pytest.mark.gpu
ORACLE_KIND = "reference" if O.reference_available() else "port"


def _oracle(fn, *args):
    O.use(ORACLE_KIND)
    try:
        return fn(*args)
    finally:
        O.use("port")


def _sym(n, seed, diag_scale=1.0):
    g = np.random.default_rng(seed)
    G = g.standard_normal((n, n))
    A = G + G.T
    A[np.diag_indices(n)] *= diag_scale
    return np.asfortranarray(A)


def _with_other(A0, uplo, v=55.0):
    A = A0.copy(order="F")
    n = A.shape[0]
    other = np.triu_indices(n, 1) if uplo == mo.LOWER else np.tril_indices(n, -1)
    A[other] = v
    return A, other


def _ldl_perm(ipiv, n, lower):
    P = np.eye(n)
    for k in (range(n) if lower else range(n - 1, -1, -1)):
        if ipiv[k] > 0:
            j = ipiv[k] - 1
            P[[k, j]] = P[[j, k]]
    return P


def _ldl_backward(A0, A, ipiv, lower):
    """max |P A0 P^T - T D T^T| / max(|T| |D| |T|^T): the backward error of
    the factors, relative to the bound of an LDL^T factorization (a few n
    eps whatever the growth of the factors)."""
    n = A0.shape[0]
    T = (np.tril(A, -1) if lower else np.triu(A, 1)) + np.eye(n)
    P = _ldl_perm(ipiv, n, lower)
    d = np.diag(A)
    R = P @ A0 @ P.T - (T * d) @ T.T
    return np.abs(R).max() / ((np.abs(T) * np.abs(d)) @ np.abs(T).T).max()


def _growth(A0, A, lower):
    """max(|T| |D| |T|^T) / max |A0| of the stored factors."""
    n = A0.shape[0]
    T = (np.tril(A, -1) if lower else np.triu(A, 1)) + np.eye(n)
    return ((np.abs(T) * np.abs(np.diag(A))) @ np.abs(T).T).max() / np.abs(A0).max()


def _eta(A0, X, B0):
    R = A0 @ X - B0
    return np.abs(R).max() / (np.abs(A0).sum(axis=1).max() * np.abs(X).max() + np.abs(B0).max())


def _oracle_ldl(A0, uplo, nb, other_val=55.0):
    n = A0.shape[0]
    Ao, _ = _with_other(A0, uplo, other_val)
    ipo = np.zeros(n, np.int64)
    W = np.zeros((n, max(nb + 1, 2)), order="F")
    if uplo == mo.LOWER and 0 < nb <= n and n % 2:
        # the reference's GEMV (calgo/dmvec_gemv.c:241) offsets Y by R times
        # X's increment on its unaligned path (odd leading dimension) once
        # it splits more than 68 rows into blocks, and the blocked LOWER
        # LDL's panel GEMV has a strided X: the reference kernels write out
        # of bounds there, so the restated ones are the oracle
        O.use("port")
        return Ao, ipo, O.DecomposeLDL(Ao, W, ipo, uplo, nb)
    msg = _oracle(O.DecomposeLDL, Ao, W, ipo, uplo, nb)
    return Ao, ipo, msg


# ------------------------------------------------------------------- LDL
@pytest.mark.parametrize("uplo", [mo.LOWER, mo.UPPER])
@pytest.mark.parametrize("n,nb", [(1, 0), (2, 0), (7, 0), (50, 0), (131, 0), (64, 16),
                                  (100, 32), (300, 68), (257, 64), (1000, 0), (1000, 96)])
def test_ldl_against_oracle(uplo, n, nb):
    A0 = _sym(n, 10 + n + nb)
    A, other = _with_other(A0, uplo)
    ipiv = np.zeros(n, np.int64)
    R, err = mo.DecomposeLDL(A, np.zeros((n, max(nb, 1)), order="F"), ipiv, uplo, nb)
    assert err is None and R is A
    np.testing.assert_array_equal(A[other], 55.0)  # the stored triangle only
    Ao, ipo, msg = _oracle_ldl(A0, uplo, nb)
    assert msg is None
    np.testing.assert_array_equal(ipiv, ipo)
    # rounding differences grow with the factors' growth (block-local
    # pivoting lets |T| |D| |T|^T reach 1e4 max|A| at n = 1000; the
    # oracle's two kernel sets then differ by 1e-5 too); the backward error
    # is the strict check
    assert np.abs(A - Ao).max() <= 100 * n * EPS * np.abs(Ao).max() * max(
        1.0, _growth(A0, A, uplo == mo.LOWER))
    assert _ldl_backward(A0, A, ipiv, uplo == mo.LOWER) <= 10 * n * EPS


@pytest.mark.parametrize("uplo", [mo.LOWER, mo.UPPER])
@pytest.mark.parametrize("nb", [0, 4])
def test_ldl_diagonal_matrix_sorts_by_magnitude(uplo, nb):
    d = np.array([3.0, -7.0, 1.0, 7.0, -2.0, 0.5])
    A = np.asfortranarray(np.diag(d))
    ipiv = np.zeros(6, np.int64)
    _, err = mo.DecomposeLDL(A, None, ipiv, uplo, nb)
    assert err is None
    D = np.diag(A)
    Ao, ipo, _ = _oracle_ldl(np.diag(d), uplo, nb, 0.0)
    np.testing.assert_array_equal(A, Ao)
    np.testing.assert_array_equal(ipiv, ipo)
    if uplo == mo.LOWER:
        np.testing.assert_array_equal(D, [-7.0, 7.0, 3.0, -2.0, 1.0, 0.5])
    elif nb == 0:
        np.testing.assert_array_equal(D, [0.5, 1.0, -2.0, 3.0, 7.0, -7.0])


@pytest.mark.parametrize("uplo,msg", [(mo.LOWER, "zero value on diagonal"),
                                      (mo.UPPER, "zero on diagonal.")])
def test_ldl_zero_pivot_error(uplo, msg):
    A = np.zeros((3, 3), order="F")
    ipiv = np.zeros(3, np.int64)
    R, err = mo.DecomposeLDL(A, None, ipiv, uplo, 0)
    assert R is A and str(err) == msg
    _, err = mo.DecomposeLDL(np.zeros((3, 4), order="F"), None, ipiv, uplo, 0)
    assert str(err) == "A not a square matrix"
    rc = _lib.lib().mops_dsytrf_ldl(ctypes.byref(_lib.MData(
        ctypes.c_void_p(A.ctypes.data), 3)), ctypes.c_void_p(ipiv.ctypes.data), 3, uplo, 3, 0)
    assert rc == (1 if uplo == mo.LOWER else 3)  # 1 + the index of the zero pivot
    mo.SetPanicOnError(True)
    try:
        with pytest.raises(RuntimeError, match=msg.rstrip(".")):
            mo.DecomposeLDL(np.zeros((3, 3), order="F"), None, ipiv, uplo, 0)
    finally:
        mo.SetPanicOnError(False)


def _tail_zero_matrix():
    # exact arithmetic: the panel (columns 0..3, nb = 4) takes the diagonal
    # 8, 4, 2, 1.5 in order; the tail's updated diagonal is [0, 1, 0, 0], so
    # its first step interchanges rows 4 and 5 and its second meets a zero
    A = np.diag([8.0, 4.0, 2.0, 1.5, 0.5, 1.25, 0.5, 0.0])
    A[4, 0] = A[0, 4] = 2.0
    A[5, 1] = A[1, 5] = 1.0
    A[6, 2] = A[2, 6] = 1.0
    return np.asfortranarray(A)


@pytest.mark.parametrize("nb", [0, 4])
def test_ldl_zero_pivot_state_matches_reference(nb):
    A0 = _tail_zero_matrix()
    A = A0.copy(order="F")
    ipiv = np.zeros(8, np.int64)
    _, err = mo.DecomposeLDL(A, None, ipiv, mo.LOWER, nb)
    assert str(err) == "zero value on diagonal"
    Ao, ipo, msg = _oracle_ldl(A0, mo.LOWER, nb, 0.0)
    assert msg == "zero value on diagonal"
    # blocked: the tail's pivots stay relative and the rows of the panel
    # columns are not interchanged, as the reference returns
    np.testing.assert_array_equal(ipiv, ipo)
    np.testing.assert_array_equal(np.tril(A), np.tril(Ao))
    if nb:
        assert ipiv[4] == 2 and A[4, 1] == 0.0


def test_ldl_upper_blocked_zero_pivot_in_second_block():
    n, nb = 12, 4
    A0 = _sym(n, 3)
    A0[:8, :8] = 0.0  # the top two blocks are zero
    A0[8:, :8] = 0.0
    A0[:8, 8:] = 0.0
    A = A0.copy(order="F")
    ipiv = np.zeros(n, np.int64)
    _, err = mo.DecomposeLDL(A, None, ipiv, mo.UPPER, nb)
    assert str(err) == "zero on diagonal."
    Ao, ipo, msg = _oracle_ldl(A0, mo.UPPER, nb, 0.0)
    assert msg == "zero on diagonal."
    np.testing.assert_array_equal(ipiv, ipo)
    assert np.abs(np.triu(A) - np.triu(Ao)).max() <= 100 * n * EPS * np.abs(Ao).max()


@pytest.mark.parametrize("uplo,nb", [(mo.LOWER, 0), (mo.UPPER, 0), (mo.UPPER, 16),
                                     (mo.LOWER, 64)])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (90, 5), (700, 3), (1100, 40)])
def test_solve_ldl(uplo, nb, n, nrhs):
    A0 = _sym(n, 7 + n)
    A = A0.copy(order="F")
    ipiv = np.zeros(n, np.int64)
    _, err = mo.DecomposeLDL(A, None, ipiv, uplo, nb)
    assert err is None
    B0 = np.asfortranarray(np.random.default_rng(8).standard_normal((n, nrhs)))
    X = B0.copy(order="F")
    mo.SolveLDL(X, A, ipiv, uplo)
    cond = np.linalg.cond(A0)
    ref = np.linalg.solve(A0, B0)
    assert np.abs(X - ref).max() <= 100 * n * EPS * cond * np.abs(ref).max()
    if n <= 700:
        Xo = B0.copy(order="F")
        _oracle(O.SolveLDL, Xo, A, ipiv, uplo)
        assert np.abs(X - Xo).max() <= 100 * n * EPS * cond * np.abs(ref).max()


def test_solve_ldl_no_triangle_bit_leaves_b():
    A = _sym(5, 1)
    ipiv = np.zeros(5, np.int64)
    B = np.asfortranarray(np.arange(10.0).reshape(5, 2))
    mo.SolveLDL(B, A, ipiv, 0)
    np.testing.assert_array_equal(B, np.arange(10.0).reshape(5, 2))


# -------------------------------------------------------------------- BK
@pytest.mark.parametrize("uplo", [mo.LOWER, mo.UPPER])
@pytest.mark.parametrize("n,scale,nb", [(1, 1.0, 0), (2, 0.01, 0), (7, 0.05, 0), (50, 0.05, 0),
                                        (200, 0.05, 0), (123, 1.0, 0), (513, 0.05, 64),
                                        (1000, 0.05, 0), (1000, 1.0, 128)])
def test_bk_against_lapack_and_oracle(uplo, n, scale, nb):
    A0 = _sym(n, 40 + n, scale)  # a small diagonal brings 2x2 pivots
    A, other = _with_other(A0, uplo)
    ipiv = np.zeros(n, np.int64)
    R, err = mo.DecomposeBK(A, np.zeros((n, nb + 1 if nb else 2), order="F"), ipiv, uplo, nb)
    assert err is None and R is A
    np.testing.assert_array_equal(A[other], 55.0)
    ldu, lpiv, info = lapack.dsytrf(A0, lower=1 if uplo == mo.LOWER else 0)
    assert info == 0
    np.testing.assert_array_equal(ipiv, lpiv.astype(np.int64))
    tri = np.tril_indices(n) if uplo == mo.LOWER else np.triu_indices(n)
    big = max(1.0, np.abs(ldu[tri]).max())
    assert np.abs(A[tri] - ldu[tri]).max() <= 100 * n * EPS * big
    if n <= 513:
        Ao, _ = _with_other(A0, uplo)
        ipo = np.zeros(n, np.int64)
        assert _oracle(O.DecomposeBK, Ao, np.zeros((n, 5), order="F"), ipo, uplo, 0) is None
        np.testing.assert_array_equal(ipiv, ipo)
        assert np.abs(A[tri] - Ao[tri]).max() <= 100 * n * EPS * big
    if n >= 50 and scale < 1.0:
        assert (ipiv < 0).any()  # 2x2 blocks were exercised


def test_bk_reference_table_matrix():
    # the 7 x 7 lower table of ldlbkl_test.go:62-69 (a 2x2 pivot at the
    # second column)
    L = np.array([[1, 0, 0, 0, 0, 0, 0], [1, 2, 0, 0, 0, 0, 0], [1, 2, 3, 0, 0, 0, 0],
                  [1, 2, 3, 4, 0, 0, 0], [1, 5, 3, 4, 5, 0, 0], [1, 2, 3, 4, 9, 6, 0],
                  [1, 2, 3, 4, 5, 6, 7]], dtype=np.float64)
    A0 = np.tril(L) + np.tril(L, -1).T
    A = np.asfortranarray(A0.copy())
    ipiv = np.zeros(7, np.int64)
    _, err = mo.DecomposeBK(A, np.zeros((7, 2), order="F"), ipiv, mo.LOWER, 0)
    assert err is None
    Ao = np.asfortranarray(A0.copy())
    ipo = np.zeros(7, np.int64)
    _oracle(O.DecomposeBK, Ao, np.zeros((7, 5), order="F"), ipo, mo.LOWER, 0)
    np.testing.assert_array_equal(ipiv, ipo)
    assert (ipiv < 0).any()
    tri = np.tril_indices(7)
    assert np.abs(A[tri] - Ao[tri]).max() <= 1e-13 * np.abs(Ao).max()


def test_bk_self_pivot_format_and_flags():
    A0 = np.diag([4.0, 3.0, 2.0, 1.0])
    for uplo in (mo.LOWER, mo.UPPER):
        A = np.asfortranarray(A0.copy())
        ipiv = np.zeros(4, np.int64)
        mo.DecomposeBK(A, np.zeros((4, 2), order="F"), ipiv, uplo, 0)
        np.testing.assert_array_equal(ipiv, [1, 2, 3, 4])
        np.testing.assert_array_equal(A, A0)
    A = np.asfortranarray(_sym(6, 2))
    A1 = A.copy(order="F")
    ipiv = np.full(6, 9, np.int64)
    _, err = mo.DecomposeBK(A, np.zeros((6, 2), order="F"), ipiv, 0, 0)  # neither triangle
    assert err is None
    np.testing.assert_array_equal(ipiv, 0)
    np.testing.assert_array_equal(A, A1)


def test_bk_workspace_checks():
    A = np.asfortranarray(_sym(10, 1))
    ip = np.zeros(10, np.int64)
    for W, nb in ((np.zeros((10, 1), order="F"), 0), (np.zeros((9, 2), order="F"), 0),
                  (np.zeros((10, 4), order="F"), 4), (None, 0)):
        R, err = mo.DecomposeBK(A.copy(order="F"), W, ip, mo.LOWER, nb)
        assert R is None and str(err) == "Workspace too small"


@pytest.mark.parametrize("uplo", [mo.LOWER, mo.UPPER])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (2, 3), (7, 2), (60, 4), (200, 4), (1000, 16)])
def test_solve_bk(uplo, n, nrhs):
    A0 = _sym(n, 60 + n, 0.05)
    A = A0.copy(order="F")
    ipiv = np.zeros(n, np.int64)
    mo.DecomposeBK(A, np.zeros((n, 2), order="F"), ipiv, uplo, 0)
    B0 = np.asfortranarray(np.random.default_rng(n).standard_normal((n, nrhs)))
    X = B0.copy(order="F")
    mo.SolveBK(X, A, ipiv, uplo)
    low = 1 if uplo == mo.LOWER else 0
    ref = lapack.dsytrs(*lapack.dsytrf(A0, lower=low)[:2], B0, lower=low)[0]
    tol = 1e3 * n * EPS * np.linalg.cond(A0) * np.abs(ref).max()
    assert np.abs(X - ref).max() <= tol
    own = lapack.dsytrs(A, ipiv.astype(np.int32), B0, lower=low)[0]  # LAPACK on our factors
    assert np.abs(X - own).max() <= tol
    if uplo == mo.LOWER and n <= 200:
        Xo = B0.copy(order="F")
        _oracle(O.SolveBK, Xo, A, ipiv, mo.LOWER)
        assert np.abs(X - Xo).max() <= tol


@pytest.mark.parametrize("uplo", [mo.LOWER, mo.UPPER])
@pytest.mark.parametrize("n", [4097, 8192, 8300])
def test_solve_bk_large_orders(uplo, n):
    # the solve keeps the right-hand side in LDS up to n = 8192 (4 rows per
    # thread to 4096, 8 beyond) and sweeps in global memory above
    A0 = _sym(n, 90 + n, 0.05)
    A = A0.copy(order="F")
    ipiv = np.zeros(n, np.int64)
    mo.DecomposeBK(A, np.zeros((n, 2), order="F"), ipiv, uplo, 0)
    assert (ipiv < 0).any()
    B0 = np.asfortranarray(np.random.default_rng(n).standard_normal((n, 3)))
    X = B0.copy(order="F")
    mo.SolveBK(X, A, ipiv, uplo)
    assert _eta(A0, X, B0) <= 10 * n * EPS


# ------------------------------------------------- full size, device twins
@pytest.mark.parametrize("uplo", [mo.LOWER, mo.UPPER])
def test_ldl_and_bk_4096_reconstruct_and_solve(uplo):
    n = 4096
    lower = uplo == mo.LOWER
    B0 = np.asfortranarray(np.random.default_rng(3).standard_normal((n, 2)))
    A0 = _sym(n, 77)
    A = A0.copy(order="F")
    ipiv = np.zeros(n, np.int64)
    _, err = mo.DecomposeLDL(A, None, ipiv, uplo, 0)
    assert err is None
    assert _ldl_backward(A0, A, ipiv, lower) <= 2 * n * EPS
    X = B0.copy(order="F")
    mo.SolveLDL(X, A, ipiv, uplo)
    assert _eta(A0, X, B0) <= 10 * n * EPS * max(1.0, _growth(A0, A, lower))
    A0 = _sym(n, 78, 0.05)
    A = A0.copy(order="F")
    _, err = mo.DecomposeBK(A, np.zeros((n, 2), order="F"), ipiv, uplo, 0)
    assert err is None
    ldu, lpiv, _ = lapack.dsytrf(A0, lower=1 if lower else 0)
    np.testing.assert_array_equal(ipiv, lpiv.astype(np.int64))
    tri = np.tril_indices(n) if lower else np.triu_indices(n)
    assert np.abs(A[tri] - ldu[tri]).max() <= 10 * n * EPS * max(1.0, np.abs(ldu[tri]).max())
    X = B0.copy(order="F")
    mo.SolveBK(X, A, ipiv, uplo)
    assert _eta(A0, X, B0) <= 10 * n * EPS


def test_device_twins_match_host():
    import torch
    dev = torch.device("cuda:0")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    Lb = _lib.lib()

    def d(X):
        return torch.from_numpy(np.ascontiguousarray(X.T)).to(dev).T

    n = 600
    A0 = _sym(n, 5, 0.05)
    B0 = np.asfortranarray(np.random.default_rng(4).standard_normal((n, 3)))
    for uplo, nb in ((mo.LOWER, 0), (mo.UPPER, 0), (mo.UPPER, 64), (mo.LOWER, 64)):
        A = A0.copy(order="F")
        ipiv = np.zeros(n, np.int64)
        mo.DecomposeLDL(A, None, ipiv, uplo, nb)
        dA = d(A0)
        dp = torch.zeros(n, dtype=torch.int64, device=dev)
        dinfo = torch.zeros(1, dtype=torch.int32, device=dev)
        assert Lb.mops_d_dsytrf_ldl(uplo, n, nb, dA.data_ptr(), n, dp.data_ptr(), n,
                                    dinfo.data_ptr(), s) == 0
        np.testing.assert_array_equal(dA.cpu().numpy(), A)
        np.testing.assert_array_equal(dp.cpu().numpy(), ipiv)
        assert int(dinfo.item()) == -1
        X = B0.copy(order="F")
        mo.SolveLDL(X, A, ipiv, uplo)
        dB = d(B0)
        assert Lb.mops_d_dsytrs_ldl(uplo, n, 3, dA.data_ptr(), n, dp.data_ptr(), n,
                                    dB.data_ptr(), n, s) == 0
        np.testing.assert_array_equal(dB.cpu().numpy(), X)
    for uplo in (mo.LOWER, mo.UPPER):
        A = A0.copy(order="F")
        ipiv = np.zeros(n, np.int64)
        mo.DecomposeBK(A, np.zeros((n, 2), order="F"), ipiv, uplo, 0)
        dA = d(A0)
        dp = torch.zeros(n, dtype=torch.int64, device=dev)
        assert Lb.mops_d_dsytrf_bk(uplo, n, 0, dA.data_ptr(), n, dp.data_ptr(), n, s) == 0
        np.testing.assert_array_equal(dA.cpu().numpy(), A)
        np.testing.assert_array_equal(dp.cpu().numpy(), ipiv)
        X = B0.copy(order="F")
        mo.SolveBK(X, A, ipiv, uplo)
        dB = d(B0)
        assert Lb.mops_d_dsytrs_bk(uplo, n, 3, dA.data_ptr(), n, dp.data_ptr(), n,
                                   dB.data_ptr(), n, s) == 0
        np.testing.assert_array_equal(dB.cpu().numpy(), X)
    # a zero pivot through the device twin's info
    Z = d(np.zeros((4, 4), order="F"))
    dp = torch.zeros(4, dtype=torch.int64, device=dev)
    dinfo = torch.zeros(1, dtype=torch.int32, device=dev)
    assert Lb.mops_d_dsytrf_ldl(mo.LOWER, 4, 0, Z.data_ptr(), 4, dp.data_ptr(), 4,
                                dinfo.data_ptr(), s) == 0
    torch.cuda.synchronize()
    assert int(dinfo.item()) == 0
    assert Lb.mops_d_status(s) == 0
