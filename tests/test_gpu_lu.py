"""GPU parity of the LU path (DecomposeLU / SolveLU through the C ABI):
pivot vectors identical to the oracle's (and to LAPACK's), factor residual
||PA - LU|| / ||A|| <= 10*n*eps, solves at O(eps) backward error, the golden
fixtures, and the edge cases the reference's code admits (nb = 0, nb >= n,
tall and wide matrices, exact ties, a zero pivot, views with ld > rows).
"""
import os
import sys

import numpy as np
import pytest
from scipy.linalg import lapack

import matops_amd as mo
from helpers import EPS, embed

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import golden_inputs as gi  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = os.path.join(ROOT, "tests", "golden", "golden.npz")


def _perm_matrix(piv, m):
    P = np.eye(m)
    for k, p in enumerate(piv):
        P[[k, p]] = P[[p, k]]
    return P


def _residual(A0, F, piv):
    m, n = A0.shape
    k = min(m, n)
    L = np.tril(F, -1)[:, :k] + np.eye(m, k)
    U = np.triu(F)[:k, :]
    return np.linalg.norm(_perm_matrix(piv[:k], m) @ A0 - L @ U) / np.linalg.norm(A0)


def test_golden_lu():
    d = np.load(GOLD)
    for n, nb in gi.LU_CASES:
        A, B = gi.lu_case(n)
        F = A.copy(order="F")
        piv = np.zeros(n, dtype=np.int64)
        _, err = mo.DecomposeLU(F, piv, nb)
        assert err is None
        assert np.array_equal(piv, d[f"lu_{n}_{nb}_piv"]), (n, nb)
        if n <= 64:
            np.testing.assert_allclose(F, d[f"lu_{n}_{nb}_F"], rtol=0, atol=1e-12 * n)
        X = B.copy(order="F")
        mo.SolveLU(X, F, piv, mo.NONE)
        np.testing.assert_allclose(X, d[f"lu_{n}_{nb}_X"], rtol=1e-9, atol=1e-10)


def test_ties_and_kat():
    d = np.load(GOLD)
    F = gi.ties_case().copy(order="F")
    piv = np.zeros(4, dtype=np.int64)
    mo.DecomposeLU(F, piv, 0)
    assert piv.tolist() == d["ties_piv"].tolist()
    np.testing.assert_allclose(F, d["ties_F"], atol=1e-15)
    A = np.asfortranarray([[3.0, 2, 2], [6, 4, 1], [4, 6, 3]])
    piv = np.zeros(3, dtype=np.int64)
    mo.DecomposeLU(A, piv, 0)
    assert piv.tolist() == [1, 2, 2]  # decomp_test.go:84, by hand


@pytest.mark.parametrize("n,nb", [(1, 0), (2, 1), (31, 0), (32, 16), (33, 32), (100, 12),
                                  (257, 64), (513, 256), (1000, 256), (1536, 128)])
def test_lu_square_vs_oracle(n, nb):
    A0 = gi.mat(n, n, 200 + n, "normal")
    F = A0.copy(order="F")
    piv = np.zeros(n, dtype=np.int64)
    mo.DecomposeLU(F, piv, nb)
    Fo = A0.copy(order="F")
    pivo = np.zeros(n, dtype=np.int64)
    O.use("port")
    O.DecomposeLU(Fo, pivo, nb)
    mism = np.nonzero(piv != pivo)[0]
    assert mism.size == 0, f"first pivot mismatch at column {mism[:1]}"
    assert _residual(A0, F, piv) <= 10 * n * EPS
    if n <= 300:
        np.testing.assert_allclose(F, Fo, rtol=0, atol=1e-10 * max(1, np.abs(Fo).max()))


@pytest.mark.parametrize("m,n,nb", [(300, 120, 64), (130, 300, 64), (97, 40, 0), (40, 97, 0)])
def test_lu_rectangular_vs_lapack(m, n, nb):
    A0 = gi.mat(m, n, 300 + m + n, "normal")
    F = A0.copy(order="F")
    piv = np.zeros(min(m, n), dtype=np.int64)
    mo.DecomposeLU(F, piv, nb)
    lu, ipiv, info = lapack.dgetrf(A0)
    assert np.array_equal(piv, ipiv)
    np.testing.assert_allclose(F, lu, rtol=0, atol=1e-11)
    assert _residual(A0, F, piv) <= 10 * max(m, n) * EPS


def test_lu_on_views_and_long_pivot_slice():
    n = 150
    A0 = gi.mat(n, n, 77, "normal")
    P, V = embed(A0, 9, 5, 3)
    piv = np.full(n + 10, -5, dtype=np.int64)  # longer than min(m,n): tail cleared
    mo.DecomposeLU(V, piv, 32)
    assert np.all(piv[n:] == 0)
    pivo = np.zeros(n, dtype=np.int64)
    Fo = A0.copy(order="F")
    O.DecomposeLU(Fo, pivo, 32)
    assert np.array_equal(piv[:n], pivo)
    assert np.isnan(P[:5]).all() and np.isnan(P[:, :3]).all()


def test_short_pivot_array_is_an_error():
    F = gi.mat(10, 10, 1)
    _, err = mo.DecomposeLU(F, np.zeros(9, dtype=np.int64), 0)
    assert str(err) == "pivot array < min(A.Rows(),A.Cols())"


def test_zero_pivot_propagates_without_error():
    n = 64
    A0 = gi.mat(n, n, 5, "normal")
    A0[:, 10] = 0.0  # singular: column 10 becomes a zero pivot column
    F = A0.copy(order="F")
    piv = np.zeros(n, dtype=np.int64)
    _, err = mo.DecomposeLU(F, piv, 16)
    assert err is None
    assert piv[10] == 10  # all candidates are 0: nothing beats the diagonal
    assert np.all(piv >= np.arange(n)) and np.all(piv < n)


@pytest.mark.parametrize("trans", [mo.NONE, mo.TRANSA])
def test_solve_lu(trans):
    n, nrhs = 700, 9
    A0 = gi.mat(n, n, 11, "normal")
    B0 = gi.mat(n, nrhs, 12, "normal")
    F = A0.copy(order="F")
    piv = np.zeros(n, dtype=np.int64)
    mo.DecomposeLU(F, piv, 128)
    X = B0.copy(order="F")
    mo.SolveLU(X, F, piv, trans)
    op = A0.T if trans else A0
    eta = np.linalg.norm(op @ X - B0) / (np.linalg.norm(op) * np.linalg.norm(X) + np.linalg.norm(B0))
    assert eta <= 10 * n * EPS


def test_device_api_perm_and_pivots_agree():
    import ctypes

    import torch
    from matops_amd import _lib
    L = _lib.lib()
    dev = torch.device("cuda:0")
    n, nrhs = 1024, 16
    A0 = gi.mat(n, n, 21, "normal")
    B0 = gi.mat(n, nrhs, 22, "normal")
    A = torch.from_numpy(A0.T.copy()).to(dev).T  # column-major on device
    ipiv = torch.zeros(n, dtype=torch.int64, device=dev)
    perm = torch.zeros(n, dtype=torch.int32, device=dev)
    st = torch.ones(1, dtype=torch.int32, device=dev)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.mops_d_dgetrf(n, n, 256, A.data_ptr(), n, ipiv.data_ptr(), perm.data_ptr(),
                           st.data_ptr(), s) == 0
    X1 = torch.from_numpy(B0.T.copy()).to(dev).T
    X2 = X1.clone()
    assert L.mops_d_dgetrs(0, n, nrhs, A.data_ptr(), n, ipiv.data_ptr(), n, perm.data_ptr(),
                           X1.data_ptr(), n, s) == 0
    assert L.mops_d_dgetrs(0, n, nrhs, A.data_ptr(), n, ipiv.data_ptr(), n, None,
                           X2.data_ptr(), n, s) == 0
    torch.cuda.synchronize()
    assert st.item() == 0
    assert torch.equal(X1, X2)
    p = ipiv.cpu().numpy()
    # perm is the pivot sequence applied to the identity
    ident = np.arange(n)
    for k, q in enumerate(p):
        ident[[k, q]] = ident[[q, k]]
    assert np.array_equal(perm.cpu().numpy(), ident)


def _bits(x):
    return np.asarray(x, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("case", ["nan_below", "nan_diag", "signalling_nan_rows"])
def test_lu_nan_pivot_rules(case):
    """pivotIndex (pivot.go:98-107): a NaN below the diagonal never wins, a
    NaN on the diagonal is never replaced; pivots equal the oracle's. The
    last case puts the signalling-NaN pattern 0x7FF5A5A5A5A5A5A5 into two
    rows that are published as pivot candidates in the first column (by the
    first and by the second workgroup of rows): their data must arrive
    unchanged, and the untouched U row keeps the exact bits."""
    n = 700
    A0 = gi.mat(n, n, 4242, "normal")
    sent = np.array([0x7FF5A5A5A5A5A5A5], dtype=np.uint64).view(np.float64)[0]
    if case == "nan_below":
        A0[50, 3] = np.nan
        A0[600, 40] = np.nan
    elif case == "nan_diag":
        A0[7, 7] = np.nan
    else:
        A0[0, 0] = 50.0            # row 0 is the first pivot (and the old row k)
        A0[600, 0] = 40.0          # the best candidate of the rows 512..1023
        A0[0, 20] = sent
        A0[600, 25] = sent
    F = A0.copy(order="F")
    piv = np.zeros(n, dtype=np.int64)
    _, err = mo.DecomposeLU(F, piv, 64)
    assert err is None
    Fo = A0.copy(order="F")
    pivo = np.zeros(n, dtype=np.int64)
    O.use("port")
    O.DecomposeLU(Fo, pivo, 64)
    mism = np.nonzero(piv != pivo)[0]
    assert mism.size == 0, f"first pivot mismatch at column {mism[:1]}"
    if case == "signalling_nan_rows":
        assert piv[0] == 0
        assert _bits(F[0, 20]) == _bits(sent)
        np.testing.assert_array_equal(_bits(np.delete(F[0], 20)), _bits(np.delete(A0[0], 20)))
