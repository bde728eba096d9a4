"""GPU parity of DecomposeLUnoPiv, DecomposeLDLnoPiv / SolveLDLnoPiv and
InverseTrm through the C ABI, against the oracle on the same inputs (the
reference's calgo kernels under the restated Go glue) and against LAPACK /
numpy. Without pivoting the factors are unique, so both sides agree to
rounding: bounds are normwise, 20 n eps times the condition number.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla

import matops_amd as mo
from matops_amd import _lib
from helpers import EPS, embed, fmat
from oracle import oracle as O

pytestmark = pytest.mark.gpu
ORACLE_KIND = "reference" if O.reference_available() else "port"


def _oracle(fn, *args):
    O.use(ORACLE_KIND)
    try:
        return fn(*args)
    finally:
        O.use("port")


def _dd(m, n, seed):
    A = fmat(m, n, seed, "normal")
    k = min(m, n)
    A[np.arange(k), np.arange(k)] += 2.0 * max(m, n)
    return np.asfortranarray(A)


def _spd(n, seed):
    G = fmat(n, n, seed, "normal")
    A = G @ G.T / n + np.eye(n)
    return np.asfortranarray(np.tril(A) + np.tril(A, -1).T)


@pytest.mark.parametrize("m,n,nb", [(1, 1, 0), (4, 4, 0), (64, 64, 16), (65, 65, 0),
                                    (300, 300, 68), (200, 500, 32), (1000, 1000, 256),
                                    (700, 300, 64)])
def test_lunopiv(m, n, nb):
    A0 = _dd(m, n, 100 + m + n)
    A = A0.copy(order="F")
    _, err = mo.DecomposeLUnoPiv(A, nb)
    assert err is None
    k = min(m, n)
    L = np.tril(A, -1)[:, :k] + np.eye(m, k)
    U = np.triu(A)[:k]
    assert np.linalg.norm(A0 - L @ U) / np.linalg.norm(A0) <= 10 * n * EPS
    if m <= n:  # the reference's loops cover square and wide matrices
        Ao = A0.copy(order="F")
        _oracle(O.DecomposeLUnoPiv, Ao, nb)
        assert np.abs(A - Ao).max() <= 20 * n * EPS * np.abs(Ao).max()
    P, _, Ul = sla.lu(A0)
    assert np.array_equal(P, np.eye(m))
    assert np.abs(U - Ul).max() <= 20 * n * EPS * np.abs(Ul).max()


def test_lunopiv_exact_integer_and_view():
    L = np.array([[1, 0, 0, 0], [2, 1, 0, 0], [-1, 3, 1, 0], [4, -2, 5, 1.0]])
    U = np.array([[2, 1, -1, 3], [0, 3, 2, 1], [0, 0, -4, 2], [0, 0, 0, 5.0]])
    P, V = embed(L @ U)
    mo.DecomposeLUnoPiv(V, 0)
    np.testing.assert_array_equal(np.tril(V, -1) + np.eye(4), L)
    np.testing.assert_array_equal(np.triu(V), U)
    mask = np.ones(P.shape, bool)
    mask[3:7, 2:6] = False
    assert np.all(np.isnan(P[mask]))


@pytest.mark.parametrize("uplo", [mo.LOWER, mo.UPPER])
@pytest.mark.parametrize("n,nb", [(1, 0), (40, 0), (64, 16), (150, 32), (300, 68), (1000, 256),
                                  (2049, 256)])
def test_ldlnopiv_and_solve(uplo, n, nb):
    A0 = _spd(n, 7 + n)
    A = A0.copy(order="F")
    other = np.triu_indices(n, 1) if uplo == mo.LOWER else np.tril_indices(n, -1)
    A[other] = 55.0
    W = np.zeros((n, max(nb, 1)), order="F")
    R, err = mo.DecomposeLDLnoPiv(A, W, uplo, nb)
    assert err is None and R is A
    np.testing.assert_array_equal(A[other], 55.0)
    F = (np.tril(A, -1) if uplo == mo.LOWER else np.triu(A, 1)) + np.eye(n)
    d = np.diag(A)
    assert np.linalg.norm(A0 - F @ np.diag(d) @ F.T) / np.linalg.norm(A0) <= 10 * n * EPS
    if n <= 1000:
        Ao = A0.copy(order="F")
        Ao[other] = 55.0
        _oracle(O.DecomposeLDLnoPiv, Ao, uplo, nb)
        assert np.abs(A - Ao).max() <= 20 * n * EPS * 5.0 * np.abs(Ao).max()
    B0 = fmat(n, 9, 9, "normal")
    X = B0.copy(order="F")
    mo.SolveLDLnoPiv(X, A, uplo)
    eta = np.linalg.norm(A0 @ X - B0) / (np.linalg.norm(A0) * np.linalg.norm(X) +
                                         np.linalg.norm(B0))
    assert eta <= 10 * n * EPS


def test_ldlnopiv_reference_errors():
    _, err = mo.DecomposeLDLnoPiv(np.zeros((3, 4), order="F"), None, mo.LOWER, 0)
    assert str(err) == "A not a square matrix"
    _, err = mo.DecomposeLDLnoPiv(_spd(100, 1), None, mo.LOWER, 32)
    assert str(err) == "No workspace for blocking invocation"


@pytest.mark.parametrize("uplo", [mo.LOWER, mo.UPPER])
@pytest.mark.parametrize("unit", [0, mo.UNIT])
@pytest.mark.parametrize("n,nb", [(1, 0), (30, 0), (150, 32), (300, 68), (1500, 256)])
def test_inverse_trm(uplo, unit, n, nb):
    T = fmat(n, n, 21 + n, "normal") * 0.3
    T[np.diag_indices(n)] = 2.0 + np.abs(fmat(n, 1, 22)[:, 0])
    T = np.triu(T) if uplo == mo.UPPER else np.tril(T)
    A0 = np.asfortranarray(T + (np.tril(np.full((n, n), -5.0), -1) if uplo == mo.UPPER
                                else np.triu(np.full((n, n), -5.0), 1)))
    Te = T.copy()
    if unit:
        Te[np.diag_indices(n)] = 1.0
    A = A0.copy(order="F")
    _, err = mo.InverseTrm(A, uplo | unit, nb)
    assert err is None
    Ti = sla.solve_triangular(Te, np.eye(n), lower=uplo == mo.LOWER)
    tri = np.triu_indices(n) if uplo == mo.UPPER else np.tril_indices(n)
    strict = np.triu_indices(n, 1) if uplo == mo.UPPER else np.tril_indices(n, -1)
    other = np.tril_indices(n, -1) if uplo == mo.UPPER else np.triu_indices(n, 1)
    sel = strict if unit else tri
    cond = np.linalg.cond(Te, np.inf)
    if sel[0].size:
        assert np.abs(A[sel] - Ti[sel]).max() <= 20 * n * EPS * cond * np.abs(Ti).max()
    np.testing.assert_array_equal(A[other], -5.0)
    if unit:
        np.testing.assert_array_equal(np.diag(A), np.diag(T))
    if n <= 300 and sel[0].size:
        Ao = A0.copy(order="F")
        _oracle(O.InverseTrm, Ao, uplo | unit, nb)
        assert np.abs(A[sel] - Ao[sel]).max() <= 40 * n * EPS * cond * np.abs(Ti).max()


def test_device_twins_match_host():
    import torch
    dev = torch.device("cuda:0")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    Lb = _lib.lib()

    def d(X):
        return torch.from_numpy(np.ascontiguousarray(X.T)).to(dev).T

    n = 700
    A0 = _dd(n, n, 5)
    A = A0.copy(order="F")
    mo.DecomposeLUnoPiv(A, 64)
    dA = d(A0)
    assert Lb.mops_d_dgetrf_nopiv(n, n, 64, dA.data_ptr(), n, s) == 0
    np.testing.assert_array_equal(dA.cpu().numpy(), A)
    S0 = _spd(n, 6)
    S = S0.copy(order="F")
    mo.DecomposeLDLnoPiv(S, np.zeros((n, 64), order="F"), mo.LOWER, 64)
    dS = d(S0)
    assert Lb.mops_d_dsytrf_nopiv(mo.LOWER, n, 64, dS.data_ptr(), n, s) == 0
    np.testing.assert_array_equal(dS.cpu().numpy(), S)
    B0 = fmat(n, 4, 7)
    X = B0.copy(order="F")
    mo.SolveLDLnoPiv(X, S, mo.LOWER)
    dB = d(B0)
    assert Lb.mops_d_dsytrs_nopiv(mo.LOWER, n, 4, dS.data_ptr(), n, dB.data_ptr(), n, s) == 0
    np.testing.assert_array_equal(dB.cpu().numpy(), X)
    T = np.asfortranarray(np.tril(fmat(n, n, 8) * 0.2) + 3 * np.eye(n))
    Ti = T.copy(order="F")
    mo.InverseTrm(Ti, mo.LOWER, 64)
    dT = d(T)
    assert Lb.mops_d_dinverse_trm(mo.LOWER, n, 64, dT.data_ptr(), n, s) == 0
    np.testing.assert_array_equal(dT.cpu().numpy(), Ti)
    torch.cuda.synchronize()
    assert Lb.mops_d_status(s) == 0
