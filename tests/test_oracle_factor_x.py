"""CPU pins of the oracle's factorizations without pivoting and of its
triangular inverse (DecomposeLUnoPiv, lu.go:38-96, 291-299;
DecomposeLDLnoPiv / SolveLDLnoPiv, ldlnp.go:23-264; InverseTrm,
trinv.go:16-245), on the restated kernels and on the reference's calgo C.

Pins: exact integer cases (factors and inverses with integer entries are
recovered exactly), LAPACK / numpy on random inputs (the LU of a strongly
diagonally dominant matrix needs no interchange, so LAPACK's partial
pivoting returns the no-pivot factors; the L D L^T factors of an SPD matrix
follow from its Cholesky factor), residuals, and the triangles that must
stay untouched.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import oracle as O
from helpers import EPS, fmat

KINDS = ["port"] + (["reference"] if O.reference_available() else [])
LOWER, UPPER, UNIT = O.LOWER, O.UPPER, O.UNIT


@pytest.fixture(params=KINDS)
def kind(request):
    O.use(request.param)
    yield request.param
    O.use("port")


def _dd(m, n, seed):
    """Strongly diagonally dominant m x n matrix."""
    A = fmat(m, n, seed, "normal")
    k = min(m, n)
    A[np.arange(k), np.arange(k)] += 2.0 * max(m, n)
    return np.asfortranarray(A)


# ------------------------------------------------------------- LU no pivot


def test_lunopiv_exact_integer(kind):
    L = np.array([[1, 0, 0, 0], [2, 1, 0, 0], [-1, 3, 1, 0], [4, -2, 5, 1.0]])
    U = np.array([[2, 1, -1, 3], [0, 3, 2, 1], [0, 0, -4, 2], [0, 0, 0, 5.0]])
    for nb in (0, 2):
        A = np.asfortranarray(L @ U)
        O.DecomposeLUnoPiv(A, nb)
        np.testing.assert_array_equal(np.tril(A, -1) + np.eye(4), L)
        np.testing.assert_array_equal(np.triu(A), U)


@pytest.mark.parametrize("m,n,nb", [(50, 50, 0), (130, 130, 16), (300, 300, 68), (60, 90, 16),
                                    (70, 200, 0)])
def test_lunopiv_matches_lapack(kind, m, n, nb):
    A0 = _dd(m, n, 100 + m + n)
    A = A0.copy(order="F")
    O.DecomposeLUnoPiv(A, nb)
    P, Ll, Ul = sla.lu(A0)
    assert np.array_equal(P, np.eye(m))  # no interchange needed
    k = min(m, n)
    L = np.tril(A, -1)[:, :k] + np.eye(m, k)
    U = np.triu(A)[:k]
    np.testing.assert_allclose(L, Ll, rtol=0, atol=50 * n * EPS)
    np.testing.assert_allclose(U, Ul, rtol=0, atol=50 * n * EPS * np.abs(Ul).max())
    assert np.linalg.norm(A0 - L @ U) / np.linalg.norm(A0) <= 10 * n * EPS


# ------------------------------------------------------------ LDL no pivot


def _spd(n, seed):
    G = fmat(n, n, seed, "normal")
    A = G @ G.T / n + np.eye(n)
    return np.asfortranarray(np.tril(A) + np.tril(A, -1).T)


def test_ldlnopiv_exact_integer(kind):
    L = np.array([[1, 0, 0, 0], [3, 1, 0, 0], [-2, 1, 1, 0], [1, -1, 2, 1.0]])
    D = np.diag([2.0, -1.0, 4.0, 3.0])  # indefinite: no pivoting needed here
    A0 = L @ D @ L.T
    for nb in (0, 2):
        A = np.asfortranarray(A0.copy())
        A[np.triu_indices(4, 1)] = 99.0
        O.DecomposeLDLnoPiv(A, LOWER, nb)
        np.testing.assert_array_equal(np.tril(A, -1) + np.eye(4), L)
        np.testing.assert_array_equal(np.diag(A), np.diag(D))
        np.testing.assert_array_equal(A[np.triu_indices(4, 1)], 99.0)
    # the upper form of the same numbers: A = U D U^T with U = J L J
    J = np.eye(4)[::-1]
    U, Du = J @ L @ J, J @ D @ J
    for nb in (0, 2):
        A = np.asfortranarray(U @ Du @ U.T)
        O.DecomposeLDLnoPiv(A, UPPER, nb)
        np.testing.assert_array_equal(np.triu(A, 1) + np.eye(4), U)
        np.testing.assert_array_equal(np.diag(A), np.diag(Du))


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("n,nb", [(40, 0), (150, 32), (300, 68)])
def test_ldlnopiv_against_cholesky_and_solve(kind, uplo, n, nb):
    A0 = _spd(n, 7 + n)
    A = A0.copy(order="F")
    assert O.DecomposeLDLnoPiv(A, uplo, nb) == 0
    d = np.diag(A)
    if uplo == LOWER:
        Lc = np.linalg.cholesky(A0)
        Lr = Lc / np.diag(Lc)
        np.testing.assert_allclose(np.tril(A, -1) + np.eye(n), Lr, rtol=0, atol=100 * n * EPS)
        np.testing.assert_allclose(d, np.diag(Lc) ** 2, rtol=100 * n * EPS)
        F = np.tril(A, -1) + np.eye(n)
    else:
        J = np.eye(n)[::-1]
        Lc = np.linalg.cholesky(J @ A0 @ J)
        Ur = J @ (Lc / np.diag(Lc)) @ J
        np.testing.assert_allclose(np.triu(A, 1) + np.eye(n), Ur, rtol=0, atol=100 * n * EPS)
        F = np.triu(A, 1) + np.eye(n)
    res = np.linalg.norm(A0 - F @ np.diag(d) @ F.T) / np.linalg.norm(A0)
    assert res <= 10 * n * EPS
    B0 = fmat(n, 5, 9, "normal")
    X = B0.copy(order="F")
    O.SolveLDLnoPiv(X, A, uplo)
    np.testing.assert_allclose(X, np.linalg.solve(A0, B0), rtol=0, atol=1e3 * n * EPS)


# --------------------------------------------------------------- InverseTrm


def test_inverse_unit_integer_exact(kind):
    L = np.array([[1, 0, 0, 0], [2, 1, 0, 0], [-1, 3, 1, 0], [4, -2, 5, 1.0]])
    for nb in (0, 2):
        for uplo, T in ((LOWER, L), (UPPER, L.T)):
            A = np.asfortranarray(T.copy())
            A[np.diag_indices(4)] = 7.0  # UNIT: never read, never written
            O.InverseTrm(A, uplo | UNIT, nb)
            Ti = np.round(np.linalg.inv(T))
            np.testing.assert_array_equal(Ti @ T, np.eye(4))  # the exact inverse
            tri = np.tril_indices(4, -1) if uplo == LOWER else np.triu_indices(4, 1)
            np.testing.assert_array_equal(A[tri], Ti[tri])
            np.testing.assert_array_equal(np.diag(A), 7.0)


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("unit", [0, UNIT])
@pytest.mark.parametrize("n,nb", [(30, 0), (150, 32), (300, 68)])
def test_inverse_trm_against_numpy(kind, uplo, unit, n, nb):
    T = fmat(n, n, 21 + n, "normal") * 0.3
    T[np.diag_indices(n)] = 2.0 + np.abs(fmat(n, 1, 22)[:, 0])
    T = np.triu(T) if uplo == UPPER else np.tril(T)
    A = np.asfortranarray(T + (np.tril(np.full((n, n), -5.0), -1) if uplo == UPPER
                               else np.triu(np.full((n, n), -5.0), 1)))
    Te = T.copy()
    if unit:
        Te[np.diag_indices(n)] = 1.0
    O.InverseTrm(A, uplo | unit, nb)
    Ti = np.linalg.inv(Te)
    tri = np.triu_indices(n) if uplo == UPPER else np.tril_indices(n)
    strict = np.triu_indices(n, 1) if uplo == UPPER else np.tril_indices(n, -1)
    other = np.tril_indices(n, -1) if uplo == UPPER else np.triu_indices(n, 1)
    cond = np.linalg.cond(Te, np.inf)
    err = np.abs(A[strict if unit else tri] - Ti[strict if unit else tri]).max()
    assert err <= 20 * n * EPS * cond * np.abs(Ti).max()
    np.testing.assert_array_equal(A[other], -5.0)
    if unit:
        np.testing.assert_array_equal(np.diag(A), np.diag(T))
