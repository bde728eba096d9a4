"""CPU pins of the oracle's symmetric indefinite factorizations with
pivoting, on the restated kernels and on the reference's calgo C:

* DecomposeLDL / SolveLDL (ldl.go:25-511): L D L^T or U D U^T with 1x1
  diagonal pivots (the largest |updated diagonal| by the reference's IAMax,
  calgo/dvec_util.c:164-195, whose pair logic is pinned on its own).
* DecomposeBK / SolveBK (ldlbk.go, ldlbks.go): Bunch-Kaufman 1x1 / 2x2
  pivots, the forms of LAPACK's dsytf2.

Pins that do not restate the algorithm: the reconstruction P A P^T =
T D T^T from the stored factors and pivots; the diagonal special case
(pivoting sorts the diagonal by magnitude); block-local pivoting of the
blocked UPPER LDL; LAPACK's dsytrf (scipy), which gives the same pivot
vector as the reference's BK on the same input (same alpha, same pivot
rule, same 1-based format with negative entries for 2x2 blocks) and the
same factors to rounding; solve residuals.

Found while pinning: with W of nb columns (as DecomposeLDL's documentation
asks) the blocked LOWER LDL (ldl.go:90-216) uses W's last column both for
the updated diagonal and for the last column of each block, so the pivot
swap of that column is undone (findAndBuildPivot swaps wcol[0] and
wcol[dmax], then the same two entries of the diagonal, which are the same
memory) and the factorization is wrong; with nb + 1 columns it equals the
unblocked one. And the reference's GEMV, calgo/dmvec_gemv.c:241, starts Y at
R times X's increment (not Y's) on its unaligned path (odd leading
dimension) once it splits more than MB = 68 rows into blocks: the blocked
LOWER LDL's panel GEMV has a row of W as X, so for an odd order above 68
the reference kernels write outside Y. Those shapes run on the restated
kernels only (test_ldl_blocked_lower_odd_order_port).
"""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from oracle import oracle as O
from helpers import EPS

KINDS = ["port"] + (["reference"] if O.reference_available() else [])
LOWER, UPPER = O.LOWER, O.UPPER


@pytest.fixture(params=KINDS)
def kind(request):
    O.use(request.param)
    yield request.param
    O.use("port")


def _sym(n, seed, diag_scale=1.0):
    g = np.random.default_rng(seed)
    G = g.standard_normal((n, n))
    A = G + G.T
    A[np.diag_indices(n)] *= diag_scale
    return A


def _perm(ipiv, n, lower):
    """P with P A P^T = T D T^T for the LDL row pivots: FORWARD order for
    LOWER, BACKWARD for UPPER (as SolveLDL applies them)."""
    P = np.eye(n)
    for k in (range(n) if lower else range(n - 1, -1, -1)):
        if ipiv[k] > 0:
            j = ipiv[k] - 1
            P[[k, j]] = P[[j, k]]
    return P


def _ldl_resid(A0, A, ipiv, lower):
    n = A0.shape[0]
    T = (np.tril(A, -1) if lower else np.triu(A, 1)) + np.eye(n)
    P = _perm(ipiv, n, lower)
    return np.abs(P @ A0 @ P.T - T @ np.diag(np.diag(A)) @ T.T).max() / np.abs(A0).max()


# --------------------------------------------------------------------- IAMax
@pytest.mark.parametrize("x,want", [([5, 1, 0], 0), ([1, 5, 0], 1), ([3, 3, 3], 0),
                                    ([0, 0, 0], 0), ([2, 1, 7, 7, 7], 2), ([1, 2, 3, 1, 2], 2),
                                    ([0, 0, 0, 0, 1], 4), ([-9, 9], 0), ([1], 0),
                                    ([np.nan, 3, 1], 2), ([1, np.nan, 3], 2),
                                    ([np.nan, np.nan], 0)])
def test_iamax_pair_logic(kind, x, want):
    # first strict maximum of |x|; a NaN in a pair's first slot hides its
    # partner, and NaN never wins
    assert O.iamax(np.array(x, dtype=np.float64)) == want


def test_iamax_random_is_first_max(kind):
    g = np.random.default_rng(1)
    for n in (2, 3, 17, 64, 1001):
        x = np.round(g.standard_normal(n) * 4)  # ties are frequent
        assert O.iamax(x) == int(np.argmax(np.abs(x)))


# ------------------------------------------------------------------- LDL
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("n", [1, 2, 7, 50, 131])
def test_ldl_unblocked_reconstructs(kind, uplo, n):
    A0 = _sym(n, 10 + n)
    A = np.asfortranarray(A0.copy())
    ipiv = np.zeros(n, np.int64)
    assert O.DecomposeLDL(A, None, ipiv, uplo, 0) is None
    assert _ldl_resid(A0, A, ipiv, uplo == LOWER) <= 50 * n * EPS
    assert np.all((ipiv == 0) | ((ipiv >= 1) & (ipiv <= n)))


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
def test_ldl_diagonal_matrix_sorts_by_magnitude(kind, uplo):
    d = np.array([3.0, -7.0, 1.0, 7.0, -2.0, 0.5])
    A = np.asfortranarray(np.diag(d))
    ipiv = np.zeros(6, np.int64)
    assert O.DecomposeLDL(A, None, ipiv, uplo, 0) is None
    D = np.diag(A)
    if uplo == LOWER:  # the largest |d| first, ties to the lowest index
        np.testing.assert_array_equal(D, [-7.0, 7.0, 3.0, -2.0, 1.0, 0.5])
    else:              # from the bottom: the largest |d| of the leading part last
        np.testing.assert_array_equal(D, [0.5, 1.0, -2.0, 3.0, 7.0, -7.0])
    np.testing.assert_array_equal(np.tril(A, -1) + np.triu(A, 1), 0.0)


@pytest.mark.parametrize("n,nb", [(50, 8), (64, 16), (100, 32), (37, 8)])
def test_ldl_blocked_upper_pivots_within_blocks(kind, n, nb):
    A0 = _sym(n, 20 + n)
    A = np.asfortranarray(A0.copy())
    ipiv = np.zeros(n, np.int64)
    W = np.zeros((n, nb), order="F")
    assert O.DecomposeLDL(A, W, ipiv, UPPER, nb) is None
    assert _ldl_resid(A0, A, ipiv, False) <= 100 * n * EPS
    for k in range(n):  # block b of nb rows from the bottom: pivots stay inside it
        if ipiv[k]:
            e = n - ((n - 1 - k) // nb) * nb
            assert e - nb <= ipiv[k] - 1 < e, (k, ipiv[k])


@pytest.mark.parametrize("n,nb", [(50, 8), (64, 16), (100, 32)])
def test_ldl_blocked_lower_equals_unblocked_with_a_wider_workspace(kind, n, nb):
    A0 = _sym(n, 30 + n)
    A, Au = np.asfortranarray(A0.copy()), np.asfortranarray(A0.copy())
    ipiv, ipu = np.zeros(n, np.int64), np.zeros(n, np.int64)
    assert O.DecomposeLDL(A, np.zeros((n, nb + 1), order="F"), ipiv, LOWER, nb) is None
    assert O.DecomposeLDL(Au, None, ipu, LOWER, 0) is None
    np.testing.assert_array_equal(ipiv, ipu)
    assert np.abs(A - Au).max() <= 100 * n * EPS * np.abs(Au).max()


@pytest.mark.parametrize("n,nb", [(129, 64), (257, 64), (201, 68)])
def test_ldl_blocked_lower_odd_order_port(n, nb):
    # restated kernels only: the reference's GEMV writes out of bounds here
    O.use("port")
    A0 = _sym(n, 31 + n)
    A, Au = np.asfortranarray(A0.copy()), np.asfortranarray(A0.copy())
    ipiv, ipu = np.zeros(n, np.int64), np.zeros(n, np.int64)
    assert O.DecomposeLDL(A, np.zeros((n, nb + 1), order="F"), ipiv, LOWER, nb) is None
    assert O.DecomposeLDL(Au, None, ipu, LOWER, 0) is None
    np.testing.assert_array_equal(ipiv, ipu)
    assert np.abs(A - Au).max() <= 100 * n * EPS * np.abs(Au).max()
    assert _ldl_resid(A0, A, ipiv, True) <= 100 * n * EPS


def test_reference_blocked_ldl_lower_defect_with_nb_columns(kind):
    n, nb = 64, 16
    A0 = _sym(n, 5)
    A = np.asfortranarray(A0.copy())
    ipiv = np.zeros(n, np.int64)
    assert O.DecomposeLDL(A, np.zeros((n, nb), order="F"), ipiv, LOWER, nb) is None
    assert _ldl_resid(A0, A, ipiv, True) > 1e-3


@pytest.mark.parametrize("uplo,msg", [(LOWER, "zero value on diagonal"),
                                      (UPPER, "zero on diagonal.")])
def test_ldl_zero_pivot_error(kind, uplo, msg):
    A = np.asfortranarray(np.zeros((3, 3)))
    assert O.DecomposeLDL(A, None, np.zeros(3, np.int64), uplo, 0) == msg
    assert O.DecomposeLDL(np.zeros((3, 4), order="F"), None, np.zeros(3, np.int64), uplo,
                          0) == "A not a square matrix"


@pytest.mark.parametrize("uplo,nb", [(LOWER, 0), (UPPER, 0), (UPPER, 16)])
def test_solve_ldl(kind, uplo, nb):
    n = 90
    A0 = _sym(n, 7)
    A = np.asfortranarray(A0.copy())
    ipiv = np.zeros(n, np.int64)
    assert O.DecomposeLDL(A, np.zeros((n, max(nb, 1)), order="F"), ipiv, uplo, nb) is None
    B0 = np.random.default_rng(8).standard_normal((n, 5))
    X = np.asfortranarray(B0.copy())
    O.SolveLDL(X, A, ipiv, uplo)
    cond = np.linalg.cond(A0)
    assert np.abs(A0 @ X - B0).max() <= 100 * n * EPS * cond * np.abs(B0).max()


# -------------------------------------------------------------------- BK
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("n,scale", [(1, 1.0), (2, 0.01), (7, 0.05), (50, 0.05), (200, 0.05),
                                     (123, 1.0)])
def test_bk_equals_lapack_dsytrf(kind, uplo, n, scale):
    A0 = _sym(n, 40 + n, scale)  # a small diagonal brings 2x2 pivots
    A = np.asfortranarray(A0.copy())
    ipiv = np.zeros(n, np.int64)
    assert O.DecomposeBK(A, np.zeros((n, 5), order="F"), ipiv, uplo, 0) is None
    ldu, lpiv, info = lapack.dsytrf(A0, lower=1 if uplo == LOWER else 0)
    assert info == 0
    np.testing.assert_array_equal(ipiv, lpiv.astype(np.int64))
    tri = np.tril_indices(n) if uplo == LOWER else np.triu_indices(n)
    assert np.abs(A[tri] - ldu[tri]).max() <= 100 * n * EPS * max(1.0, np.abs(ldu[tri]).max())


def test_bk_reference_table_matrix(kind):
    # the 7 x 7 lower table of ldlbkl_test.go:62-69 (a 2x2 pivot at the
    # second column), against LAPACK
    L = np.array([[1, 0, 0, 0, 0, 0, 0], [1, 2, 0, 0, 0, 0, 0], [1, 2, 3, 0, 0, 0, 0],
                  [1, 2, 3, 4, 0, 0, 0], [1, 5, 3, 4, 5, 0, 0], [1, 2, 3, 4, 9, 6, 0],
                  [1, 2, 3, 4, 5, 6, 7]], dtype=np.float64)
    A0 = np.tril(L) + np.tril(L, -1).T
    A = np.asfortranarray(A0.copy())
    ipiv = np.zeros(7, np.int64)
    assert O.DecomposeBK(A, np.zeros((7, 5), order="F"), ipiv, LOWER, 0) is None
    ldu, lpiv, _ = lapack.dsytrf(A0, lower=1)
    np.testing.assert_array_equal(ipiv, lpiv.astype(np.int64))
    assert (ipiv < 0).any()
    tri = np.tril_indices(7)
    assert np.abs(A[tri] - ldu[tri]).max() <= 1e-13 * np.abs(ldu).max()


def test_bk_self_pivot_format(kind):
    # no interchange: a 1x1 pivot records its own 1-based row, LOWER and UPPER
    A0 = np.diag([4.0, 3.0, 2.0, 1.0])
    for uplo in (LOWER, UPPER):
        A = np.asfortranarray(A0.copy())
        ipiv = np.zeros(4, np.int64)
        assert O.DecomposeBK(A, np.zeros((4, 2), order="F"), ipiv, uplo, 0) is None
        np.testing.assert_array_equal(ipiv, [1, 2, 3, 4])


def test_bk_workspace_checks(kind):
    A = np.asfortranarray(_sym(10, 1))
    ip = np.zeros(10, np.int64)
    assert O.DecomposeBK(A.copy(order="F"), np.zeros((10, 1), order="F"), ip, LOWER, 0) == \
        "Workspace too small"
    assert O.DecomposeBK(A.copy(order="F"), np.zeros((9, 2), order="F"), ip, LOWER, 0) == \
        "Workspace too small"
    assert O.DecomposeBK(A.copy(order="F"), np.zeros((10, 4), order="F"), ip, LOWER, 4) == \
        "Workspace too small"


@pytest.mark.parametrize("n", [1, 2, 7, 60, 200])
def test_solve_bk_lower(kind, n):
    A0 = _sym(n, 60 + n, 0.05)
    A = np.asfortranarray(A0.copy())
    ipiv = np.zeros(n, np.int64)
    assert O.DecomposeBK(A, np.zeros((n, 5), order="F"), ipiv, LOWER, 0) is None
    B0 = np.random.default_rng(n).standard_normal((n, 4))
    X = np.asfortranarray(B0.copy())
    O.SolveBK(X, A, ipiv, LOWER)
    ref = lapack.dsytrs(*lapack.dsytrf(A0, lower=1)[:2], B0, lower=1)[0]
    assert np.abs(X - ref).max() <= 1e3 * n * EPS * np.linalg.cond(A0) * np.abs(ref).max()
