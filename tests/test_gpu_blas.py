"""GPU parity of the BLAS-3 drop-ins (GEMM, TRSM, SYRK) and the Cholesky
path, called through the C ABI on host memory (the reference's calling
convention) and compared with numpy/scipy FP64 on the same seeded inputs.
"""
import numpy as np
import pytest
import scipy.linalg as sla

import matops_amd as mo
from helpers import EPS, embed, fmat, gemm_bound, spd

pytestmark = pytest.mark.gpu


def _op(X, t):
    return X.T if t else X


@pytest.mark.parametrize("flags", [mo.NOTRANS, mo.TRANSA, mo.TRANSB, mo.TRANSA | mo.TRANSB])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (64, 48, 40), (129, 65, 200), (300, 257, 33),
                                   (128, 64, 32), (256, 192, 48)])
def test_mult_flags(flags, m, n, k):
    ta, tb = bool(flags & mo.TRANSA), bool(flags & mo.TRANSB)
    A = fmat(*((k, m) if ta else (m, k)), seed=11)
    B = fmat(*((n, k) if tb else (k, n)), seed=12)
    C0 = fmat(m, n, seed=13)
    for alpha, beta in [(1.0, 0.0), (-1.0, 1.0), (0.5, -2.0)]:
        C = C0.copy(order="F")
        assert mo.Mult(C, A, B, alpha, beta, flags) is None
        ref = beta * C0 + alpha * (_op(A, ta) @ _op(B, tb))
        bound = abs(alpha) * gemm_bound(np.abs(_op(A, ta)), np.abs(_op(B, tb)), k) \
            + 2 * EPS * np.abs(beta * C0) + 1e-300
        assert np.all(np.abs(C - ref) <= bound), (alpha, beta)


def test_mult_beta_zero_clears_nan_and_views():
    m, n, k = 70, 90, 50
    A, B = fmat(m, k, 1), fmat(k, n, 2)
    Cp, Cv = embed(np.full((m, n), np.nan))
    Ap, Av = embed(A, 7, 1, 4)
    Bp, Bv = embed(B, 3, 5, 1)
    mo.Mult(Cv, Av, Bv, 1.0, 0.0, mo.NOTRANS)
    assert np.all(np.isfinite(Cv))
    np.testing.assert_allclose(Cv, A @ B, rtol=0, atol=k * 4 * EPS * 50)
    # parent outside the view untouched
    mask = np.ones_like(Cp, dtype=bool)
    mask[3:3 + m, 2:2 + n] = False
    assert np.all(np.isnan(Cp[mask]))


def test_mult_size_mismatch_and_empty():
    A, B, C = fmat(4, 3, 1), fmat(4, 5, 2), fmat(4, 5, 3)
    err = mo.Mult(C, A, B, 1.0, 0.0, mo.NOTRANS)
    assert isinstance(err, mo.MatopsError) and str(err) == "Mult: size mismatch"
    E = np.zeros((0, 3), order="F")
    assert mo.Mult(np.zeros((0, 5), order="F"), E, fmat(3, 5, 4), 1.0, 0.0, 0) is None


@pytest.mark.parametrize("flags", [mo.NOTRANS, mo.TRANSA, mo.TRANSB, mo.TRANSA | mo.TRANSB])
@pytest.mark.parametrize("m,n,k", [(64, 64, 4000), (100, 300, 5000), (1, 1, 2048), (700, 65, 777),
                                   (1000, 520, 1024), (2048, 16, 2048), (2048, 13, 2050)])
def test_mult_long_k(flags, m, n, k):
    """Few output tiles and a long inner dimension: k is split across
    workgroups and the partial tiles are added in a fixed order (twice the
    same call gives the same bits)."""
    ta, tb = bool(flags & mo.TRANSA), bool(flags & mo.TRANSB)
    A = fmat(*((k, m) if ta else (m, k)), seed=21)
    B = fmat(*((n, k) if tb else (k, n)), seed=22)
    C0 = fmat(m, n, seed=23)
    outs = []
    for _ in range(2):
        C = C0.copy(order="F")
        assert mo.Mult(C, A, B, -1.0, 0.5, flags) is None
        outs.append(C)
    ref = 0.5 * C0 - _op(A, ta) @ _op(B, tb)
    bound = gemm_bound(np.abs(_op(A, ta)), np.abs(_op(B, tb)), k) + 2 * EPS * np.abs(0.5 * C0)
    assert np.all(np.abs(outs[0] - ref) <= bound)
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("flags", [mo.LOWER, mo.UPPER, mo.LOWER | mo.TRANSA])
def test_rankupdatesym_long_k(flags):
    n, k = 64, 4000
    A = fmat(*((k, n) if flags & mo.TRANSA else (n, k)), seed=31)
    C0 = fmat(n, n, seed=32)
    C = C0.copy(order="F")
    mo.RankUpdateSym(C, A, -1.0, 1.0, flags)
    AA = A.T @ A if flags & mo.TRANSA else A @ A.T
    ref = C0 - AA
    idx = np.tril_indices(n) if flags & mo.LOWER else np.triu_indices(n)
    other = np.triu_indices(n, 1) if flags & mo.LOWER else np.tril_indices(n, -1)
    bound = k * EPS * (np.abs(A.T) @ np.abs(A) if flags & mo.TRANSA else np.abs(A) @ np.abs(A.T)) \
        + 2 * EPS * np.abs(C0)
    assert np.all((np.abs(C - ref) <= bound)[idx])
    np.testing.assert_array_equal(C[other], C0[other])


@pytest.mark.parametrize("flags", [mo.NOTRANS, mo.TRANSA, mo.TRANSB, mo.TRANSA | mo.TRANSB])
@pytest.mark.parametrize("m,n,k", [(2944, 2944, 48), (2945, 2900, 50)])
def test_mult_128_tiles(flags, m, n, k):
    """Outputs of 512 or more 128x128 tiles: the 128x128 instance, with
    (2944 = 23 x 128, k = 48) and without bounds checks compiled in."""
    ta, tb = bool(flags & mo.TRANSA), bool(flags & mo.TRANSB)
    A = fmat(*((k, m) if ta else (m, k)), seed=41)
    B = fmat(*((n, k) if tb else (k, n)), seed=42)
    C0 = fmat(m, n, seed=43)
    C = C0.copy(order="F")
    assert mo.Mult(C, A, B, 1.5, -0.5, flags) is None
    ref = -0.5 * C0 + 1.5 * (_op(A, ta) @ _op(B, tb))
    bound = 1.5 * gemm_bound(np.abs(_op(A, ta)), np.abs(_op(B, tb)), k) + 2 * EPS * np.abs(0.5 * C0)
    assert np.all(np.abs(C - ref) <= bound)


def test_mult_dgemm_1024():
    n = 1024
    A, B = fmat(n, n, 1), fmat(n, n, 2)
    C = np.full((n, n), np.nan, order="F")
    mo.Mult(C, A, B, 1.0, 0.0, mo.NOTRANS)
    ref = A @ B
    assert np.all(np.abs(C - ref) <= gemm_bound(A, B, n))


TRSM_FLAGS = [
    side | uplo | t | u
    for side in (mo.LEFT, mo.RIGHT)
    for uplo in (mo.LOWER, mo.UPPER)
    for t in (0, mo.TRANSA)
    for u in (0, mo.UNIT)
]


def _tri(n, flags, seed):
    T = fmat(n, n, seed, "normal") * 0.3
    T[np.diag_indices(n)] = 2.0 + np.abs(fmat(n, 1, seed + 1)[:, 0])
    if flags & mo.UPPER:
        T = np.triu(T)
    else:
        T = np.tril(T)
    # garbage in the other triangle must never be read
    G = np.asfortranarray(T + (np.triu(np.full((n, n), 7e300), 1) if not flags & mo.UPPER
                               else np.tril(np.full((n, n), 7e300), -1)))
    return np.asfortranarray(T), G


@pytest.mark.parametrize("flags", TRSM_FLAGS)
@pytest.mark.parametrize("n,nrhs", [(7, 3), (96, 40), (700, 37), (700, 100), (1000, 130)])
def test_solvetrm_all_variants(flags, n, nrhs):
    T, G = _tri(n, flags, 5)
    if flags & mo.UNIT:
        Teff = T.copy()
        Teff[np.diag_indices(n)] = 1.0
        G = G.copy(order="F")
        G[np.diag_indices(n)] = np.nan  # diagonal must not be read
    else:
        Teff = T
    opT = Teff.T if flags & mo.TRANSA else Teff
    right = bool(flags & mo.RIGHT)
    B0 = fmat(*((nrhs, n) if right else (n, nrhs)), seed=6)
    B = B0.copy(order="F")
    alpha = 1.5
    assert mo.SolveTrm(B, G, alpha, flags) is None
    X = B
    res = (X @ opT if right else opT @ X) - alpha * B0
    scale = (np.abs(X) @ np.abs(opT) if right else np.abs(opT) @ np.abs(X)) + np.abs(alpha * B0)
    assert np.max(np.abs(res) / scale) < 50 * n * EPS


@pytest.mark.parametrize("flags", [mo.LOWER, mo.UPPER, mo.LOWER | mo.TRANSA, mo.UPPER | mo.TRANSA])
@pytest.mark.parametrize("n,k", [(1, 1), (100, 70), (515, 256), (512, 256), (192, 64)])
def test_rankupdatesym(flags, n, k):
    ta = bool(flags & mo.TRANSA)
    A = fmat(*((k, n) if ta else (n, k)), seed=8)
    C0 = fmat(n, n, seed=9)
    C = C0.copy(order="F")
    mo.RankUpdateSym(C, A, -1.0, 1.0, flags)
    AA = A.T @ A if ta else A @ A.T
    ref = C0 - AA
    lower = bool(flags & mo.LOWER)
    idx = np.tril_indices(n) if lower else np.triu_indices(n)
    other = np.triu_indices(n, 1) if lower else np.tril_indices(n, -1)
    bound = k * EPS * (np.abs(A.T) @ np.abs(A) if ta else np.abs(A) @ np.abs(A.T)) \
        + 2 * EPS * np.abs(C0)
    assert np.all((np.abs(C - ref) <= bound)[idx])
    # the other strict triangle is untouched
    np.testing.assert_array_equal(C[other], C0[other])


@pytest.mark.parametrize("uplo", [mo.LOWER, mo.UPPER])
@pytest.mark.parametrize("n,nb", [(7, 0), (60, 16), (300, 68), (513, 256), (1024, 256), (700, 0),
                                  (1000, 200), (2500, 256), (1536, 384), (1300, 1000), (2100, 0)])
def test_chol_factor_and_solve(uplo, n, nb):
    A0 = spd(n, 3)
    A = A0.copy(order="F")
    if uplo == mo.LOWER:
        A[np.triu_indices(n, 1)] = -9.0  # must stay untouched
    else:
        A[np.tril_indices(n, -1)] = -9.0
    R, err = mo.DecomposeCHOL(A, uplo, nb)
    assert err is None
    if uplo == mo.LOWER:
        L = np.tril(R)
        np.testing.assert_array_equal(R[np.triu_indices(n, 1)], -9.0)
        res = np.linalg.norm(A0 - L @ L.T) / np.linalg.norm(A0)
        ref = sla.cholesky(A0, lower=True)
        F = L
    else:
        U = np.triu(R)
        np.testing.assert_array_equal(R[np.tril_indices(n, -1)], -9.0)
        res = np.linalg.norm(A0 - U.T @ U) / np.linalg.norm(A0)
        ref = sla.cholesky(A0, lower=False)
        F = U
    assert res <= 10 * n * EPS
    np.testing.assert_allclose(F, ref, rtol=0, atol=1e-10 * np.abs(ref).max())
    B0 = fmat(n, 5, 4, "normal")
    X = B0.copy(order="F")
    mo.SolveCHOL(X, R, uplo)
    eta = np.linalg.norm(A0 @ X - B0) / (np.linalg.norm(A0) * np.linalg.norm(X) + np.linalg.norm(B0))
    assert eta < 10 * n * EPS


def test_chol_kat_3x3():
    # decomp_test.go:150 _TestCHOL3x3: L=[[3,0,0],[6,4,0],[4,6,3]], A = L L^T
    L = np.array([[3.0, 0, 0], [6, 4, 0], [4, 6, 3]])
    A = np.asfortranarray(L @ L.T)
    mo.DecomposeCHOL(A, mo.LOWER, 0)
    np.testing.assert_array_equal(np.tril(A), L)


def test_chol_not_spd_panics_with_index():
    n = 200
    A = spd(n, 3)
    A[150, 150] = -1e6
    with pytest.raises(RuntimeError, match=r"illegal value at 150: "):
        mo.DecomposeCHOL(A, mo.LOWER, 64)


@pytest.mark.parametrize("uplo", [mo.LOWER, mo.UPPER])
def test_chol_not_spd_in_a_512_column_step(uplo):
    # blocks of 256 run as 512-column panels: the first bad diagonal (index
    # 777, in the second panel's fifth 64-column block) is the one reported
    n = 1200
    A = spd(n, 3)
    A[777, 777] = -1e6
    A = np.asfortranarray(A)
    with pytest.raises(RuntimeError, match=r"illegal value at 777: "):
        mo.DecomposeCHOL(A.copy(order="F"), uplo, 256)
