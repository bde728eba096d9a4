"""Seeded inputs of the golden cases, shared by make_golden.py and the
tests (CPU oracle checks and GPU parity checks)."""
import hashlib

import numpy as np

GEMM_FLAGS = (0, 1, 2, 3)
GEMM_AB = ((1.0, 0.0), (-1.0, 1.0))
TRSM_FLAGS = tuple(side | uplo | t | u for side in (0x10, 0x20) for uplo in (0x4, 0x8)
                   for t in (0, 0x1) for u in (0, 0x40))
TRSM_N, TRSM_NRHS, TRSM_ALPHA = 48, 20, 1.5
SYRK_FLAGS = (0x4, 0x8, 0x5, 0x9)
CHOL_CASES = ((7, 0), (64, 16), (150, 68))
LU_CASES = ((7, 0), (64, 12), (300, 68))


def mat(r, c, seed, kind="uniform"):
    g = np.random.Generator(np.random.Philox(seed))
    a = g.uniform(-1, 1, (c, r)).T if kind == "uniform" else g.standard_normal((c, r)).T
    return np.asfortranarray(a)


def digest(a):
    """SHA-1 of the column-major bytes (pins regenerated inputs)."""
    return str(hashlib.sha1(np.asfortranarray(a).tobytes(order="F")).hexdigest())


def gemm_case(flags):
    ta, tb = flags & 1, flags & 2
    A = mat(40, 64, 11) if ta else mat(64, 40, 11)
    B = mat(48, 40, 12) if tb else mat(40, 48, 12)
    return A, B, mat(64, 48, 13)


def trsm_case(flags):
    n = TRSM_N
    T = mat(n, n, 21, "normal") * 0.3
    T[np.diag_indices(n)] = 2.0 + np.abs(mat(n, 1, 22)[:, 0])
    T = np.asfortranarray(np.tril(T) if flags & 0x4 else np.triu(T))
    B0 = mat(TRSM_NRHS, n, 23) if flags & 0x20 else mat(n, TRSM_NRHS, 23)
    return T, B0


def syrk_case(flags):
    A = mat(40, 60, 31) if flags & 1 else mat(60, 40, 31)
    return A, mat(60, 60, 32)


def chol_case(n):
    G = mat(n, n, 41, "normal")
    S = G @ G.T / n + np.eye(n)
    return np.asfortranarray(np.tril(S) + np.tril(S, -1).T), mat(n, 5, 42, "normal")


def lu_case(n):
    return mat(n, n, 51, "normal"), mat(n, 4, 52, "normal")


def ties_case():
    """Equal |candidates| in exact arithmetic in columns 0 and 1."""
    return np.asfortranarray(np.array([
        [1.0, 2.0, 0.0, 1.0],
        [3.0, 1.0, 2.0, 0.0],
        [-3.0, 1.0, 1.0, 2.0],
        [3.0, -1.0, 0.0, 1.0]]))
