"""Shared test helpers: seeded inputs and error measures."""
import numpy as np

EPS = np.finfo(np.float64).eps


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def fmat(rows, cols, seed, kind="uniform"):
    """Column-major (Fortran-ordered) float64 matrix from a seeded Philox
    stream: U[-1,1) or N(0,1)."""
    g = rng(seed)
    if kind == "uniform":
        a = g.uniform(-1.0, 1.0, size=(cols, rows)).T
    else:
        a = g.standard_normal(size=(cols, rows)).T
    return np.asfortranarray(a)


def spd(n, seed):
    """A = G G^T / n + I, exactly symmetric."""
    G = fmat(n, n, seed, "normal")
    A = G @ G.T / n + np.eye(n)
    A = np.tril(A) + np.tril(A, -1).T
    return np.asfortranarray(A)


def gemm_bound(A, B, k):
    """Componentwise k*eps*|A||B| bound."""
    return k * EPS * (np.abs(A) @ np.abs(B))


def embed(X, ld_extra=5, r0=3, c0=2, fill=np.nan):
    """X copied into a bigger Fortran array (ld > rows, odd offset); returns
    (parent, view)."""
    r, c = X.shape
    P = np.full((r + r0 + ld_extra, c + c0 + 3), fill, order="F")
    V = P[r0:r0 + r, c0:c0 + c]
    V[...] = X
    return P, V
