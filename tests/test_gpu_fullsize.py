"""GPU parity at the workloads' full sizes, through the host-memory C ABI:

* DecomposeLU(A, piv, 256) + SolveLU at n = 16384, 16 RHS, A i.i.d. N(0,1)
  (seed 5): the pivot vector equals LAPACK's (tests/golden/fullsize_lu.npz,
  made by make_fullsize.py; the reference's pivots are LAPACK's ipiv - 1,
  test/perf_lu.go:105-115), ||PA - LU||_F / ||A||_F <= 10 n eps and the
  solve's backward error <= 10 n eps, both evaluated in FP64 on the GPU.
* DecomposeCHOL(A, LOWER, 256) + SolveCHOL at n = 8192, 64 RHS,
  A = G G^T / n + I: ||A - L L^T||_F / ||A||_F <= 10 n eps, the factor
  against LAPACK's dpotrf, the upper triangle untouched, the solve's
  backward error <= 10 n eps.
"""
import os
import sys

import numpy as np
import pytest

import matops_amd as mo
from helpers import EPS, fmat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import golden_inputs as gi  # noqa: E402

pytestmark = pytest.mark.gpu
FULL_LU = os.path.join(ROOT, "tests", "golden", "fullsize_lu.npz")


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch, torch.device("cuda:0")


def test_lu_16384_pivots_equal_lapack_and_residual():
    torch, dev = _torch()
    fx = np.load(FULL_LU)
    n = int(fx["n"])
    print(f"[C3] generating A {n}x{n} seed {int(fx['seed'])}", flush=True)
    A = gi.mat(n, n, int(fx["seed"]), "normal")
    assert gi.digest(A) == str(fx["a_sha"]), "regenerated input differs from the fixture's"
    F = A.copy(order="F")
    piv = np.zeros(n, dtype=np.int64)
    print("[C3] DecomposeLU", flush=True)
    _, err = mo.DecomposeLU(F, piv, 256)
    assert err is None
    ref = fx["piv"].astype(np.int64)
    bad = np.nonzero(piv != ref)[0]
    if bad.size:
        k = int(bad[0])
        pytest.fail(f"pivots differ first at column {k}: {piv[k]} vs LAPACK {ref[k]} "
                    f"({bad.size} columns differ); runner-up/pivot ratio there "
                    f"{float(fx['runner_up'][k]):.17g}")
    assert np.all(piv >= np.arange(n))
    # P A - L U on the GPU in FP64
    print("[C3] residual", flush=True)
    perm = np.arange(n)
    for k, p in enumerate(piv):
        perm[[k, p]] = perm[[p, k]]
    tA = torch.from_numpy(A.T).to(dev).T  # column-major on the device
    tF = torch.from_numpy(F.T).to(dev).T
    Lm = torch.tril(tF, -1) + torch.eye(n, dtype=torch.float64, device=dev)
    Um = torch.triu(tF)
    R = tA[torch.from_numpy(perm).to(dev)] - Lm @ Um
    res = (torch.linalg.norm(R) / torch.linalg.norm(tA)).item()
    del R, Lm, Um
    assert res <= 10 * n * EPS, res
    # SolveLU, 16 RHS
    B = fmat(n, 16, 6, "normal")
    X = B.copy(order="F")
    assert mo.SolveLU(X, F, piv, mo.NONE) is None
    tX = torch.from_numpy(X).to(dev)
    tB = torch.from_numpy(B).to(dev)
    eta = (torch.linalg.norm(tA @ tX - tB) /
           (torch.linalg.norm(tA) * torch.linalg.norm(tX) + torch.linalg.norm(tB))).item()
    assert eta <= 10 * n * EPS, eta
    print(f"[C3] pivots equal LAPACK's, residual {res:.3e}, eta {eta:.3e}", flush=True)


def test_chol_8192_residual_lapack_and_solve():
    import scipy.linalg as sla
    torch, dev = _torch()
    n, nb, nrhs = 8192, 256, 64
    print(f"[C2] generating A = G G^T / n + I, n = {n}", flush=True)
    G = torch.from_numpy(fmat(n, n, 3, "normal")).to(dev)
    tA = G @ G.T / n + torch.eye(n, dtype=torch.float64, device=dev)
    tA = torch.tril(tA) + torch.tril(tA, -1).T
    del G
    A0 = np.asfortranarray(tA.cpu().numpy())
    F = A0.copy(order="F")
    F[np.triu_indices(n, 1)] = -3.0  # must stay untouched
    print("[C2] DecomposeCHOL", flush=True)
    _, err = mo.DecomposeCHOL(F, mo.LOWER, nb)
    assert err is None
    assert np.all(F[np.triu_indices(n, 1)] == -3.0)
    L = np.tril(F)
    tL = torch.from_numpy(L).to(dev)
    res = (torch.linalg.norm(tA - tL @ tL.T) / torch.linalg.norm(tA)).item()
    assert res <= 10 * n * EPS, res
    print("[C2] LAPACK dpotrf", flush=True)
    Ll = sla.cholesky(A0, lower=True)
    d = np.abs(L - Ll).max() / np.abs(Ll).max()
    assert d <= 10 * n * EPS * 5.0, d  # cond(A) <= 5 for G G^T / n + I
    B = fmat(n, nrhs, 4, "normal")
    X = B.copy(order="F")
    mo.SolveCHOL(X, F, mo.LOWER)
    tX = torch.from_numpy(X).to(dev)
    tB = torch.from_numpy(B).to(dev)
    eta = (torch.linalg.norm(tA @ tX - tB) /
           (torch.linalg.norm(tA) * torch.linalg.norm(tX) + torch.linalg.norm(tB))).item()
    assert eta <= 10 * n * EPS, eta
    print(f"[C2] residual {res:.3e}, |L - L_lapack| {d:.3e}, eta {eta:.3e}", flush=True)
