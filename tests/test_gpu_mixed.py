"""GPU parity of the mixed precision LU + iterative refinement
(mops_dgesv_ir / mops_d_dgesv_ir): the solution reaches a normwise
backward error eta = max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf +
||b_j||_inf) <= 1e-12 (recomputed here in FP64 on the CPU), agrees with the
FP64 LAPACK solution to the accuracy the condition number allows, and the
report is honest about iterations and non-convergence.
"""
import ctypes

import numpy as np
import pytest

import matops_amd as mo

pytestmark = pytest.mark.gpu


def _eta(A, X, B):
    R = B - A @ X
    an = np.abs(A).sum(axis=1).max()
    return max(np.abs(R[:, j]).max() / (an * np.abs(X[:, j]).max() + np.abs(B[:, j]).max())
               for j in range(B.shape[1]))


def _problem(n, nrhs, seed):
    g = np.random.default_rng(seed)
    return np.asfortranarray(g.standard_normal((n, n))), np.asfortranarray(g.standard_normal((n, nrhs)))


@pytest.mark.parametrize("n,nrhs", [(1, 1), (100, 3), (1000, 16), (2500, 16)])
def test_fp32_refinement_reaches_1e12(n, nrhs):
    A, B = _problem(n, nrhs, 10 + n)
    A0, B0 = A.copy(), B.copy()
    X = np.zeros_like(B, order="F")
    rep = mo.SolveIR(X, A, B, mo.PREC_FP32, 1e-12, 20)
    assert np.array_equal(A, A0) and np.array_equal(B, B0)  # inputs untouched
    assert rep["converged"] == 1, rep
    assert rep["backward_error"] <= 1e-12
    assert _eta(A, X, B) <= 1.5e-12  # recomputed on the CPU
    assert 1 <= rep["iterations"] <= 10
    if n > 1:
        assert rep["first_error"] > rep["backward_error"]  # FP32 factors: refinement did work
    Xr = np.linalg.solve(A, B)
    fwd = np.abs(X - Xr).max() / np.abs(Xr).max()
    assert fwd <= 1e-16 * np.linalg.cond(A, np.inf) * n


def test_fp64_path_needs_one_solve():
    A, B = _problem(800, 4, 3)
    X = np.zeros_like(B, order="F")
    rep = mo.SolveIR(X, A, B, mo.PREC_FP64, 1e-12, 5)
    assert rep["converged"] == 1 and rep["iterations"] <= 2
    assert _eta(A, X, B) <= 1e-14


def test_ill_conditioned_reports_without_pretending():
    # cond(A) ~ 1e9: FP32 factors are too coarse for the refinement to
    # contract, and the report has to say so
    n = 300
    g = np.random.default_rng(5)
    U, _ = np.linalg.qr(g.standard_normal((n, n)))
    V, _ = np.linalg.qr(g.standard_normal((n, n)))
    A = np.asfortranarray((U * np.logspace(0, -9, n)) @ V.T)
    B = np.asfortranarray(g.standard_normal((n, 2)))
    X = np.zeros_like(B, order="F")
    rep = mo.SolveIR(X, A, B, mo.PREC_FP32, 1e-12, 6)
    assert abs(rep["backward_error"] - _eta(A, X, B)) <= 1e-3 * max(rep["backward_error"], 1e-300) + 1e-16
    if not rep["converged"]:
        assert rep["iterations"] == 6 and rep["backward_error"] > 1e-12


def test_unknown_precision_is_an_error():
    A, B = _problem(10, 1, 1)
    X = np.zeros_like(B, order="F")
    for prec in (3, -1):
        with pytest.raises(Exception):
            mo.SolveIR(X, A, B, prec, 1e-12, 5)


def test_device_api_full_size():
    """n = 16384, 16 RHS, N(0,1): converged to eta <= 1e-12 through the
    device entry point."""
    import torch
    from matops_amd import _lib
    dev = torch.device("cuda:0")
    n, nrhs = 16384, 16
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    A = torch.empty((n, n), dtype=torch.float64, device=dev).t()
    A.normal_(generator=g)
    B = torch.empty((nrhs, n), dtype=torch.float64, device=dev).t()
    B.normal_(generator=g)
    X = torch.zeros((nrhs, n), dtype=torch.float64, device=dev).t()
    rep = _lib.IrReport()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = _lib.lib().mops_d_dgesv_ir(n, nrhs, A.data_ptr(), n, B.data_ptr(), n, X.data_ptr(), n,
                                    mo.PREC_FP32, 1e-12, 10, ctypes.byref(rep), s)
    assert rc == 0, _lib.last_error()
    assert rep.converged == 1 and rep.backward_error <= 1e-12
    R = B - A @ X
    an = A.abs().sum(dim=1).max()
    eta = (R.abs().amax(dim=0) / (an * X.abs().amax(dim=0) + B.abs().amax(dim=0))).max().item()
    assert eta <= 1.5e-12
    print(f"n={n}: iterations {rep.iterations}, eta {rep.backward_error:.2e}, "
          f"factor {rep.factor_ms:.1f} ms, refine {rep.refine_ms:.1f} ms")
