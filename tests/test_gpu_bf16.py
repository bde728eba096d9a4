"""GPU tests of the BF16 form of the mixed-precision LU solve.

* The BF16 product (mops_d_gemm_bf16): C += alpha * A * B with FP32
  operands rounded to BF16 (round to nearest even, A after its power-of-two
  scaling) equals the FP64 product of the rounded operands to FP32
  accumulation accuracy, K * 2^-24 * |A||B| componentwise. This also pins
  the lane layout of the BF16 matrix instruction: a wrong row, column or k
  assignment gives errors of the size of the product itself.
* SolveIR with PREC_BF16 converges to eta <= 1e-12 on a well-conditioned
  (diagonally dominant) system, with a first solve much less accurate than
  the FP32 factors give, and reports non-convergence honestly when the
  BF16 factors cannot contract the error (the N(0,1) system at n = 16384:
  classical refinement converges only when cond(A) * 2^-8 is well below 1).
"""
import numpy as np
import pytest

import matops_amd as mo
from helpers import fmat

pytestmark = pytest.mark.gpu


def _bf16_round(torch, X):
    return X.to(torch.bfloat16).to(torch.float64)


@pytest.mark.parametrize("m,n,k,scale", [(128, 128, 32, 1.0), (300, 257, 96, 1.0),
                                         (1000, 700, 256, 1e20), (77, 5, 33, 1e-20),
                                         (4096, 4096, 256, 3.0)])
def test_gemm_bf16_against_rounded_operands(m, n, k, scale):
    import ctypes
    import torch
    from matops_amd import _lib
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(m + n + k)
    A = (torch.randn((k, m), generator=g, device=dev, dtype=torch.float32) * scale).T
    B = torch.randn((n, k), generator=g, device=dev, dtype=torch.float32).T
    C0 = torch.randn((n, m), generator=g, device=dev, dtype=torch.float32).T
    C = C0.clone().T.contiguous().T
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = _lib.lib().mops_d_gemm_bf16(m, n, k, ctypes.c_float(-0.5), A.data_ptr(), m,
                                     B.data_ptr(), k, C.data_ptr(), m, s)
    assert rc == 0
    torch.cuda.synchronize()
    # the scaling of A is a power of two: rounding A * 2^-e equals rounding A
    Ar = _bf16_round(torch, A.double())
    Br = _bf16_round(torch, B.double())
    ref = C0.double() - 0.5 * (Ar @ Br)
    bound = 0.5 * (k + 2) * 2.0 ** -24 * (Ar.abs() @ Br.abs()) + 2.0 ** -24 * C0.double().abs()
    err = (C.double() - ref).abs()
    assert bool((err <= bound + 1e-300).all()), float((err / (bound + 1e-300)).max())


def test_bf16_refinement_converges_on_a_well_conditioned_system():
    n, nrhs = 2048, 4
    A = fmat(n, n, 3, "normal")
    A[np.diag_indices(n)] += 4.0 * np.sqrt(n)  # cond(A) ~ 3
    A = np.asfortranarray(A)
    B = fmat(n, nrhs, 4, "normal")
    X = np.zeros((n, nrhs), order="F")
    rep = mo.SolveIR(X, A, B, prec=mo.PREC_BF16, tol=1e-12, maxit=20)
    assert rep["converged"] == 1, rep
    R = B - A @ X
    eta = np.max(np.abs(R).max(axis=0) / (np.abs(A).sum(axis=1).max() * np.abs(X).max(axis=0)
                                         + np.abs(B).max(axis=0)))
    assert eta <= 1e-12 and rep["backward_error"] <= 1e-12
    X32 = np.zeros((n, nrhs), order="F")
    rep32 = mo.SolveIR(X32, A, B, prec=mo.PREC_FP32, tol=1e-12, maxit=20)
    # BF16 factors: a first solve far less accurate than with FP32 factors
    assert rep["first_error"] > 30 * rep32["first_error"], (rep, rep32)
    assert rep["iterations"] >= rep32["iterations"]


def test_bf16_refinement_reports_its_outcome_at_full_size():
    """n = 16384, 16 RHS, N(0,1) entries (the LU workload): whatever the
    outcome, the report agrees with the backward error recomputed from X."""
    import ctypes
    import torch
    from matops_amd import _lib
    n, nrhs = 16384, 16
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(5)
    A = torch.randn((n, n), generator=g, device=dev, dtype=torch.float64).T
    B = torch.randn((nrhs, n), generator=g, device=dev, dtype=torch.float64).T
    X = torch.zeros((nrhs, n), device=dev, dtype=torch.float64).T
    rep = _lib.IrReport()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    print("[bf16] n=16384 SolveIR BF16, maxit 8", flush=True)
    rc = _lib.lib().mops_d_dgesv_ir(n, nrhs, A.data_ptr(), n, B.data_ptr(), n, X.data_ptr(), n,
                                    mo.PREC_BF16, 1e-12, 8, ctypes.byref(rep), s)
    assert rc == 0
    torch.cuda.synchronize()
    R = B - A @ X
    an = A.abs().sum(dim=1).max()
    eta = (R.abs().amax(dim=0) / (an * X.abs().amax(dim=0) + B.abs().amax(dim=0))).max().item()
    print(f"[bf16] iterations {rep.iterations} converged {rep.converged} "
          f"eta first {rep.first_error:.3e} last {rep.backward_error:.3e} (check {eta:.3e})",
          flush=True)
    assert rep.iterations >= 1
    assert bool(rep.converged) == (rep.backward_error <= 1e-12)
    if np.isfinite(eta):
        assert abs(eta - rep.backward_error) <= 1e-3 * max(eta, rep.backward_error) + 1e-18
    else:
        assert not rep.converged
