"""GPU parity of the Cholesky adjoint (mops_dpotrf_adjoint /
mops_d_dpotrf_adjoint; the GPU evaluates the closed form in place by a
blocked TRMM and two blocked TRSMs, the oracle with its own GEMM and two
triangular solves): against the oracle (block orders 64 to 1024, with
partial last blocks), the A = I special case, central finite differences
through the Cholesky (h = 1e-6, 20 symmetric directions, unit Frobenius
norm), torch's own Cholesky backward, and at the full n = 4096 workload
(A = Q diag(logspace(0, 4)) Q^T, W lower N(0,1)) by finite differences.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import matops_amd as mo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import golden_inputs as gi  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


def _spd(n, seed):
    G = gi.mat(n, n, seed, "normal")
    return G @ G.T / n + np.eye(n)


def _sym_dirs(n, count, seed):
    g = np.random.default_rng(seed)
    for _ in range(count):
        Z = g.standard_normal((n, n))
        E = (Z + Z.T) / 2
        yield E / np.linalg.norm(E)


@pytest.mark.parametrize("n,nb", [(1, 0), (5, 0), (64, 64), (100, 32), (200, 64), (333, 0),
                                  (700, 256), (1030, 0), (1100, 1024)])
def test_adjoint_vs_oracle_closed_form(n, nb):
    A = _spd(n, 40 + n)
    L = np.linalg.cholesky(A)
    W = np.tril(gi.mat(n, n, 41 + n, "normal"))
    ref = O.chol_adjoint(L, W)
    Ab = np.asfortranarray(W.copy())
    Ab[np.triu_indices(n, 1)] = 7.0  # the upper triangle of Lbar is ignored
    mo.CholeskyAdjoint(np.asfortranarray(L), Ab, nb)
    np.testing.assert_allclose(Ab, ref, rtol=0, atol=1e-11 * max(1.0, np.abs(ref).max()))
    assert np.array_equal(Ab, Ab.T)


def test_adjoint_identity():
    # A = I: L = I and Abar = 1/2 sym(Phi(Lbar)), Phi = lower, halved diagonal
    n = 70
    W = np.tril(gi.mat(n, n, 5, "normal"))
    Ab = np.asfortranarray(W.copy())
    mo.CholeskyAdjoint(np.asfortranarray(np.eye(n)), Ab, 64)
    Phi = W - 0.5 * np.diag(np.diag(W))
    np.testing.assert_allclose(Ab, 0.5 * (Phi + Phi.T), atol=1e-15)


def _gpu_chol(A_dev, nb=256):
    import torch
    from matops_amd import _lib
    n = A_dev.shape[0]
    L = torch.empty((n, n), dtype=torch.float64, device=A_dev.device).t()
    L.copy_(A_dev)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert _lib.lib().mops_d_dpotrf(mo.LOWER, n, nb, L.data_ptr(), n, None, s) == 0
    return torch.tril(L)


def _fd_check(A, W, Ab, n_dirs, seed, rtol):
    """Central differences of f(A) = sum(W * chol(A)) along symmetric E of
    unit Frobenius norm. The error of each directional derivative is
    measured against ||Abar||_F, the largest value <Abar, E> can take: the
    quotient's rounding noise (~eps*|f|/h) does not shrink with <Abar, E>."""
    import torch
    dev = torch.device("cuda:0")
    n = A.shape[0]
    Ad = torch.from_numpy(A).to(dev)
    Wd = torch.from_numpy(W).to(dev)
    Abd = torch.from_numpy(Ab).to(dev)
    h = 1e-6
    scale = max(1.0, float(np.linalg.norm(Ab)))
    worst = 0.0
    for E in _sym_dirs(n, n_dirs, seed):
        Ed = torch.from_numpy(E).to(dev)
        fp = torch.sum(Wd * _gpu_chol(Ad + h * Ed)).item()
        fm = torch.sum(Wd * _gpu_chol(Ad - h * Ed)).item()
        fd = (fp - fm) / (2 * h)
        an = torch.sum(Abd * Ed).item()
        worst = max(worst, abs(fd - an) / scale)
    print(f"finite differences, n={n}: worst |fd - <Abar,E>| / ||Abar||_F = {worst:.2e}")
    assert worst < rtol, worst


def test_adjoint_finite_differences():
    n = 256
    A = _spd(n, 77)
    L = np.linalg.cholesky(A)
    W = np.tril(gi.mat(n, n, 78, "normal"))
    Ab = np.asfortranarray(W.copy())
    mo.CholeskyAdjoint(np.asfortranarray(L), Ab, 64)
    _fd_check(A, W, np.ascontiguousarray(Ab), 20, 9, 1e-6)


def test_autograd_wrapper_matches_torch_cholesky_backward():
    import torch
    from matops_amd.autograd import cholesky
    dev = torch.device("cuda:0")
    n = 300
    A = torch.from_numpy(_spd(n, 90)).to(dev)
    W = torch.from_numpy(np.tril(gi.mat(n, n, 91, "normal"))).to(dev)
    A1 = A.clone().requires_grad_(True)
    (W * cholesky(A1)).sum().backward()
    # torch's reference: float64, computed on the CPU
    A2 = A.cpu().clone().requires_grad_(True)
    (W.cpu() * torch.linalg.cholesky(A2)).sum().backward()
    g1, g2 = A1.grad.cpu().numpy(), A2.grad.numpy()
    np.testing.assert_allclose(g1, g2, rtol=0, atol=1e-10 * max(1.0, np.abs(g2).max()))


def test_adjoint_full_size_logspace_spectrum():
    """n = 4096, A = Q diag(logspace(0, 4)) Q^T, W lower N(0,1): 20 FD
    directions through the GPU Cholesky, and symmetry of the result."""
    import torch
    dev = torch.device("cuda:0")
    n = 4096
    g = torch.Generator(device="cpu").manual_seed(7)
    X = torch.randn((n, n), dtype=torch.float64, generator=g).to(dev)
    Q, _ = torch.linalg.qr(X)
    sig = torch.logspace(0, 4, n, dtype=torch.float64, device=dev)
    A = (Q * sig) @ Q.T
    A = 0.5 * (A + A.T)
    W = torch.tril(torch.randn((n, n), dtype=torch.float64, generator=g)).to(dev)
    L = _gpu_chol(A)
    Ab = torch.empty((n, n), dtype=torch.float64, device=dev).t()
    Ab.copy_(W)
    Lc = torch.empty((n, n), dtype=torch.float64, device=dev).t()
    Lc.copy_(L)
    from matops_amd import _lib
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert _lib.lib().mops_d_dpotrf_adjoint(n, 64, Lc.data_ptr(), n, Ab.data_ptr(), n, s) == 0
    torch.cuda.synchronize()
    assert torch.equal(Ab, Ab.T)
    assert torch.isfinite(Ab).all()
    _fd_check(A.cpu().numpy(), W.cpu().numpy(), Ab.contiguous().cpu().numpy(), 20, 10, 1e-6)
