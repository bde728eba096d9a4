"""torch.autograd binding of the GPU Cholesky and its hand-written adjoint.

For checking the adjoint kernel against torch's own machinery (gradcheck,
finite differences); the product entry points are the C ABI functions
mops_d_dpotrf and mops_d_dpotrf_adjoint, which this wraps without adding
any computation of its own.
"""
import ctypes

import torch

from . import _lib

LOWER = 0x4


def _colmajor_copy(X):
    """A column-major (Fortran-ordered) float64 copy of a 2-D cuda tensor."""
    n, m = X.shape
    Y = torch.empty((m, n), dtype=torch.float64, device=X.device).t()
    Y.copy_(X)
    return Y


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class CholeskyFunction(torch.autograd.Function):
    """A (symmetric, lower triangle read) -> L = chol(A), lower.

    Forward: DecomposeCHOL(A, LOWER, nb) on the GPU (mops_d_dpotrf).
    Backward: the closed-form adjoint (mops_d_dpotrf_adjoint); the
    gradient returned is the symmetric d loss / dA."""

    @staticmethod
    def forward(ctx, A, nb=256):
        if A.dtype != torch.float64 or not A.is_cuda or A.dim() != 2 or A.shape[0] != A.shape[1]:
            raise TypeError("CholeskyFunction takes a square float64 cuda tensor")
        n = A.shape[0]
        L = _colmajor_copy(A.detach())
        info = torch.full((1,), -1, dtype=torch.int32, device=A.device)
        rc = _lib.lib().mops_d_dpotrf(LOWER, n, nb, L.data_ptr(), n, info.data_ptr(), _stream())
        _lib.check("mops_d_dpotrf", rc)
        bad = int(info.item())
        if bad >= 0:
            raise RuntimeError(f"illegal value at {bad}")
        L = torch.tril(L)
        ctx.save_for_backward(L)
        return L

    @staticmethod
    def backward(ctx, gL):
        (L,) = ctx.saved_tensors
        n = L.shape[0]
        Lc = _colmajor_copy(L)
        Ab = _colmajor_copy(torch.tril(gL))
        rc = _lib.lib().mops_d_dpotrf_adjoint(n, 0, Lc.data_ptr(), n, Ab.data_ptr(), n,
                                              _stream())
        _lib.check("mops_d_dpotrf_adjoint", rc)
        return Ab.contiguous(), None


def cholesky(A, nb=256):
    """L = chol(A) with the GPU adjoint in the backward pass."""
    return CholeskyFunction.apply(A, nb)
