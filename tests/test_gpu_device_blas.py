"""GPU parity of the device-resident BLAS-3 entry points (mops_d_dgemm,
mops_d_dtrsm, mops_d_dsyrk): operands already in device memory, column
major, called through the C ABI on torch's current stream, compared with
numpy FP64 on the same seeded inputs (forward-error bounds of the
respective BLAS routine).
"""
import ctypes

import numpy as np
import pytest

import matops_amd as mo
from helpers import EPS, fmat, gemm_bound

pytestmark = pytest.mark.gpu


def _dev(X):
    import torch
    return torch.from_numpy(np.ascontiguousarray(X.T)).to("cuda:0").T  # column major


def _host(T):
    import torch
    torch.cuda.synchronize()
    return np.asfortranarray(T.cpu().numpy())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _op(X, t):
    return X.T if t else X


@pytest.mark.parametrize("flags", [mo.NOTRANS, mo.TRANSA, mo.TRANSB, mo.TRANSA | mo.TRANSB])
@pytest.mark.parametrize("m,n,k", [(33, 70, 19), (500, 300, 260), (64, 128, 3000)])
def test_d_dgemm(flags, m, n, k):
    from matops_amd import _lib
    ta, tb = bool(flags & mo.TRANSA), bool(flags & mo.TRANSB)
    A = fmat(*((k, m) if ta else (m, k)), seed=41)
    B = fmat(*((n, k) if tb else (k, n)), seed=42)
    C0 = fmat(m, n, seed=43)
    dA, dB, dC = _dev(A), _dev(B), _dev(C0)
    rc = _lib.lib().mops_d_dgemm(flags, m, n, k, -1.5, dA.data_ptr(), A.shape[0], dB.data_ptr(),
                                 B.shape[0], 0.25, dC.data_ptr(), m, _stream())
    assert rc == 0, _lib.lib().mops_last_error()
    C = _host(dC)
    ref = 0.25 * C0 - 1.5 * (_op(A, ta) @ _op(B, tb))
    bound = 1.5 * gemm_bound(np.abs(_op(A, ta)), np.abs(_op(B, tb)), k) \
        + 2 * EPS * np.abs(0.25 * C0)
    assert np.all(np.abs(C - ref) <= bound)


@pytest.mark.parametrize("flags", [mo.NOTRANS, mo.TRANSA, mo.TRANSB, mo.TRANSA | mo.TRANSB])
@pytest.mark.parametrize("k", [333, 32, 16])
def test_d_dgemm_big_tiles(flags, k):
    """Outputs of >= 512 tiles of 128x128 (the big-tile instance, two k-slabs
    in flight) with edge tiles in m and n and 21, 2 or 1 slabs of 16 in k."""
    from matops_amd import _lib
    m, n = 2900, 2950
    ta, tb = bool(flags & mo.TRANSA), bool(flags & mo.TRANSB)
    A = fmat(*((k, m) if ta else (m, k)), seed=44)
    B = fmat(*((n, k) if tb else (k, n)), seed=45)
    C0 = fmat(m, n, seed=46)
    dA, dB, dC = _dev(A), _dev(B), _dev(C0)
    rc = _lib.lib().mops_d_dgemm(flags, m, n, k, 0.75, dA.data_ptr(), A.shape[0], dB.data_ptr(),
                                 B.shape[0], -1.0, dC.data_ptr(), m, _stream())
    assert rc == 0, _lib.lib().mops_last_error()
    C = _host(dC)
    ref = 0.75 * (_op(A, ta) @ _op(B, tb)) - C0
    bound = 0.75 * gemm_bound(np.abs(_op(A, ta)), np.abs(_op(B, tb)), k) + 2 * EPS * np.abs(C0)
    assert np.all(np.abs(C - ref) <= bound)


@pytest.mark.parametrize("flags", [mo.NOTRANS, mo.TRANSA, mo.TRANSB, mo.TRANSA | mo.TRANSB])
@pytest.mark.parametrize("m,n,k", [(3000, 16, 4000), (2100, 5, 2500), (2048, 1, 2048)])
def test_d_dgemm_skinny(flags, m, n, k):
    """Outputs of at most 16 columns with a long k (the refinement's
    residual shape): the 128x16-tile instance with k split into chunks whose
    partial tiles are added in chunk order."""
    from matops_amd import _lib
    ta, tb = bool(flags & mo.TRANSA), bool(flags & mo.TRANSB)
    A = fmat(*((k, m) if ta else (m, k)), seed=47)
    B = fmat(*((n, k) if tb else (k, n)), seed=48)
    C0 = fmat(m, n, seed=49)
    dA, dB, dC = _dev(A), _dev(B), _dev(C0)
    rc = _lib.lib().mops_d_dgemm(flags, m, n, k, -1.0, dA.data_ptr(), A.shape[0], dB.data_ptr(),
                                 B.shape[0], 1.0, dC.data_ptr(), m, _stream())
    assert rc == 0, _lib.lib().mops_last_error()
    C = _host(dC)
    ref = C0 - _op(A, ta) @ _op(B, tb)
    bound = gemm_bound(np.abs(_op(A, ta)), np.abs(_op(B, tb)), k) + 2 * EPS * np.abs(C0)
    assert np.all(np.abs(C - ref) <= bound)


def test_d_dgemm_bad_arguments():
    from matops_amd import _lib
    L = _lib.lib()
    assert L.mops_d_dgemm(0, -1, 4, 4, 1.0, None, 1, None, 1, 0.0, None, 1, None) == -2
    assert L.mops_d_dgemm(0, 8, 4, 4, 1.0, None, 7, None, 4, 0.0, None, 8, None) == -7
    assert L.mops_d_dgemm(0, 8, 4, 4, 1.0, None, 8, None, 4, 0.0, None, 7, None) == -12


@pytest.mark.parametrize("flags", [mo.LOWER, mo.UPPER, mo.LOWER | mo.TRANSA,
                                   mo.UPPER | mo.TRANSA])
@pytest.mark.parametrize("n,k", [(200, 50), (700, 256), (96, 4096)])
def test_d_dsyrk(flags, n, k):
    """(96, 4096) has few output tiles and a long k: the split-k path."""
    from matops_amd import _lib
    ta = bool(flags & mo.TRANSA)
    A = fmat(*((k, n) if ta else (n, k)), seed=51)
    C0 = fmat(n, n, seed=52)
    dA, dC = _dev(A), _dev(C0)
    rc = _lib.lib().mops_d_dsyrk(flags, n, k, -1.0, dA.data_ptr(), A.shape[0], 1.0,
                                 dC.data_ptr(), n, _stream())
    assert rc == 0, _lib.lib().mops_last_error()
    C = _host(dC)
    Aop = A.T if ta else A
    ref = C0 - Aop @ Aop.T
    bound = k * EPS * (np.abs(Aop) @ np.abs(Aop.T)) + 2 * EPS * np.abs(C0)
    idx = np.tril_indices(n) if flags & mo.LOWER else np.triu_indices(n)
    other = np.triu_indices(n, 1) if flags & mo.LOWER else np.tril_indices(n, -1)
    assert np.all((np.abs(C - ref) <= bound)[idx])
    np.testing.assert_array_equal(C[other], C0[other])


@pytest.mark.parametrize("flags", [mo.LEFT | mo.LOWER, mo.LEFT | mo.UPPER | mo.TRANSA,
                                   mo.RIGHT | mo.LOWER | mo.TRANSA, mo.RIGHT | mo.UPPER | mo.UNIT])
@pytest.mark.parametrize("n,nrhs", [(300, 17), (1024, 64)])
def test_d_dtrsm(flags, n, nrhs):
    from matops_amd import _lib
    T = fmat(n, n, 61, "normal") * 0.3
    T[np.diag_indices(n)] = 2.0 + np.abs(fmat(n, 1, 62)[:, 0])
    T = np.asfortranarray(np.triu(T) if flags & mo.UPPER else np.tril(T))
    right = bool(flags & mo.RIGHT)
    B0 = fmat(*((nrhs, n) if right else (n, nrhs)), seed=63)
    dT, dB = _dev(T), _dev(B0)
    rc = _lib.lib().mops_d_dtrsm(flags, n, nrhs, 2.0, dT.data_ptr(), n, dB.data_ptr(),
                                 B0.shape[0], _stream())
    assert rc == 0, _lib.lib().mops_last_error()
    X = _host(dB)
    Te = T.copy()
    if flags & mo.UNIT:
        Te[np.diag_indices(n)] = 1.0
    Top = Te.T if flags & mo.TRANSA else Te
    # residual of the triangular system relative to its data
    R = (X @ Top - 2.0 * B0) if right else (Top @ X - 2.0 * B0)
    scale = (np.abs(X) @ np.abs(Top)) if right else (np.abs(Top) @ np.abs(X))
    assert np.all(np.abs(R) <= 4 * n * EPS * (scale + 2.0 * np.abs(B0)) + 1e-300)


def test_calls_on_two_streams_share_the_workspace():
    """Two GEMMs with few output tiles and a long k (their partial tiles
    live in the calling thread's workspace) queued back to back on two
    streams from one host thread, three times over: each call waits for
    the previous call's use of the workspace, so both results are right."""
    import torch
    from matops_amd import _lib
    L = _lib.lib()
    m, n, k = 64, 128, 6000
    A1, B1, A2, B2 = fmat(m, k, 71), fmat(k, n, 72), fmat(m, k, 73), fmat(k, n, 74)
    dA1, dB1, dA2, dB2 = _dev(A1), _dev(B1), _dev(A2), _dev(B2)
    C1 = torch.zeros((n, m), dtype=torch.float64, device="cuda:0").T
    C2 = torch.zeros((n, m), dtype=torch.float64, device="cuda:0").T
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for _ in range(3):
        assert L.mops_d_dgemm(0, m, n, k, 1.0, dA1.data_ptr(), m, dB1.data_ptr(), k, 0.0,
                              C1.data_ptr(), m, ctypes.c_void_p(s1.cuda_stream)) == 0
        assert L.mops_d_dgemm(0, m, n, k, 1.0, dA2.data_ptr(), m, dB2.data_ptr(), k, 0.0,
                              C2.data_ptr(), m, ctypes.c_void_p(s2.cuda_stream)) == 0
    torch.cuda.synchronize()
    for C, A, B in ((C1, A1, B1), (C2, A2, B2)):
        Ch = np.asfortranarray(C.cpu().numpy())
        assert np.all(np.abs(Ch - A @ B) <= gemm_bound(np.abs(A), np.abs(B), k))
