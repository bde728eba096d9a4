"""GPU tests of failure reporting.

* Bounded waits: with MOPS_SPIN_LIMIT=0 (read once per process) every
  bounded wait gives up at once instead of polling, so the give-up path runs
  without any kernel waiting. A split-k GEMM and a chained triangular solve
  must then report the timeout: the host-memory calls fail with a runtime
  error naming the wait, the device-resident ones through mops_d_status(),
  which clears the word. Run in a child process so the setting stays there.
* Refinement on a singular matrix: Inf/NaN in the solution must never read
  as converged.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import matops_amd as mo
from matops_amd import _lib
from helpers import fmat

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import ctypes, sys
sys.path.insert(0, sys.argv[1])
sys.path.insert(0, sys.argv[1] + "/tests")
import numpy as np
import torch
import matops_amd as mo
from matops_amd import _lib
from helpers import fmat
L = _lib.lib()
# host-memory split-k GEMM (one 64x64 output tile, k = 8192: split in chunks)
A, B = fmat(64, 8192, 1), fmat(8192, 64, 2)
C = np.zeros((64, 64), order="F")
try:
    mo.Mult(C, A, B, 1.0, 0.0, mo.NOTRANS)
    print("gemm-host: no error")
except _lib.MopsError as e:
    print("gemm-host:", e.code, "split-k" in e.msg)
print("status-after-host:", L.mops_d_status(None))
# device-resident split-k GEMM, then the status word
dev = torch.device("cuda:0")
tA = torch.from_numpy(A.T.copy()).to(dev).T
tB = torch.from_numpy(B.T.copy()).to(dev).T
tC = torch.zeros((64, 64), dtype=torch.float64, device=dev).T
s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
rc = L.mops_d_dgemm(0, 64, 64, 8192, 1.0, tA.data_ptr(), 64, tB.data_ptr(), 8192, 0.0,
                    tC.data_ptr(), 64, s)
print("gemm-dev-rc:", rc)
st = L.mops_d_status(s)
print("status1:", st, "split-k" in _lib.last_error())
print("status2:", L.mops_d_status(s))
# chained triangular solve (LEFT, few right-hand sides)
n = 512
T = np.asfortranarray(np.tril(fmat(n, n, 3, "normal") * 0.1) + 3.0 * np.eye(n))
X = fmat(n, 8, 4)
try:
    mo.SolveTrm(X, T, 1.0, mo.LEFT | mo.LOWER)
    print("trsm-host: no error")
except _lib.MopsError as e:
    print("trsm-host:", e.code, "triangular solve" in e.msg)
print("status-end:", L.mops_d_status(s))
"""


def test_bounded_waits_report_timeouts():
    env = dict(os.environ, MOPS_SPIN_LIMIT="0")
    print("[status] child with MOPS_SPIN_LIMIT=0", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], env=env, capture_output=True,
                       text=True, timeout=300)
    out = r.stdout
    assert r.returncode == 0, out + r.stderr
    got = dict(line.split(":", 1) for line in out.strip().splitlines() if ":" in line)
    assert got["gemm-host"].split() == ["-1000", "True"], out
    assert got["status-after-host"].strip() == "0", out
    assert got["gemm-dev-rc"].strip() == "0", out
    assert got["status1"].split() == ["-1000", "True"], out
    assert got["status2"].strip() == "0", out
    assert got["trsm-host"].split() == ["-1000", "True"], out
    assert got["status-end"].strip() == "0", out


def test_status_clear_after_normal_calls():
    A, B = fmat(64, 8192, 1), fmat(8192, 64, 2)
    C = np.zeros((64, 64), order="F")
    assert mo.Mult(C, A, B, 1.0, 0.0, mo.NOTRANS) is None
    assert np.allclose(C, A @ B, rtol=0, atol=1e-10)
    assert _lib.lib().mops_d_status(None) == 0


def test_refinement_singular_matrix_not_converged():
    n, nrhs = 300, 3
    A = fmat(n, n, 7, "normal")
    A[:, 17] = 0.0  # singular: a zero column
    B = fmat(n, nrhs, 8, "normal")
    for prec in (mo.PREC_FP64, mo.PREC_FP32):
        X = np.zeros((n, nrhs), order="F")
        rep = mo.SolveIR(X, A, B, prec=prec, tol=1e-12, maxit=3)
        assert rep["converged"] == 0, (prec, rep)
        assert not (rep["backward_error"] <= 1e-12), (prec, rep)


def test_dmult_alpha_zero_propagates_inf_like_the_reference():
    """alpha = 0 still forms the product (calgo/dmmat_gemm.c:297,413): an Inf
    in A gives NaN in C; beta = 0 clears what C held."""
    from oracle import oracle as O
    m, n, k = 40, 30, 20
    A, B = fmat(m, k, 1), fmat(k, n, 2)
    A[5, 3] = np.inf
    C = np.full((m, n), 7.0, order="F")
    mo.Mult(C, A, B, 0.0, 0.0, mo.NOTRANS)
    Co = np.full((m, n), 7.0, order="F")
    O.use("reference" if O.reference_available() else "port")
    O.Mult(Co, A, B, 0.0, 0.0, 0)
    O.use("port")
    np.testing.assert_array_equal(np.isnan(C), np.isnan(Co))
    assert np.isnan(C[5]).all() and not np.isnan(np.delete(C, 5, axis=0)).any()
    assert np.all(C[~np.isnan(C)] == 0.0)


def test_split_flags_reused_across_many_launches():
    """The chunk flags of split-k launches come from a per-thread pool, and
    each launch marks them with a value no earlier launch used (nothing is
    cleared between launches). 1100 launches that take 1024 flags each go
    around the pool of 2^20 flags once; every result must equal the first
    (the split adds its partial tiles in a fixed order), alternating with
    chained triangular solves that take their flags from the same pool."""
    import torch
    dev = torch.device("cuda:0")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = _lib.lib()
    g = torch.Generator(device=dev).manual_seed(3)
    m, k = 1000, 8192  # 16 x 16 tiles of 64, k split in 4 chunks
    A = torch.randn((k, m), generator=g, device=dev, dtype=torch.float64).T
    B = torch.randn((m, k), generator=g, device=dev, dtype=torch.float64).T
    C = torch.empty((m, m), device=dev, dtype=torch.float64).T
    ref = None
    n = 512
    T = torch.tril(torch.randn((n, n), generator=g, device=dev, dtype=torch.float64) * 0.1)
    T = (T + 3.0 * torch.eye(n, device=dev, dtype=torch.float64)).T.contiguous().T
    X0 = torch.randn((8, n), generator=g, device=dev, dtype=torch.float64).T
    X = torch.empty_like(X0.T).T
    xref = None
    print("[flags] 1100 split-k GEMMs 1000x1000x8192 with chained solves", flush=True)
    for it in range(1100):
        assert L.mops_d_dgemm(0, m, m, k, 1.0, A.data_ptr(), m, B.data_ptr(), k, 0.0,
                              C.data_ptr(), m, s) == 0
        if ref is None:
            ref = C.clone()
            assert torch.allclose(ref, A @ B, rtol=0, atol=1e-9)
        elif not torch.equal(C, ref):
            pytest.fail(f"launch {it}: result differs from the first")
        if it % 100 == 0:
            X.copy_(X0)
            assert L.mops_d_dtrsm(mo.LEFT | mo.LOWER, n, 8, 1.0, T.data_ptr(), n, X.data_ptr(),
                                  n, s) == 0
            if xref is None:
                xref = X.clone()
                assert torch.allclose(T @ xref, X0, rtol=0, atol=1e-12)
            else:
                assert torch.equal(X, xref), it
    torch.cuda.synchronize()
    assert L.mops_d_status(s) == 0
