"""GPU test of mops_d_dgemm_noalloc, the GEMM entry point that allocates
nothing: through the C ABI its output equals mops_d_dgemm's bit for bit
(the same kernels and the same split plan) on two calls in a row, the
second one smaller, for split-k, skinny split-k, large-tile and K = 0
products; and in a fresh process the device's free memory does not change
across a series of calls once the library's code is loaded.
"""
import ctypes
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu


def _setup():
    import torch
    from matops_amd import _lib
    dev = torch.device("cuda:0")
    return torch, _lib.lib(), dev


def _mat(torch, dev, r, c, seed):
    g = torch.Generator(device=dev).manual_seed(seed)
    return torch.empty((c, r), dtype=torch.float64, device=dev).uniform_(-1, 1, generator=g).T


# (flags, M, N, K) pairs: the second call of each pair is smaller
PAIRS = [
    ((0, 64, 64, 8192), (0, 40, 33, 3000)),          # split k, 64x64 tiles
    ((3, 100, 300, 5000), (1, 64, 64, 700)),         # split k, transposes
    ((0, 19585, 16, 2048), (0, 4000, 8, 2048)),      # skinny split k (128x16 tiles)
    ((0, 1024, 1024, 1024), (2, 512, 300, 200)),     # 128x128 and 64x64 tiles, no split
    ((0, 300, 200, 0), (0, 100, 50, 0)),             # K = 0: C = beta C
]


@pytest.mark.parametrize("first,second", PAIRS)
def test_noalloc_equals_dgemm_twice(first, second):
    torch, L, dev = _setup()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for flags, M, N, K in (first, second):
        print(f"[noalloc] flags={flags} M={M} N={N} K={K}", flush=True)
        ta, tb = flags & 1, flags & 2
        A = _mat(torch, dev, *((K, M) if ta else (M, K)), seed=M + 1)
        B = _mat(torch, dev, *((N, K) if tb else (K, N)), seed=N + 2)
        C0 = _mat(torch, dev, M, N, seed=K + 3)
        lda, ldb = max(A.shape[0], 1), max(B.shape[0], 1)
        C1, C2 = C0.clone(), C0.clone()
        C1 = C1.T.contiguous().T
        C2 = C2.T.contiguous().T
        rc = L.mops_d_dgemm(flags, M, N, K, -1.25, A.data_ptr(), lda, B.data_ptr(), ldb, 0.5,
                            C1.data_ptr(), M, s)
        assert rc == 0, rc
        rc = L.mops_d_dgemm_noalloc(flags, M, N, K, -1.25, A.data_ptr(), lda, B.data_ptr(), ldb,
                                    0.5, C2.data_ptr(), M, s)
        assert rc == 0, rc
        torch.cuda.synchronize()
        assert L.mops_d_status(s) == 0
        assert torch.equal(C1, C2), (flags, M, N, K)
        opA = A.T if ta else A
        opB = B.T if tb else B
        ref = 0.5 * C0 - 1.25 * (opA @ opB)
        bound = 1.25 * 2 * max(K, 1) * 2.0 ** -52 * (opA.abs() @ opB.abs()) \
            + 2.0 ** -52 * C0.abs()
        assert bool(((C2 - ref).abs() <= bound).all())


CHILD = r"""
import ctypes, sys
import torch
sys.path.insert(0, sys.argv[1])
from matops_amd import _lib
L = _lib.lib()
dev = torch.device("cuda:0")
s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
def mat(r, c):
    return torch.rand((c, r), dtype=torch.float64, device=dev).T
shapes = [(64, 64, 8192), (19585, 16, 2048), (1024, 1024, 1024), (300, 200, 0), (40, 33, 3000)]
ops = {sh: (mat(sh[0], sh[2]), mat(sh[2], sh[1]), mat(sh[0], sh[1])) for sh in shapes}
torch.cuda.synchronize()
def run():
    for (M, N, K), (A, B, C) in ops.items():
        rc = L.mops_d_dgemm_noalloc(0, M, N, K, 1.0, A.data_ptr(), max(M, 1), B.data_ptr(),
                                    max(K, 1), 0.0, C.data_ptr(), M, s)
        assert rc == 0, rc
    torch.cuda.synchronize()
run()  # loads the library's code objects
f0 = torch.cuda.mem_get_info()[0]
for _ in range(3):
    run()
f1 = torch.cuda.mem_get_info()[0]
print("free-before", f0, "free-after", f1, "status", L.mops_d_status(s))
"""


def test_noalloc_allocates_nothing_in_a_fresh_process():
    print("[noalloc] child process", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("free-before")][-1].split()
    f0, f1, st = int(line[1]), int(line[3]), int(line[5])
    assert st == 0
    assert f0 == f1, (f0, f1)
