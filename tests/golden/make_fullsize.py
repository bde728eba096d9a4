"""Writes tests/golden/fullsize_lu.npz: LAPACK's pivots of the LU workload
at its full size, n = 16384, A of i.i.d. N(0,1) entries from the seeded
generator of golden_inputs.py (seed 5).

The reference's pivots are LAPACK's ipiv - 1 (the reference's own check,
test/perf_lu.go:105-115, compares them entry by entry), so dgetrf's ipiv is
the expected pivot vector of DecomposeLU at this size. Also kept, per
column k, the largest |L[i, k]| below the diagonal of LAPACK's factor: the
ratio of the runner-up candidate to the pivot, which tells how close a
column's pivot choice was when a comparison fails. And the SHA-1 of the
input, so the GPU test knows it regenerated the same matrix.

Run from the repository root:  python tests/golden/make_fullsize.py
(about a minute on 8 cores; needs ~6 GB of memory).
"""
import os
import sys
import time

import numpy as np
from scipy.linalg import lapack

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import golden_inputs as gi  # noqa: E402

OUT = os.path.join(HERE, "fullsize_lu.npz")
N, SEED = 16384, 5


def main():
    t0 = time.time()
    A = gi.mat(N, N, SEED, "normal")
    sha = gi.digest(A)
    lu, ipiv, info = lapack.dgetrf(A, overwrite_a=True)
    assert info == 0
    del A
    # runner-up ratio per column: max_{i>k} |L[i,k]| (the pivot's is 1)
    second = np.zeros(N, dtype=np.float32)
    for k in range(N - 1):
        second[k] = np.abs(lu[k + 1:, k]).max()
    np.savez_compressed(OUT, n=N, seed=SEED, a_sha=np.array(sha),
                        piv=ipiv.astype(np.int32), runner_up=second)
    print(OUT, os.path.getsize(OUT), "bytes", f"{time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
