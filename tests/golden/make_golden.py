"""Writes tests/golden/golden.npz: the oracle's outputs for small cases of
the path (GEMM, TRSM, SYRK, Cholesky, LU with pivots, solves). Inputs are
regenerated from the seeds in golden_inputs.py; the file keeps SHA-1
digests of them so a changed generator is detected.

The oracle runs with the reference's own calgo kernels when oracle/_ref was
built (else with the restated ones); which one is recorded in the file.
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
from oracle import oracle as O  # noqa: E402
import golden_inputs as gi  # noqa: E402

OUT = os.path.join(HERE, "golden.npz")


def main():
    kind = "reference" if O.reference_available() else "port"
    O.use(kind)
    d = {"kernels": np.array(kind)}
    for flags in gi.GEMM_FLAGS:
        A, B, C0 = gi.gemm_case(flags)
        d[f"gemm_{flags}_sha"] = np.array(gi.digest(np.concatenate([A.ravel("F"), B.ravel("F")])))
        for ab, (al, be) in enumerate(gi.GEMM_AB):
            C = C0.copy(order="F")
            O.Mult(C, A, B, al, be, flags)
            d[f"gemm_{flags}_{ab}_C"] = C
    for flags in gi.TRSM_FLAGS:
        T, B0 = gi.trsm_case(flags)
        B = B0.copy(order="F")
        O.SolveTrm(B, T, gi.TRSM_ALPHA, flags)
        d[f"trsm_{flags}_X"] = B
    for flags in gi.SYRK_FLAGS:
        A, C0 = gi.syrk_case(flags)
        C = C0.copy(order="F")
        O.RankUpdateSym(C, A, -1.0, 1.0, flags)
        d[f"syrk_{flags}_C"] = C
    for n, nb in gi.CHOL_CASES:
        S, B = gi.chol_case(n)
        for uplo in (0x4, 0x8):
            F = S.copy(order="F")
            rc, _, _ = O.DecomposeCHOL(F, uplo, nb)
            assert rc == 0
            X = B.copy(order="F")
            O.SolveCHOL(X, F, uplo)
            key = f"chol_{n}_{nb}_{uplo}"
            d[key + "_Asha"], d[key + "_X"] = np.array(gi.digest(S)), X
            if n <= 64:
                d[key + "_F"] = F
    for n, nb in gi.LU_CASES:
        A, B = gi.lu_case(n)
        F = A.copy(order="F")
        piv = np.zeros(n, dtype=np.int64)
        O.DecomposeLU(F, piv, nb)
        X = B.copy(order="F")
        O.SolveLU(X, F, piv, 0)
        key = f"lu_{n}_{nb}"
        d[key + "_Asha"], d[key + "_piv"], d[key + "_X"] = np.array(gi.digest(A)), piv, X
        if n <= 64:
            d[key + "_F"] = F
    T = gi.ties_case()
    F = T.copy(order="F")
    piv = np.zeros(4, dtype=np.int64)
    O.DecomposeLU(F, piv, 0)
    d["ties_F"], d["ties_piv"] = F, piv
    np.savez_compressed(OUT, **d)
    print(OUT, os.path.getsize(OUT), "bytes, kernels:", kind)


if __name__ == "__main__":
    main()
