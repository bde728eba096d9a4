"""matops_amd: the dense FP64 path of matops (GEMM, TRSM, SYRK, Cholesky,
LU with partial pivoting) on AMD MI355X (gfx950), behind a C ABI.

The compute lives in libmops.so (hand-written HIP kernels); this package is
the host-side mirror of the reference's Go API and the ctypes binding.
"""
from .matops import (  # noqa: F401
    LEFT, LOWER, NONE, NOTRANS, RIGHT, TRANS, TRANSA, TRANSB, UNIT, UPPER,
    BlockingParams, CholeskyAdjoint, DecomposeCHOL, DecomposeLU, MatopsError, Mult, NumWorkers,
    PREC_BF16, PREC_FP32, PREC_FP64, RankUpdateSym, SetPanicOnError, SolveCHOL, SolveIR, SolveLU, SolveTrm,
    blockIndex4, divideWork, scheduleWork, MultSym, MultTrm, RankUpdate2Sym, UpdateTrm,
    DecomposeLUnoPiv, DecomposeLDLnoPiv, SolveLDLnoPiv, InverseTrm,
    DecomposeLDL, SolveLDL, DecomposeBK, SolveBK,
)
