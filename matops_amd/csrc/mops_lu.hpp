// Device-side LU driver (LU with partial pivoting and its solve).
#pragma once

#include "mops_kernels.hpp"
#include "mops_solvers.hpp"

namespace mops {

size_t getrf_ws_bytes(int m, int n, int nb);
// A (m x n, column-major) = P^T L U in place. ipiv (device, may be null):
// min(m,n) 0-based global pivots as the reference stores them (lu.go:257).
// perm_out (device int[m], may be null): perm[i] = original row now at i.
// err_flag (device int, may be null): 1 if a leaf's grid barrier timed out.
int getrf_d(hipStream_t s, int m, int n, int nb, double* A, long lda, long long* ipiv,
            int* perm_out, int* err_flag);

size_t getrs_ws_bytes(int n, int nrhs);
// B (n x nrhs) = A^-1 B (or A^-T B with TRANSA) from getrf_d's factors.
// perm may be null: it is then rebuilt from the npiv pivots.
int getrs_d(hipStream_t s, int flags, int n, int nrhs, const double* A, long lda,
            const long long* ipiv, int npiv, const int* perm, double* B, long ldb);

// getrs_d without TRANSA, repeated with one set of factors (iterative
// refinement): getrs_prep_d computes the inverses of the diagonal blocks of
// L and U once (workspace getrs_prep_ws_bytes, held until the last solve);
// getrs_prep_solve_d then sets Out = A^-1 R (R is read only; Out may not
// alias R) with them.
struct GetrsPrep {
  TrsmPrep l, u;
};
size_t getrs_prep_ws_bytes(int n, int nrhs);
int getrs_prep_d(hipStream_t s, int n, int nrhs, const double* A, long lda, GetrsPrep& g);
int getrs_prep_solve_d(hipStream_t s, int n, int nrhs, const double* A, long lda, const int* perm,
                       const double* R, long ldr, double* Out, long ldo, const GetrsPrep& g);

// Mixed precision LU of the n x n matrix A (read only): the trailing matrix
// in FP32 (FP32 matrix cores), each panel widened to FP64 and factored by the
// FP64 panel code. LU (n x n, FP64) receives the factors (their FP32
// values), perm_out the row permutation, ipiv the pivots. bf16: the
// trailing products A22 -= L21 U12 take BF16 operands (U12 scaled by a
// power of two per panel) on the BF16 matrix cores, accumulating in FP32.
size_t getrf_mixed_ws_bytes(int n, int nb);
int getrf_mixed_d(hipStream_t s, int n, int nb, const double* A, long lda, double* LU, long ldlu,
                  long long* ipiv, int* perm_out, int* err_flag, bool bf16 = false);

}  // namespace mops

#include "../../include/mops.h"

namespace mops {

// A X = B by LU in precision prec (0 = FP64, 1 = FP32) plus FP64 iterative
// refinement until the normwise backward error is <= tol (at most maxit
// solves). X may not alias B.
size_t gesv_ir_ws_bytes(int n, int nrhs, int prec);
int gesv_ir_d(hipStream_t s, int n, int nrhs, const double* A, long lda, const double* B,
              long ldb, double* X, long ldx, int prec, double tol, int maxit,
              mops_ir_report* res);

}  // namespace mops
