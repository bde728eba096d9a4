// Device-side LU driver (LU with partial pivoting and its solve).
#pragma once

#include "mops_kernels.hpp"

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

}  // namespace mops
