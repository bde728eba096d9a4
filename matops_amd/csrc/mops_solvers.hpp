// Device-side drivers of the factorizations and triangular solves.
// All pointers are device pointers; work is queued on `s`.
#pragma once

#include "mops_kernels.hpp"

namespace mops {

// Bytes of workspace trtri_d needs beyond X for an n x n triangle.
size_t trtri_ws_bytes(int n);
// Bytes of workspace trtri_blocks_d needs beyond X (the ceil(n/bs) inverses
// of bs x bs blocks).
size_t trtri_blocks_ws_bytes(int n, int bs);
// X (n x n, dense, zeros in the other triangle) = inverse of the lower
// (lower=true) or upper triangle of A; unit: the diagonal is taken as 1 and
// never read. Inverts all 64x64 diagonal blocks in one launch, then combines
// them level by level with batched GEMMs. Uses workspace (carve).
int trtri_d(hipStream_t s, bool lower, bool unit, int n, const double* A, long lda, double* X,
            long ldx);

// Inverse of every bs x bs diagonal block of the n x n triangle of A
// (the last block may be smaller), stored block b at X + b*bs*bs with
// leading dimension bs.
int trtri_blocks_d(hipStream_t s, bool lower, bool unit, int n, int bs, const double* A,
                   long lda, double* X);

// Workspace needed by trsm_d.
size_t trsm_ws_bytes(int flags, int N, int nrhs);
// B = alpha * op(A)^-1 * B (LEFT, B is N x nrhs) or alpha * B * op(A)^-1
// (RIGHT, B is nrhs x N). Flag semantics of dmmat_solve_blk
// (calgo/dmmat_solve.c:864-909): RIGHT selects the side, otherwise LEFT;
// UPPER selects the upper triangle, otherwise lower; TRANSA transposes A;
// UNIT never reads the diagonal.
int trsm_d(hipStream_t s, int flags, int N, int nrhs, double alpha, const double* A, long lda,
           double* B, long ldb);

// Two LEFT solves with the same triangle of A (flags differing only in
// TRANSA), alpha = 1, one after the other on B; with few right-hand sides
// the inverses of the diagonal blocks are computed once for both.
int trsm_pair_d(hipStream_t s, int flags1, int flags2, int N, int nrhs, const double* A,
                long lda, double* B, long ldb);

// LEFT solves repeated with one factor (the refinement's solves): the
// inverses of the 64x64 diagonal blocks are computed once by trsm_prep_d,
// trsm_prep_solve_d then solves B = op(A)^-1 B (alpha = 1) with them. With
// many right-hand sides (no chained solve) nothing is prepared and the
// solve is trsm_d. W, when not null, is the sentinel-filled copy of the
// solution the chained solve needs (trsm_prep_w_bytes; may be shared by
// preparations used one after the other on one stream).
struct TrsmPrep {
  double* D = nullptr;
  double* W = nullptr;
  long ldw = 0;
  int nb64 = 0, ngrp = 0;
};
size_t trsm_prep_ws_bytes(int N, int nrhs);  // D, W and the inversion's scratch
size_t trsm_prep_w_bytes(int N, int nrhs);
int trsm_prep_d(hipStream_t s, int flags, int N, int nrhs, const double* A, long lda, double* W,
                TrsmPrep& p);
int trsm_prep_solve_d(hipStream_t s, int flags, int N, int nrhs, const double* A, long lda,
                      double* B, long ldb, const TrsmPrep& p);

// The other BLAS-3 operations (blas3x.hip). symm_d: C (M x N) = beta C +
// alpha S B (left: S = sym(A), M x M) or alpha B S (S N x N), sym(A) the
// symmetric matrix whose lower (else upper) triangle A stores; needs
// symm_ws_bytes(order of S). trmm_d: B = alpha op(A) B (LEFT, B N x nrhs) or
// alpha B op(A) (RIGHT, B nrhs x N), triangle and UNIT as in trsm_d; needs
// trmm_ws_bytes. syr2k_d: one triangle (LOWER, else UPPER) of C (N x N) =
// beta C + alpha (op(A) op(B)^T + op(B) op(A)^T), op = transpose with
// TRANSA, inner length P. trmupd_d: one triangle (UPPER, else LOWER) of
// C = beta C + alpha op(A) op(B) (TRANSA / TRANSB).
size_t symm_ws_bytes(int n);
int sym_expand_d(hipStream_t s, bool lower, int n, const double* A, long lda, double* Out,
                 long ldo);
int symm_d(hipStream_t s, bool left, bool lower, int M, int N, double alpha, const double* A,
           long lda, const double* B, long ldb, double beta, double* C, long ldc);
size_t trmm_ws_bytes(int N, int nrhs);
int trmm_d(hipStream_t s, int flags, int N, int nrhs, double alpha, const double* A, long lda,
           double* B, long ldb);
int syr2k_d(hipStream_t s, int flags, int N, int P, double alpha, const double* A, long lda,
            const double* B, long ldb, double beta, double* C, long ldc);
int trmupd_d(hipStream_t s, int flags, int N, int P, double alpha, const double* A, long lda,
             const double* B, long ldb, double beta, double* C, long ldc);

// Factorizations without pivoting and the triangular inverse (factor_x.hip).
// getrf_nopiv_d: A (m x n) = L U, unit L below the diagonal, no row
// interchanges. sytrf_nopiv_d: A (n x n) = L D L^T (lower) or U D U^T, unit
// L / U off the diagonal, D on it, the other triangle untouched;
// sytrs_nopiv_d: B = A^-1 B from those factors. inverse_trm_d: the lower
// (else upper) triangle of A replaced by that of its inverse; with unit the
// diagonal is taken as 1 and left as stored.
size_t getrf_nopiv_ws_bytes(int m, int n);
int getrf_nopiv_d(hipStream_t s, int m, int n, double* A, long lda);
size_t sytrf_nopiv_ws_bytes(int n);
int sytrf_nopiv_d(hipStream_t s, bool lower, int n, double* A, long lda);
size_t sytrs_nopiv_ws_bytes(int n, int nrhs);
int sytrs_nopiv_d(hipStream_t s, bool lower, int n, int nrhs, const double* A, long lda,
                  double* B, long ldb);
size_t inverse_trm_ws_bytes(int n);
int inverse_trm_d(hipStream_t s, bool lower, bool unit, int n, double* A, long lda);

// Symmetric indefinite factorizations with pivoting (sym_piv.hip).
// ldl_piv_d: A (n x n) = P^T L D L^T P (lower) or P^T U D U^T P in place,
// 1x1 pivots, the reference's ipiv format (1-based, 0 = no interchange);
// the npiv entries of ipiv are cleared first. A zero pivot stops it:
// info (device int, may be null) receives its 0-based index or -1, and
// err_host (may be null; then the call synchronizes s) 1 + that index or 0.
// The blocked upper form (nb > 0, n >= nb) synchronizes s once per block.
// bk_d: Bunch-Kaufman A = L D L^T / U D U^T, LAPACK's ipiv format.
// ldl_piv_solve_d / bk_solve_d: B (n x nrhs) = A^-1 B from those factors.
size_t ldl_piv_ws_bytes(int n, int nb, bool upper);
int ldl_piv_d(hipStream_t s, bool upper, int n, int nb, double* A, long lda, int64_t* ipiv,
              long npiv, int* info, int* err_host);
size_t bk_ws_bytes(int n);
int bk_d(hipStream_t s, bool upper, int n, double* A, long lda, int64_t* ipiv, long npiv);
size_t ldl_piv_solve_ws_bytes(int n, int nrhs);
int ldl_piv_solve_d(hipStream_t s, bool upper, int n, int nrhs, const double* A, long lda,
                    const int64_t* ipiv, int npiv, double* B, long ldb);
int bk_solve_d(hipStream_t s, bool upper, int n, int nrhs, const double* A, long lda,
               const int64_t* ipiv, double* B, long ldb);

// Cholesky of the N x N matrix A in place (LOWER or UPPER flag), outer block
// nb. `info` (device int, may be null) receives the 0-based global index of
// the first NaN square root, or stays -1. `badval` (device double, may be
// null) receives the value under that square root.
size_t potrf_ws_bytes(int N, int nb);
int potrf_d(hipStream_t s, int flags, int N, int nb, double* A, long lda, int* info,
            double* badval);

// B = A^-1 B from the Cholesky factor (SolveCHOL, chol.go:147-158).
size_t potrs_ws_bytes(int flags, int N, int nrhs);
int potrs_d(hipStream_t s, int flags, int N, int nrhs, const double* A, long lda, double* B,
            long ldb);

// Reverse-mode derivative of the Cholesky factorization: L (lower factor,
// N x N) and, in Ab, Lbar = d loss / dL (lower triangle read); on return Ab
// holds the symmetric d loss / dA. Blocked reverse sweep, block nb <= 64.
size_t chol_adjoint_ws_bytes(int n, int nb);
int chol_adjoint_d(hipStream_t s, int n, int nb, const double* L, long ldl, double* Ab, long lda);

}  // namespace mops
