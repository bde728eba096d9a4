// Internal device-side building blocks (device pointers, explicit stream).
// Every routine is asynchronous on `s` and returns a status code.
#pragma once

#include "gemm_plan.hpp"
#include "mops_common.hpp"

namespace mops {

// Which part of C a GEMM writes. GM_LOWER / GM_UPPER require M == N and
// only tiles that meet the triangle are launched (SYRK-shaped updates).
// GM_LOWER_T / GM_UPPER_T: any M x N, only elements with row >= col
// (row <= col) are written (trapezoids: a diagonal block plus a rectangle).
enum : int { GM_FULL = 0, GM_LOWER = 1, GM_UPPER = 2, GM_LOWER_T = 3, GM_UPPER_T = 4 };

// Known-zero structure of an operand, used only to skip k-slabs that are
// multiplied by stored zeros. The zeros must really be stored: a skipped
// slab is never read, a partially covered one is read in full.
//   TRI_A_LOWER: op(A)[i][k] == 0 for k > i      TRI_A_UPPER: k < i
//   TRI_B_LOWER: op(B)[k][j] == 0 for k < j      TRI_B_UPPER: k > j
//   TRI_AB_LOWER: both TRI_A_LOWER and TRI_B_LOWER (j <= k <= i)
//   TRI_AU_BL:    both TRI_A_UPPER and TRI_B_LOWER (k >= max(i, j))
enum : int {
  TRI_NONE = 0,
  TRI_A_LOWER = 1,
  TRI_A_UPPER = 2,
  TRI_B_LOWER = 3,
  TRI_B_UPPER = 4,
  TRI_AB_LOWER = 5,
  TRI_AU_BL = 6
};

// While one is in scope, the split-k launches of gemm_d / gemm_s on the
// calling thread take their partial tiles and flags from this device buffer
// (bytes long, 256-byte aligned) instead of the workspace; a split that
// does not fit then runs unsplit.
class SplitScratchOverride {
 public:
  SplitScratchOverride(void* base, size_t bytes);
  ~SplitScratchOverride();
  SplitScratchOverride(const SplitScratchOverride&) = delete;
  SplitScratchOverride& operator=(const SplitScratchOverride&) = delete;

 private:
  void* prev_base_;
  size_t prev_bytes_;
};

// C(MxN) = alpha*op(A)*op(B) + beta*C, column-major. beta == 0 never reads C
// (NaN/Inf in C are cleared, as dscale_tile does at calgo/vec_scale.c:114).
int gemm_d(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
           const double* A, long lda, const double* B, long ldb, double beta,
           double* C, long ldc, int mode = GM_FULL, int tri = TRI_NONE);
int gemm_s(hipStream_t s, bool ta, bool tb, int M, int N, int K, float alpha,
           const float* A, long lda, const float* B, long ldb, float beta,
           float* C, long ldc, int mode = GM_FULL, int tri = TRI_NONE);

// C (M x N, FP32) += alpha * A (M x K, FP32) * B (K x N, FP32), the
// operands rounded to BF16 (A first scaled by the power of two that brings
// max |A| into [1/2, 1)), products on the BF16 matrix cores with FP32
// accumulation. amax_word: a device word the call uses for max |A| (calls
// on one stream may share it).
int gemm_bf16_s(hipStream_t s, int M, int N, int K, float alpha, const float* A, long lda,
                const float* B, long ldb, float* C, long ldc, unsigned* amax_word);

// Batched form: entry b uses A + b*sA, B + b*sB, C + b*sC (full C only).
int gemm_batched_d(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                   const double* A, long lda, long sA, const double* B, long ldb, long sB,
                   double beta, double* C, long ldc, long sC, int batch, int tri = TRI_NONE);
int gemm_batched_s(hipStream_t s, bool ta, bool tb, int M, int N, int K, float alpha,
                   const float* A, long lda, long sA, const float* B, long ldb, long sB,
                   float beta, float* C, long ldc, long sC, int batch, int tri = TRI_NONE);

// Two batch dimensions: entry (b, c) uses A + b*sA + c*sA2, and so on.
int gemm_batched2_d(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                    const double* A, long lda, long sA, long sA2, const double* B, long ldb,
                    long sB, long sB2, double beta, double* C, long ldc, long sC, long sC2,
                    int batch, int batch2, int tri = TRI_NONE);

// C = beta*C on an MxN tile (beta == 0 writes zeros, beta == 1 is a no-op).
int scale_tile_d(hipStream_t s, int M, int N, double beta, double* C, long ldc);
// Copy an MxN column-major tile.
int copy_tile_d(hipStream_t s, int M, int N, const double* A, long lda, double* B, long ldb);
int copy_tile_s(hipStream_t s, int M, int N, const float* A, long lda, float* B, long ldb);
// Precision conversions of an MxN tile.
int convert_d2s(hipStream_t s, int M, int N, const double* A, long lda, float* B, long ldb);
int convert_s2d(hipStream_t s, int M, int N, const float* A, long lda, double* B, long ldb);

}  // namespace mops
