// Reverse-mode derivative of the Cholesky factorization (no counterpart in
// matops): given L = chol(A) (lower) and Lbar = d loss / dL (lower), the
// symmetric Abar = d loss / dA, computed in place in the buffer that holds
// Lbar.
//
// Closed form (I. Murray, "Differentiation of the Cholesky decomposition",
// arXiv:1602.07527, 2016, eq. 10): from dA = dL L^T + L dL^T,
// dL = L Phi(L^-1 dA L^-T), where Phi keeps the lower triangle and halves
// the diagonal; Phi is its own adjoint, so
//   Abar = (X + X^T) / 2,   X = L^-T Phi(L^T Lbar) L^-1
// (the symmetric part of the gradient over all n^2 entries of A).
//
// Evaluated in place in the adjoint buffer by three blocked passes over
// block rows / columns of order bs, each a sequence of rank-bs updates (one
// large GEMM that fills the GPU) and bs-wide diagonal steps (a GEMM with a
// stored bs x bs block):
//   P = tril(L^T Lbar)   (TRMM, U = L^T upper, lower output), block rows K
//                        ascending: rows above += L[K, :K]^T Lbar[K, :],
//                        then Lbar[K, :] = L_KK^T Lbar[K, :]       n^3/3 flops
//   Q = Phi(P) L^-1      (TRSM from the right, lower result), block columns
//                        J descending: Q[:, J] = Q[:, J] L_JJ^-1, then
//                        columns left -= Q[:, J] L[J, :J]          n^3/3
//   X = L^-T Q           (TRSM from the left with L^T, full result), block
//                        rows I descending: X[I, :] = L_II^-T X[I, :], then
//                        rows above -= L[I, :I]^T X[I, :]          n^3
//   Abar = (X + X^T) / 2 (tile pairs through LDS)
// The diagonal steps use the inverses of the diagonal blocks of L, all made
// in one batched launch up front, and clean copies of those blocks (the
// upper triangle of L's buffer is never read).
//
// Memory: L and the adjoint buffer, plus O(bs n) of workspace (the block
// inverses and copies, and two bs-wide block rows or columns). The same
// formula with an explicit L^-1 and out-of-place products runs faster
// (2.7 ms at n = 4096) but needs three n x n buffers.
#include "mops_solvers.hpp"

namespace mops {
namespace {

// Default block order: rank-512 updates run near the GEMM's full rate, and
// the workspace (about 5 * 512 * n doubles) stays a small multiple of n.
// n = 4096 measured 10.6 (64), 6.6 (128), 4.6 (256), 3.7 (512), 3.3 ms
// (1024, where the workspace reaches n^2).
constexpr int ADJ_BS = 512;

// zeros in the strict upper triangle of the n x n A
__global__ void k_adj_zero_upper(int n, double* A, long lda) {
  const long total = (long)n * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (i < j) A[i + (long)j * lda] = 0.0;
  }
}

// D[b] (bs x bs, leading dimension bs, the last block jb x jb) = lower
// triangle of the b-th diagonal block of L, zeros above
__global__ void k_adj_diag_blocks(int n, int bs, const double* __restrict__ L, long ldl,
                                  double* __restrict__ D) {
  const int b = blockIdx.y, o = b * bs, jb = min(bs, n - o);
  double* Db = D + (long)b * bs * bs;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < jb * jb; e += gridDim.x * blockDim.x) {
    const int i = e % jb, j = e / jb;
    Db[i + (long)j * bs] = i >= j ? L[(o + i) + (long)(o + j) * ldl] : 0.0;
  }
}

// B (m x n at global offset (r0, c0) of the matrix) = T (leading dimension
// ldt); with lower != 0 only the elements with global row >= column
__global__ void k_adj_put(int m, int n, const double* __restrict__ T, long ldt,
                          double* __restrict__ B, long ldb, int r0, int c0, int lower) {
  const long total = (long)m * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % m), j = (int)(e / m);
    if (!lower || r0 + i >= c0 + j) B[i + (long)j * ldb] = T[i + (long)j * ldt];
  }
}

// Phi on the diagonal: P[i][i] /= 2
__global__ void k_adj_half_diag(int n, double* P, long ldp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) P[i * (ldp + 1)] *= 0.5;
}

// A = (A + A^T) / 2 in place: tile pair (bi, bj), bj <= bi, of 32 x 32,
// both tiles staged in LDS before either is written. Each element and its
// mirror get the same sum, so A is exactly symmetric.
__global__ __launch_bounds__(256) void k_adj_sym_half(int n, double* A, long lda) {
  __shared__ double lo[32][33], up[32][33];
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bj > bi) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int c = ty; c < 32; c += 8) {
    const int i = bi * 32 + tx, j = bj * 32 + c;  // tile (bi, bj): row i, column j
    lo[c][tx] = (i < n && j < n) ? A[i + (long)j * lda] : 0.0;
    const int i2 = bj * 32 + tx, j2 = bi * 32 + c;  // tile (bj, bi)
    up[c][tx] = (i2 < n && j2 < n) ? A[i2 + (long)j2 * lda] : 0.0;
  }
  __syncthreads();
  for (int c = ty; c < 32; c += 8) {
    // lo[c][tx] = A[bi*32+tx][bj*32+c]; its mirror A[bj*32+c][bi*32+tx] = up[tx][c]
    const int i = bi * 32 + tx, j = bj * 32 + c;
    if (i < n && j < n) A[i + (long)j * lda] = 0.5 * (lo[c][tx] + up[tx][c]);
    const int i2 = bj * 32 + tx, j2 = bi * 32 + c;
    if (bi != bj && i2 < n && j2 < n) A[i2 + (long)j2 * lda] = 0.5 * (up[c][tx] + lo[tx][c]);
  }
}

inline int grid_for(long total) {
  long b = (total + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

inline int adj_block(int n, int nb) {
  int bs = nb <= 0 ? ADJ_BS : (nb + 63) / 64 * 64;
  if (bs > 1024) bs = 1024;
  if (bs > n) bs = (n + 63) / 64 * 64;
  return bs < 64 ? 64 : bs;
}

}  // namespace

size_t chol_adjoint_ws_bytes(int n, int nb) {
  if (n <= 0) return 0;
  const int bs = adj_block(n, nb);
  const size_t nblk = (size_t)ceil_div(n, bs);
  const size_t blocks = align_up(nblk * bs * bs * sizeof(double), 256);
  const size_t strip = align_up((size_t)n * bs * sizeof(double), 256);
  // inverses and copies of the diagonal blocks, two block rows / columns,
  // the inverses' construction, the split of the bs-row products
  return 2 * blocks + 2 * strip + trtri_blocks_ws_bytes(n, bs) +
         gemm_split_ws_bytes(bs, n, n, sizeof(double)) + 4096;
}

int chol_adjoint_d(hipStream_t s, int n, int nb, const double* L, long ldl, double* Ab,
                   long lda) {
  if (n <= 0) return ST_OK;
  const int bs = adj_block(n, nb);
  const int nblk = ceil_div(n, bs);
  double* Dinv = (double*)ws_carve((size_t)nblk * bs * bs * sizeof(double));
  double* Ld = (double*)ws_carve((size_t)nblk * bs * bs * sizeof(double));
  double* T2[2] = {(double*)ws_carve((size_t)n * bs * sizeof(double)),
                   (double*)ws_carve((size_t)n * bs * sizeof(double))};
  if (!Dinv || !Ld || !T2[0] || !T2[1]) {
    set_error("chol_adjoint: workspace too small");
    return ST_NOMEM;
  }
  hipStream_t s2 = aux_stream();
  hipEvent_t e_main = sync_event(0), e_side = sync_event(1), e_free = sync_event(2);
  if (!s2 || !e_main || !e_side || !e_free) {
    set_error("chol_adjoint: could not create the second stream");
    return ST_HIP;
  }
  {
    const size_t mark = ws_mark();
    const int st = trtri_blocks_d(s, true, false, n, bs, L, ldl, Dinv);
    ws_release(mark);
    MOPS_TRY(st);
  }
  hipLaunchKernelGGL(k_adj_diag_blocks, dim3(ceil_div(bs * bs, 256), nblk), dim3(256), 0, s, n,
                     bs, L, ldl, Ld);
  hipLaunchKernelGGL(k_adj_zero_upper, dim3(grid_for((long)n * n)), dim3(256), 0, s, n, Ab, lda);
  MOPS_CHECK_HIP(hipGetLastError());
  // the second stream starts after the set-up
  MOPS_CHECK_HIP(hipEventRecord(e_main, s));
  MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_main, 0));

  // P = tril(L^T Lbar), block rows ascending. Block row K of the product
  // sums L[K', K]^T Lbar[K', :] over K' >= K, so block row K of Lbar is read
  // by the rows above (rank-bs update, main stream) and by its own diagonal
  // block (second stream, into T2[K % 2], beside that update) before it is
  // overwritten; no earlier step writes block row K.
  for (int b = 0; b < nblk; ++b) {
    const int k0 = b * bs, kb = min(bs, n - k0), k1 = k0 + kb;
    double* T = T2[b & 1];
    if (b >= 2) MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_free, 0));  // T2[b % 2] consumed
    MOPS_TRY(gemm_d(s2, true, false, kb, k1, kb, 1.0, Ld + (long)b * bs * bs, bs, Ab + k0, lda,
                    0.0, T, kb, GM_FULL, TRI_A_UPPER));
    MOPS_CHECK_HIP(hipEventRecord(e_side, s2));
    if (k0 > 0)
      MOPS_TRY(gemm_d(s, true, false, k0, k1, kb, 1.0, L + k0, ldl, Ab + k0, lda, 1.0, Ab, lda,
                      GM_LOWER_T));
    MOPS_CHECK_HIP(hipStreamWaitEvent(s, e_side, 0));
    hipLaunchKernelGGL(k_adj_put, dim3(grid_for((long)kb * k1)), dim3(256), 0, s, kb, k1, T,
                       (long)kb, Ab + k0, lda, k0, 0, 1);
    MOPS_CHECK_HIP(hipGetLastError());
    // T2[(b - 1) % 2] was consumed by the previous put, which the main stream
    // has queued before this point
    MOPS_CHECK_HIP(hipEventRecord(e_free, s));
  }
  hipLaunchKernelGGL(k_adj_half_diag, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, Ab, lda);
  MOPS_CHECK_HIP(hipGetLastError());

  // Q = Phi(P) L^-1 (Q L = Phi), block columns descending; Q is lower, so
  // block column J has rows j0.. only. Look-ahead: the update by block
  // column J first brings column block J - 1 up to date (main stream), whose
  // diagonal step then runs on the second stream beside the rest of the
  // update.
  double* T = T2[0];
  {
    const int b = nblk - 1, j0 = b * bs, jb = n - j0, m = jb;
    double* AJ = Ab + j0 + (long)j0 * lda;
    MOPS_TRY(gemm_d(s, false, false, m, jb, jb, 1.0, AJ, lda, Dinv + (long)b * bs * bs, bs, 0.0,
                    T, m, GM_FULL, TRI_B_LOWER));
    hipLaunchKernelGGL(k_adj_put, dim3(grid_for((long)m * jb)), dim3(256), 0, s, m, jb, T,
                       (long)m, AJ, lda, j0, j0, 1);
    MOPS_CHECK_HIP(hipGetLastError());
  }
  for (int b = nblk - 1; b >= 1; --b) {
    const int j0 = b * bs, jb = min(bs, n - j0), m = n - j0;
    const int p0 = j0 - bs, mp = n - p0;  // the next block column (order bs)
    double* AJ = Ab + j0 + (long)j0 * lda;
    double* AP = Ab + p0 + (long)p0 * lda;
    // column block b - 1, rows j0..: -= Q[:, J] L[J, J-1]
    MOPS_TRY(gemm_d(s, false, false, m, bs, jb, -1.0, AJ, lda, L + j0 + (long)p0 * ldl, ldl, 1.0,
                    Ab + j0 + (long)p0 * lda, lda, GM_FULL, TRI_A_LOWER));
    MOPS_CHECK_HIP(hipEventRecord(e_main, s));
    MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_main, 0));
    MOPS_TRY(gemm_d(s2, false, false, mp, bs, bs, 1.0, AP, lda, Dinv + (long)(b - 1) * bs * bs, bs,
                    0.0, T, mp, GM_FULL, TRI_B_LOWER));
    hipLaunchKernelGGL(k_adj_put, dim3(grid_for((long)mp * bs)), dim3(256), 0, s2, mp, bs, T,
                       (long)mp, AP, lda, p0, p0, 1);
    MOPS_CHECK_HIP(hipGetLastError());
    MOPS_CHECK_HIP(hipEventRecord(e_side, s2));
    if (p0 > 0)
      MOPS_TRY(gemm_d(s, false, false, m, p0, jb, -1.0, AJ, lda, L + j0, ldl, 1.0, Ab + j0, lda,
                      GM_FULL, TRI_A_LOWER));
    MOPS_CHECK_HIP(hipStreamWaitEvent(s, e_side, 0));
  }

  // X = L^-T Q, block rows descending (X is full); the same look-ahead by
  // block rows
  {
    const int b = nblk - 1, i0 = b * bs, ib = n - i0;
    MOPS_TRY(gemm_d(s, true, false, ib, n, ib, 1.0, Dinv + (long)b * bs * bs, bs, Ab + i0, lda,
                    0.0, T, ib, GM_FULL, TRI_A_UPPER));
    hipLaunchKernelGGL(k_adj_put, dim3(grid_for((long)ib * n)), dim3(256), 0, s, ib, n, T,
                       (long)ib, Ab + i0, lda, i0, 0, 0);
    MOPS_CHECK_HIP(hipGetLastError());
  }
  for (int b = nblk - 1; b >= 1; --b) {
    const int i0 = b * bs, ib = min(bs, n - i0);
    const int p0 = i0 - bs;  // the next block row (order bs)
    // block row b - 1 -= L[I, I-1]^T X[I, :]
    MOPS_TRY(gemm_d(s, true, false, bs, n, ib, -1.0, L + i0 + (long)p0 * ldl, ldl, Ab + i0, lda,
                    1.0, Ab + p0, lda));
    MOPS_CHECK_HIP(hipEventRecord(e_main, s));
    MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_main, 0));
    MOPS_TRY(gemm_d(s2, true, false, bs, n, bs, 1.0, Dinv + (long)(b - 1) * bs * bs, bs, Ab + p0,
                    lda, 0.0, T, bs, GM_FULL, TRI_A_UPPER));
    hipLaunchKernelGGL(k_adj_put, dim3(grid_for((long)bs * n)), dim3(256), 0, s2, bs, n, T,
                       (long)bs, Ab + p0, lda, p0, 0, 0);
    MOPS_CHECK_HIP(hipGetLastError());
    MOPS_CHECK_HIP(hipEventRecord(e_side, s2));
    if (p0 > 0)
      MOPS_TRY(gemm_d(s, true, false, p0, n, ib, -1.0, L + i0, ldl, Ab + i0, lda, 1.0, Ab, lda));
    MOPS_CHECK_HIP(hipStreamWaitEvent(s, e_side, 0));
  }

  const int nt = ceil_div(n, 32);
  hipLaunchKernelGGL(k_adj_sym_half, dim3(nt, nt), dim3(256), 0, s, n, Ab, lda);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

}  // namespace mops
