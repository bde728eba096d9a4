// Reverse-mode derivative of the Cholesky factorization (no counterpart in
// matops): given L = chol(A) (lower) and Lbar = d loss / dL (lower), the
// symmetric Abar = d loss / dA, computed in place in the buffer that holds
// Lbar.
//
// Closed form (I. Murray, "Differentiation of the Cholesky decomposition",
// arXiv:1602.07527, 2016, eq. 10): from dA = dL L^T + L dL^T,
// dL = L Phi(L^-1 dA L^-T), where Phi keeps the lower triangle and halves
// the diagonal; Phi is its own adjoint, so with W = L^-1
//   Abar = 1/2 (X + X^T),   X = W^T Phi(L^T Lbar) W
// (the symmetric part of the gradient over all n^2 entries of A). The five
// steps are large products whose operands are triangular, each skipping the
// k-slabs that multiply stored zeros:
//   P = tril(L^T Lbar)      lower output, k >= row            n^3/3 flops
//   W = L^-1                trtri (64x64 leaves, batched GEMM levels) n^3/3
//   Q = Phi(P) W            lower x lower, lower output, col <= k <= row  n^3/3
//   X = W^T Q               upper x lower, full output, k >= max(row, col) 2n^3/3
//   Abar = (X + X^T) / 2    one pass, tile pairs through LDS
// 5n^3/3 flops in a few launches that fill the GPU, against the 2n^3/3 of
// the blocked reverse sweep, whose n/64 diagonal steps are a chain of
// small dependent launches. The explicit inverse costs accuracy in
// proportion to cond(L) = sqrt(cond(A)), as the triangular solves of the
// sweep do.
//
// Memory: L and the adjoint buffer, plus three n x n matrices of workspace
// (and trtri's own), with the triangles the products read as zeros stored.
// The inverse runs on a second stream beside P, which it does not depend on.
#include "mops_solvers.hpp"

namespace mops {
namespace {

// One pass before the products: Lz = tril(L) with zeros above; zeros in the
// strict upper triangles of Pb and Ab (the products read the diagonal
// tiles of their triangular operands in full, so those zeros must be stored).
__global__ __launch_bounds__(256) void k_adj_prep(int n, const double* __restrict__ L, long ldl,
                                                  double* __restrict__ Lz, double* __restrict__ Pb,
                                                  double* __restrict__ Ab, long lda) {
  const long total = (long)n * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (i >= j) {
      Lz[e] = L[i + (long)j * ldl];
    } else {
      Lz[e] = 0.0;
      Pb[e] = 0.0;
      Ab[i + (long)j * lda] = 0.0;
    }
  }
}

// Phi on the diagonal: P[i][i] /= 2
__global__ void k_adj_half_diag(int n, double* P, long ldp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) P[i * (ldp + 1)] *= 0.5;
}

// Ab = (X + X^T) / 2 for the n x n X (leading dimension n): tile pair
// (bi, bj), bj <= bi, of 32 x 32, both tiles staged in LDS. Each element
// and its mirror get the same sum, so Ab is exactly symmetric.
__global__ __launch_bounds__(256) void k_adj_sym_half(int n, const double* __restrict__ X,
                                                      double* __restrict__ Ab, long lda) {
  __shared__ double lo[32][33], up[32][33];
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bj > bi) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int c = ty; c < 32; c += 8) {
    const int i = bi * 32 + tx, j = bj * 32 + c;  // tile (bi, bj): rows i, column j
    lo[c][tx] = (i < n && j < n) ? X[i + (long)j * n] : 0.0;
    const int i2 = bj * 32 + tx, j2 = bi * 32 + c;  // tile (bj, bi)
    up[c][tx] = (i2 < n && j2 < n) ? X[i2 + (long)j2 * n] : 0.0;
  }
  __syncthreads();
  for (int c = ty; c < 32; c += 8) {
    // lo[c][tx] = X[bi*32+tx][bj*32+c], its mirror X[bj*32+c][bi*32+tx] = up[tx][c]
    const int i = bi * 32 + tx, j = bj * 32 + c;
    if (i < n && j < n) Ab[i + (long)j * lda] = 0.5 * (lo[c][tx] + up[tx][c]);
    const int i2 = bj * 32 + tx, j2 = bi * 32 + c;
    if (bi != bj && i2 < n && j2 < n) Ab[i2 + (long)j2 * lda] = 0.5 * (up[c][tx] + lo[tx][c]);
  }
}

inline int grid_for(long total) {
  long b = (total + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

}  // namespace

size_t chol_adjoint_ws_bytes(int n, int nb) {
  (void)nb;
  if (n <= 0) return 0;
  return 3 * align_up((size_t)n * n * sizeof(double), 256) + trtri_ws_bytes(n) + 4096;
}

int chol_adjoint_d(hipStream_t s, int n, int nb, const double* L, long ldl, double* Ab,
                   long lda) {
  (void)nb;  // the closed form has no block size of its own
  if (n <= 0) return ST_OK;
  const size_t nn = (size_t)n * n;
  double* Lz = (double*)ws_carve(nn * sizeof(double));  // tril(L), zeros above
  double* Wv = (double*)ws_carve(nn * sizeof(double));  // W = L^-1
  double* Pb = (double*)ws_carve(nn * sizeof(double));  // P, then X
  if (!Lz || !Wv || !Pb) {
    set_error("chol_adjoint: workspace too small");
    return ST_NOMEM;
  }
  hipStream_t s2 = aux_stream();
  hipEvent_t e_prep = sync_event(0), e_inv = sync_event(1);
  if (!s2 || !e_prep || !e_inv) {
    set_error("chol_adjoint: could not create the second stream");
    return ST_HIP;
  }
  hipLaunchKernelGGL(k_adj_prep, dim3(grid_for((long)nn)), dim3(256), 0, s, n, L, ldl, Lz, Pb, Ab,
                     lda);
  MOPS_CHECK_HIP(hipGetLastError());
  MOPS_CHECK_HIP(hipEventRecord(e_prep, s));
  // P = tril(L^T Lbar): op(A) = L^T is upper triangular (k >= row)
  MOPS_TRY(gemm_d(s, true, false, n, n, n, 1.0, Lz, n, Ab, lda, 0.0, Pb, n, GM_LOWER,
                  TRI_A_UPPER));
  hipLaunchKernelGGL(k_adj_half_diag, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, Pb, (long)n);
  MOPS_CHECK_HIP(hipGetLastError());
  // W = L^-1 on the second stream, beside P (both read only Lz; P is queued
  // first, so it does not wait for the host to queue the inverse's launches)
  MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_prep, 0));
  MOPS_TRY(trtri_d(s2, true, false, n, Lz, n, Wv, n));
  MOPS_CHECK_HIP(hipEventRecord(e_inv, s2));
  MOPS_CHECK_HIP(hipStreamWaitEvent(s, e_inv, 0));
  // Q = Phi(P) W into the adjoint buffer (Lbar has been consumed; the zeros
  // above the diagonal stay: only the lower triangle is written)
  MOPS_TRY(gemm_d(s, false, false, n, n, n, 1.0, Pb, n, Wv, n, 0.0, Ab, lda, GM_LOWER,
                  TRI_AB_LOWER));
  // X = W^T Q into Pb
  MOPS_TRY(gemm_d(s, true, false, n, n, n, 1.0, Wv, n, Ab, lda, 0.0, Pb, n, GM_FULL,
                  TRI_AU_BL));
  const int nt = ceil_div(n, 32);
  hipLaunchKernelGGL(k_adj_sym_half, dim3(nt, nt), dim3(256), 0, s, n, Pb, Ab, lda);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

}  // namespace mops
