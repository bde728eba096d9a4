// Reverse-mode derivative of the Cholesky factorization (no counterpart in
// matops): given L = chol(A) (lower) and Lbar = d loss / dL (lower), the
// symmetric Abar = d loss / dA, computed in place in the buffer that holds
// Lbar.
//
// Blocked reverse sweep (I. Murray, "Differentiation of the Cholesky
// decomposition", arXiv:1602.07527, 2016, the level-3 algorithm). With the
// block column J partitioned as
//   R = L[J, :J0]   D = L[J, J]   B = L[J1:, :J0]   C = L[J1:, J]
// and the same slices of the adjoint buffer (Rb, Db, Bb, Cb), blocks are
// visited from the last to the first:
//   Cb = Cb D^-1                    (TRSM, RIGHT | LOWER)
//   Bb -= Cb R                      (GEMM)
//   Db -= tril(Cb^T C)              (GEMM on the lower triangle)
//   Db = unblocked reverse sweep of chol(D) applied to Db
//   Rb -= Cb^T B + (Db + Db^T) R    (two GEMMs)
// Every step reads L and writes only the adjoint buffer, whose lower
// triangle then holds the gradient with respect to the lower triangle of
// A. A last pass turns that into the symmetric gradient: off-diagonal
// entries are shared by (i, j) and (j, i), so each gets half.
//
// Memory: L and the adjoint buffer, plus O(nb * n) of workspace for the
// triangular solve and one nb x nb block.
#include "mops_solvers.hpp"

namespace mops {
namespace {

constexpr int ADJ_NB = 64;  // diagonal block order (the unblocked kernel's limit)
constexpr int ADJ_LD = ADJ_NB + 1;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Unblocked reverse sweep on one jb x jb diagonal block (jb <= 64), in LDS.
// Column j, from the last to the first, with r = D[j, :j], d = D[j][j],
// Bp = D[j+1:, :j], c = D[j+1:, j] and the same slices of Db:
//   db -= c . cb / d;  db /= d;  cb /= d
//   rb -= db r + cb^T Bp;  Bpb -= cb r;  db /= 2
// Wave 0 does the dot product and the scaling (phase A); all 256 threads
// then update row j (the matrix-vector product split four ways over the
// rows) and the block below it (phase B). S receives Db + Db^T (dense).
__global__ __launch_bounds__(256) void k_chol_adj_diag(int jb, const double* __restrict__ D,
                                                       long ldl, double* __restrict__ Db,
                                                       long lda, double* __restrict__ S) {
  __shared__ double Ls[ADJ_NB * ADJ_LD];
  __shared__ double Bs[ADJ_NB * ADJ_LD];
  __shared__ double part[4][ADJ_NB];
  __shared__ double dbar;
  const int t = threadIdx.x;
  for (int idx = t; idx < ADJ_NB * ADJ_NB; idx += 256) {
    const int i = idx % ADJ_NB, k = idx / ADJ_NB;
    const bool in = i < jb && k < jb && k <= i;
    Ls[i * ADJ_LD + k] = in ? D[i + (long)k * ldl] : 0.0;
    Bs[i * ADJ_LD + k] = in ? Db[i + (long)k * lda] : 0.0;
  }
  __syncthreads();
  const int k = t & 63, g = t >> 6;
  for (int j = jb - 1; j >= 0; --j) {
    if (t < 64) {
      const int i = t;
      const double d = Ls[j * ADJ_LD + j];
      const bool below = i > j && i < jb;
      const double p = wave_sum(below ? Ls[i * ADJ_LD + j] * Bs[i * ADJ_LD + j] : 0.0);
      const double db = (Bs[j * ADJ_LD + j] - p / d) / d;
      if (below) Bs[i * ADJ_LD + j] /= d;
      if (t == 0) dbar = db;
    }
    __syncthreads();
    const double db = dbar;
    // one pass over the rows below j: thread (k, g) takes column k < j and
    // rows j+1+g, j+5+g, ...; (i) its partial sum of cb^T Bp for row j and
    // (ii) Bpb -= cb r on its elements (column j is only read here)
    double acc = 0.0;
    if (k < j) {
      const double lj = Ls[j * ADJ_LD + k];
      for (int i = j + 1 + g; i < jb; i += 4) {
        const double b = Bs[i * ADJ_LD + j];
        acc += b * Ls[i * ADJ_LD + k];
        Bs[i * ADJ_LD + k] -= b * lj;
      }
    }
    part[g][k] = acc;
    __syncthreads();
    // row j and the diagonal; wave 0 reads them next, in program order
    if (t < j)
      Bs[j * ADJ_LD + t] -= db * Ls[j * ADJ_LD + t] + part[0][t] + part[1][t] + part[2][t] +
                            part[3][t];
    if (t == 0) Bs[j * ADJ_LD + j] = 0.5 * db;
  }
  __syncthreads();
  for (int idx = t; idx < jb * jb; idx += 256) {
    const int i = idx % jb, c = idx / jb;
    if (c <= i) Db[i + (long)c * lda] = Bs[i * ADJ_LD + c];
    S[i + (long)c * jb] = c <= i ? Bs[i * ADJ_LD + c] + (c == i ? Bs[i * ADJ_LD + c] : 0.0)
                                 : Bs[c * ADJ_LD + i];
  }
}

// Lower-triangle gradient -> symmetric gradient, in place: for i > j,
// A[i][j] = A[j][i] = A[i][j] / 2. One 32 x 32 tile pair per workgroup.
__global__ __launch_bounds__(256) void k_adj_symmetrize(int n, double* A, long lda) {
  __shared__ double tile[32][33];
  const int bi = blockIdx.x, bj = blockIdx.y;  // tile row, tile column
  if (bj > bi) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int c = ty; c < 32; c += 8) {
    const int i = bi * 32 + tx, j = bj * 32 + c;
    double v = 0.0;
    if (i < n && j < n && i > j) {
      v = 0.5 * A[i + (long)j * lda];
      A[i + (long)j * lda] = v;
    }
    tile[c][tx] = v;
  }
  __syncthreads();
  // upper tile (bj, bi): A[j][i] for j in tile bj (rows), i in tile bi (columns)
  for (int c = ty; c < 32; c += 8) {
    const int j = bj * 32 + tx, i = bi * 32 + c;
    if (i < n && j < n && i > j) A[j + (long)i * lda] = tile[tx][c];
  }
}

}  // namespace

size_t chol_adjoint_ws_bytes(int n, int nb) {
  if (n <= 0) return 0;
  if (nb <= 0 || nb > ADJ_NB) nb = ADJ_NB;
  // the TN products of the sweep may split their long k (output 64 x n, k < n)
  const size_t split = gemm_split_ws_bytes(ADJ_NB, n, n, sizeof(double)) +
                       gemm_split_ws_bytes(ADJ_NB, ADJ_NB, n, sizeof(double));
  // the inverses of all diagonal blocks (and their construction), the
  // solution of one block column
  const size_t inv = (size_t)ceil_div(n, nb) * nb * nb * sizeof(double) +
                     trtri_blocks_ws_bytes(n, nb);
  const size_t w = (size_t)n * nb * sizeof(double);
  return (size_t)ADJ_NB * ADJ_NB * sizeof(double) + inv + w + split + 4096;
}

int chol_adjoint_d(hipStream_t s, int n, int nb, const double* L, long ldl, double* Ab,
                   long lda) {
  if (n <= 0) return ST_OK;
  if (nb <= 0 || nb > ADJ_NB) nb = ADJ_NB;
  double* S = (double*)ws_carve((size_t)ADJ_NB * ADJ_NB * sizeof(double));
  // inverses of every diagonal block of L, all in one launch up front (they
  // depend on L only): each block step's triangular solve is then one GEMM
  const int nblk = ceil_div(n, nb);
  double* Dinv = (double*)ws_carve((size_t)nblk * nb * nb * sizeof(double));
  double* W = (double*)ws_carve((size_t)n * nb * sizeof(double));
  if (!S || !Dinv || !W) {
    set_error("chol_adjoint: workspace too small");
    return ST_NOMEM;
  }
  {
    const size_t mark = ws_mark();
    const int st = trtri_blocks_d(s, true, false, n, nb, L, ldl, Dinv);
    ws_release(mark);
    MOPS_TRY(st);
  }
  hipStream_t s2 = aux_stream();
  hipEvent_t e_db = sync_event(0), e_sweep = sync_event(1);
  if (!s2 || !e_db || !e_sweep) {
    set_error("chol_adjoint: could not create the second stream");
    return ST_HIP;
  }
  for (int J0 = ((n - 1) / nb) * nb; J0 >= 0; J0 -= nb) {
    const int J1 = min(J0 + nb, n), jb = J1 - J0, m = n - J1;
    const double* D = L + J0 + (long)J0 * ldl;
    const double* R = L + J0;
    const double* Bm = L + J1;
    const double* C = L + J1 + (long)J0 * ldl;
    double* Db = Ab + J0 + (long)J0 * lda;
    double* Rb = Ab + J0;
    double* Bb = Ab + J1;
    double* Cb = Ab + J1 + (long)J0 * lda;
    if (m > 0) {
      // Cb = Cb D^-1
      const double* Di = Dinv + (long)(J0 / nb) * nb * nb;
      MOPS_TRY(gemm_d(s, false, false, m, jb, jb, 1.0, Cb, lda, Di, nb, 0.0, W, m, GM_FULL,
                      TRI_B_LOWER));
      MOPS_TRY(copy_tile_d(s, m, jb, W, m, Cb, lda));
      MOPS_TRY(gemm_d(s, true, false, jb, jb, m, -1.0, Cb, lda, C, ldl, 1.0, Db, lda, GM_LOWER));
    }
    // The diagonal sweep (one workgroup, latency-bound) runs on the second
    // stream beside the products that do not need its result; both become
    // ready at the same moment, so the sweep gets its CU first.
    MOPS_CHECK_HIP(hipEventRecord(e_db, s));
    MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_db, 0));
    hipLaunchKernelGGL(k_chol_adj_diag, dim3(1), dim3(256), 0, s2, jb, D, ldl, Db, lda, S);
    MOPS_CHECK_HIP(hipGetLastError());
    MOPS_CHECK_HIP(hipEventRecord(e_sweep, s2));
    if (J0 > 0 && m > 0) {
      MOPS_TRY(gemm_d(s, false, false, m, J0, jb, -1.0, Cb, lda, R, ldl, 1.0, Bb, lda));
      MOPS_TRY(gemm_d(s, true, false, jb, J0, m, -1.0, Cb, lda, Bm, ldl, 1.0, Rb, lda));
    }
    MOPS_CHECK_HIP(hipStreamWaitEvent(s, e_sweep, 0));
    if (J0 > 0) MOPS_TRY(gemm_d(s, false, false, jb, J0, jb, -1.0, S, jb, R, ldl, 1.0, Rb, lda));
  }
  const int nt = ceil_div(n, 32);
  hipLaunchKernelGGL(k_adj_symmetrize, dim3(nt, nt), dim3(256), 0, s, n, Ab, lda);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

}  // namespace mops
