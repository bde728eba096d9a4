// The other BLAS-3 operations of matops' front end on the GPU GEMM:
//
//   symm_d    C = beta C + alpha sym(A) B  (LEFT) or alpha B sym(A) (RIGHT)
//             (MultSym, mmult.go:266-326 -> dmult_symm_blocked2,
//             calgo/dmmat_symm.c:253)
//   trmm_d    B = alpha op(A) B (LEFT) or alpha B op(A) (RIGHT), A triangular
//             (MultTrm, mmult.go:338-372 -> dmmat_trmm_blk,
//             calgo/dmmat_trid.c:780)
//   syr2k_d   C = beta C + alpha (op(A) op(B)^T + op(B) op(A)^T), one triangle
//             (RankUpdate2Sym, mmult.go:453-476 -> dmmat_rank2_blk,
//             calgo/dmmat_rank.c:132)
//   trmupd_d  C = beta C + alpha op(A) op(B), one triangle
//             (UpdateTrm, mmult.go:501-520 -> dmmat_trmupd_blk,
//             calgo/dmmat_rank.c:292)
//
// Design: the symmetric operand of symm is first expanded into a dense
// matrix (one tiled pass, 16 bytes of traffic per element, against 2 m n^2
// flops of product), the triangle of trmm into a dense triangle with stored
// zeros (so the GEMM skips the zero k-slabs); syr2k and trmupd are the
// GEMM's triangle-only instance, two launches for syr2k. Everything runs on
// the FP64 matrix cores through gemm_d.
#include "mops_solvers.hpp"

namespace mops {
namespace {

constexpr int XT = 32;  // expansion tile edge

// Out (n x n) = the symmetric matrix whose lower (lower != 0) or upper
// triangle is stored in A. Workgroup = one 32x32 output tile; tiles wholly
// in the stored triangle are copied, tiles wholly in the other one are read
// transposed through LDS, diagonal tiles element by element.
__global__ __launch_bounds__(256) void k_sym_expand(int n, const double* __restrict__ A,
                                                    long lda, int lower,
                                                    double* __restrict__ Out, long ldo) {
  __shared__ double t[XT][XT + 1];
  const int i0 = blockIdx.x * XT, j0 = blockIdx.y * XT;
  const int tx = threadIdx.x % XT, ty = threadIdx.x / XT;  // ty in 0..7
  const bool stored_all = lower ? (i0 >= j0 + XT - 1) : (j0 >= i0 + XT - 1);
  const bool mirror_all = lower ? (j0 > i0 + XT - 1) : (i0 > j0 + XT - 1);
  if (stored_all) {
    const int i = i0 + tx;
    for (int r = ty; r < XT; r += 8) {
      const int j = j0 + r;
      if (i < n && j < n) Out[i + (long)j * ldo] = A[i + (long)j * lda];
    }
  } else if (mirror_all) {
    // t[tx][r] = A[j0 + tx][i0 + r]; Out[i0 + tx][j0 + r] = A[j0 + r][i0 + tx] = t[r][tx]
    for (int r = ty; r < XT; r += 8) {
      const int ar = j0 + tx, ac = i0 + r;
      t[tx][r] = (ar < n && ac < n) ? A[ar + (long)ac * lda] : 0.0;
    }
    __syncthreads();
    const int i = i0 + tx;
    for (int r = ty; r < XT; r += 8) {
      const int j = j0 + r;
      if (i < n && j < n) Out[i + (long)j * ldo] = t[r][tx];
    }
  } else {
    const int i = i0 + tx;
    for (int r = ty; r < XT; r += 8) {
      const int j = j0 + r;
      if (i >= n || j >= n) continue;
      const bool stored = lower ? i >= j : i <= j;
      Out[i + (long)j * ldo] = stored ? A[i + (long)j * lda] : A[j + (long)i * lda];
    }
  }
}

// Out (n x n) = the lower (lower != 0) or upper triangle of A with zeros in
// the other one; unit: ones on the diagonal, which is never read.
__global__ __launch_bounds__(256) void k_tri_expand(int n, const double* __restrict__ A, long lda,
                                                    int lower, int unit,
                                                    double* __restrict__ Out, long ldo) {
  const int i = blockIdx.x * XT + threadIdx.x % XT;
  const int j0 = blockIdx.y * XT;
  if (i >= n) return;
  for (int r = threadIdx.x / XT; r < XT; r += 8) {
    const int j = j0 + r;
    if (j >= n) break;
    double v = 0.0;
    if (i == j)
      v = unit ? 1.0 : A[i + (long)j * lda];
    else if (lower ? i > j : i < j)
      v = A[i + (long)j * lda];
    Out[i + (long)j * ldo] = v;
  }
}

dim3 tiles_of(int n) { return dim3((unsigned)ceil_div(n, XT), (unsigned)ceil_div(n, XT)); }

}  // namespace

size_t symm_ws_bytes(int n) { return n > 0 ? (size_t)n * n * sizeof(double) + 512 : 0; }

int sym_expand_d(hipStream_t s, bool lower, int n, const double* A, long lda, double* Out,
                 long ldo) {
  if (n <= 0) return ST_OK;
  hipLaunchKernelGGL(k_sym_expand, tiles_of(n), dim3(256), 0, s, n, A, lda, lower ? 1 : 0, Out,
                     ldo);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

int symm_d(hipStream_t s, bool left, bool lower, int M, int N, double alpha, const double* A,
           long lda, const double* B, long ldb, double beta, double* C, long ldc) {
  if (M <= 0 || N <= 0) return ST_OK;
  const int na = left ? M : N;
  double* S = (double*)ws_carve((size_t)na * na * sizeof(double));
  if (!S) {
    set_error("symm: workspace too small");
    return ST_NOMEM;
  }
  MOPS_TRY(sym_expand_d(s, lower, na, A, lda, S, na));
  if (left) return gemm_d(s, false, false, M, N, M, alpha, S, na, B, ldb, beta, C, ldc);
  return gemm_d(s, false, false, M, N, N, alpha, B, ldb, S, na, beta, C, ldc);
}

size_t trmm_ws_bytes(int N, int nrhs) {
  if (N <= 0 || nrhs <= 0) return 0;
  return (size_t)N * N * sizeof(double) + (size_t)N * nrhs * sizeof(double) + 1024 +
         gemm_split_ws_bytes(N, nrhs, N, sizeof(double)) +
         gemm_split_ws_bytes(nrhs, N, N, sizeof(double));
}

int trmm_d(hipStream_t s, int flags, int N, int nrhs, double alpha, const double* A, long lda,
           double* B, long ldb) {
  if (N <= 0 || nrhs <= 0) return ST_OK;
  const bool right = flags & F_RIGHT;
  const bool upper = flags & F_UPPER;  // dmmat_trmm_blk tests UPPER first (:799-827)
  const bool trans = flags & F_TRANSA;
  const bool unit = flags & F_UNIT;
  double* T = (double*)ws_carve((size_t)N * N * sizeof(double));
  double* W = (double*)ws_carve((size_t)N * nrhs * sizeof(double));
  if (!T || !W) {
    set_error("trmm: workspace too small");
    return ST_NOMEM;
  }
  hipLaunchKernelGGL(k_tri_expand, tiles_of(N), dim3(256), 0, s, N, A, lda, upper ? 0 : 1,
                     unit ? 1 : 0, T, (long)N);
  MOPS_CHECK_HIP(hipGetLastError());
  // op(T) is lower triangular when (lower, no trans) or (upper, trans)
  const bool op_lower = upper == trans;
  if (!right) {
    // W = alpha op(T) B, B N x nrhs
    MOPS_TRY(gemm_d(s, trans, false, N, nrhs, N, alpha, T, N, B, ldb, 0.0, W, N, GM_FULL,
                    op_lower ? TRI_A_LOWER : TRI_A_UPPER));
    return copy_tile_d(s, N, nrhs, W, N, B, ldb);
  }
  // W = alpha B op(T), B nrhs x N
  MOPS_TRY(gemm_d(s, false, trans, nrhs, N, N, alpha, B, ldb, T, N, 0.0, W, nrhs, GM_FULL,
                  op_lower ? TRI_B_LOWER : TRI_B_UPPER));
  return copy_tile_d(s, nrhs, N, W, nrhs, B, ldb);
}

int syr2k_d(hipStream_t s, int flags, int N, int P, double alpha, const double* A, long lda,
            const double* B, long ldb, double beta, double* C, long ldc) {
  if (N <= 0) return ST_OK;
  const bool ta = flags & F_TRANSA;
  const int mode = (flags & F_LOWER) ? GM_LOWER : GM_UPPER;
  // op(A) op(B)^T: NOTRANS A B^T, TRANSA A^T B
  MOPS_TRY(gemm_d(s, ta, !ta, N, N, P, alpha, A, lda, B, ldb, beta, C, ldc, mode));
  return gemm_d(s, ta, !ta, N, N, P, alpha, B, ldb, A, lda, 1.0, C, ldc, mode);
}

int trmupd_d(hipStream_t s, int flags, int N, int P, double alpha, const double* A, long lda,
             const double* B, long ldb, double beta, double* C, long ldc) {
  if (N <= 0) return ST_OK;
  const int mode = (flags & F_UPPER) ? GM_UPPER : GM_LOWER;
  return gemm_d(s, (flags & F_TRANSA) != 0, (flags & F_TRANSB) != 0, N, N, P, alpha, A, lda, B,
                ldb, beta, C, ldc, mode);
}

}  // namespace mops
