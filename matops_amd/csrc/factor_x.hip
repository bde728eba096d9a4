// Factorizations without pivoting and the triangular inverse, built on the
// GPU GEMM and triangular solves:
//
//   getrf_nopiv_d  A = L U, no row interchanges (DecomposeLUnoPiv,
//                  lu.go:38-96, 291-299)
//   sytrf_nopiv_d  A = L D L^T or U D U^T, no pivoting (DecomposeLDLnoPiv,
//                  ldlnp.go:23-241)
//   inverse_trm_d  A = inv(A) for a triangular A in place (InverseTrm,
//                  trinv.go:16-245)
//
// Design: recursive halving (Gustavson's recursive LU/LDL): an order-n block
// is split into n1 = a multiple of 64 near n/2 and n - n1; the leading part is
// factored recursively, the off-diagonal block is one triangular solve, the
// trailing block one GEMM (LU) or one triangle-only GEMM (LDL), then the
// trailing part is factored recursively. Blocks of at most 64 columns are
// factored by one workgroup in LDS. The products then have inner dimension
// n/2, n/4, ..., so most of the flops run at full-tile MFMA shapes; the
// results equal the reference's to rounding (both factorizations are unique
// without pivoting), with divisions where the reference divides.
#include "mops_solvers.hpp"

namespace mops {
namespace {

constexpr int DB = 64;        // order of the LDS-resident diagonal blocks
constexpr int DBP = DB + 1;   // padded LDS row

// LU without pivoting of the r x c block A (r, c <= 64) in LDS, right
// looking (lu.go:38-61): per column j, a21 /= a11 (division), A22 -= a21 a12.
__global__ __launch_bounds__(256) void k_lu_nopiv_leaf(int r, int c, double* A, long lda) {
  __shared__ double a[DB * DBP];  // a[col * DBP + row]
  const int t = threadIdx.x;
  for (int e = t; e < DB * DB; e += 256) {
    const int i = e % DB, j = e / DB;
    a[j * DBP + i] = (i < r && j < c) ? A[i + (long)j * lda] : 0.0;
  }
  __syncthreads();
  const int kmax = r < c ? r : c;
  for (int j = 0; j < kmax; ++j) {
    const double piv = a[j * DBP + j];
    for (int i = j + 1 + t; i < r; i += 256) a[j * DBP + i] = a[j * DBP + i] / piv;
    __syncthreads();
    // rank-1 update of the trailing (r-j-1) x (c-j-1) block
    const int nr = r - j - 1, nc = c - j - 1;
    for (int e = t; e < nr * nc; e += 256) {
      const int i = j + 1 + e % nr, q = j + 1 + e / nr;
      a[q * DBP + i] -= a[j * DBP + i] * a[q * DBP + j];
    }
    __syncthreads();
  }
  for (int e = t; e < DB * DB; e += 256) {
    const int i = e % DB, j = e / DB;
    if (i < r && j < c) A[i + (long)j * lda] = a[j * DBP + i];
  }
}

// LDL^T without pivoting of the n x n block A (n <= 64) in LDS. LOWER
// (ldlnp.go:23-61): per column j from the first, A22 -= a21 a21^T / a11 on
// the lower triangle, then l21 = a21 / a11. UPPER (ldlnp.go:64-101): per
// column j from the last, A00 -= a01 a01^T / a11 on the upper triangle, then
// u01 = a01 / a11. The other triangle is not written.
__global__ __launch_bounds__(256) void k_ldl_nopiv_leaf(int n, int lower, double* A, long lda) {
  __shared__ double a[DB * DBP];
  const int t = threadIdx.x;
  for (int e = t; e < DB * DB; e += 256) {
    const int i = e % DB, j = e / DB;
    a[j * DBP + i] = (i < n && j < n) ? A[i + (long)j * lda] : 0.0;
  }
  __syncthreads();
  for (int s = 0; s < n; ++s) {
    const int j = lower ? s : n - 1 - s;
    const double d = a[j * DBP + j];
    const double f = -1.0 / d;
    if (lower) {
      const int m = n - j - 1;
      for (int e = t; e < m * m; e += 256) {
        const int i = j + 1 + e % m, q = j + 1 + e / m;
        if (i >= q) a[q * DBP + i] += f * a[j * DBP + i] * a[j * DBP + q];
      }
      __syncthreads();
      for (int i = j + 1 + t; i < n; i += 256) a[j * DBP + i] = a[j * DBP + i] / d;
    } else {
      // a01 is column j above the diagonal: a[j * DBP + i], i < j
      const int m = j;
      for (int e = t; e < m * m; e += 256) {
        const int i = e % m, q = e / m;
        if (i <= q) a[q * DBP + i] += f * a[j * DBP + i] * a[j * DBP + q];
      }
      __syncthreads();
      for (int i = t; i < j; i += 256) a[j * DBP + i] = a[j * DBP + i] / d;
    }
    __syncthreads();
  }
  for (int e = t; e < DB * DB; e += 256) {
    const int i = e % DB, j = e / DB;
    if (i < n && j < n && (lower ? i >= j : i <= j)) A[i + (long)j * lda] = a[j * DBP + i];
  }
}

// X = X * diag(A)^-1 for an r x c X: column j times 1 / A[off + j][off + j],
// the reciprocal multiply of SolveDiag(X, D, RIGHT) (diag.go:96-101)
__global__ void k_div_cols_by_diag(int r, int c, double* X, long ldx, const double* A, long lda,
                                   int off) {
  const long total = (long)r * c;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % r), j = (int)(e / r);
    X[i + (long)j * ldx] *= 1.0 / A[(off + j) * (lda + 1)];
  }
}

// B = diag(A)^-1 B for an n x nrhs B: row i times 1 / A[i][i], the
// reciprocal multiply of SolveDiag(B, A, LEFT) with a matrix D
// (diag.go:111-116)
__global__ void k_div_rows_by_diag(int n, int nrhs, double* B, long ldb, const double* A,
                                   long lda) {
  const long total = (long)n * nrhs;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    B[i + (long)j * ldb] *= 1.0 / A[i * (lda + 1)];
  }
}

// copies the lower (lower != 0) or upper triangle of the n x n X into A;
// with the diagonal iff diag != 0
__global__ void k_copy_triangle(int n, int lower, int diag, const double* X, long ldx, double* A,
                                long lda) {
  const long total = (long)n * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    const bool in = i == j ? diag != 0 : (lower ? i > j : i < j);
    if (in) A[i + (long)j * lda] = X[i + (long)j * ldx];
  }
}

inline int grid_for(long total) {
  long b = (total + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

// split point of an order-n block: a multiple of 64 near n / 2
inline int split_at(int n) {
  int n1 = (n / 2 + DB - 1) / DB * DB;
  if (n1 >= n) n1 = n - DB;
  return n1;
}

int lu_nopiv_rec(hipStream_t s, int m, int n, double* A, long lda) {
  if (m <= 0 || n <= 0) return ST_OK;
  const int k = m < n ? m : n;
  if (m <= DB && n <= DB) {
    hipLaunchKernelGGL(k_lu_nopiv_leaf, dim3(1), dim3(256), 0, s, m, n, A, lda);
    MOPS_CHECK_HIP(hipGetLastError());
    return ST_OK;
  }
  // columns [0, n1) then the rest; for a tall or wide block of <= 64
  // columns (rows) the leaf is applied to the square part and the rest is
  // one triangular solve
  const int n1 = k <= DB ? k : split_at(k);
  MOPS_TRY(lu_nopiv_rec(s, n1, n1, A, lda));  // A11 = L11 U11
  double* A12 = A + (long)n1 * lda;
  double* A21 = A + n1;
  double* A22 = A + n1 + (long)n1 * lda;
  // L21 = A21 U11^-1, U12 = L11^-1 A12 (lu.go:86-89)
  if (m > n1) MOPS_TRY(trsm_d(s, F_RIGHT | F_UPPER, n1, m - n1, 1.0, A, lda, A21, lda));
  if (n > n1) MOPS_TRY(trsm_d(s, F_LEFT | F_LOWER | F_UNIT, n1, n - n1, 1.0, A, lda, A12, lda));
  if (m > n1 && n > n1) {
    MOPS_TRY(gemm_d(s, false, false, m - n1, n - n1, n1, -1.0, A21, lda, A12, lda, 1.0, A22, lda));
    MOPS_TRY(lu_nopiv_rec(s, m - n1, n - n1, A22, lda));
  }
  return ST_OK;
}

// LOWER: A = [A11 .; A21 A22], A11 n1 x n1 factored first (ldlnp.go:104-145).
// UPPER: A = [A00 A01; . A11], the trailing A11 (n - n0) factored first
// (ldlnp.go:148-208). W: n x n/2 workspace for L21 D (or U01 D).
int ldl_nopiv_rec(hipStream_t s, bool lower, int n, double* A, long lda, double* W) {
  if (n <= 0) return ST_OK;
  if (n <= DB) {
    hipLaunchKernelGGL(k_ldl_nopiv_leaf, dim3(1), dim3(256), 0, s, n, lower ? 1 : 0, A, lda);
    MOPS_CHECK_HIP(hipGetLastError());
    return ST_OK;
  }
  const int n1 = split_at(n), n2 = n - n1;
  if (lower) {
    double* A21 = A + n1;
    double* A22 = A + n1 + (long)n1 * lda;
    MOPS_TRY(ldl_nopiv_rec(s, true, n1, A, lda, W));
    // A21 = A21 L11^-T; W = A21 (= L21 D1); A21 = A21 D1^-1;
    // A22 -= L21 W^T on the lower triangle
    MOPS_TRY(trsm_d(s, F_RIGHT | F_LOWER | F_UNIT | F_TRANSA, n1, n2, 1.0, A, lda, A21, lda));
    MOPS_TRY(copy_tile_d(s, n2, n1, A21, lda, W, n2));
    hipLaunchKernelGGL(k_div_cols_by_diag, dim3(grid_for((long)n2 * n1)), dim3(256), 0, s, n2, n1,
                       A21, lda, A, lda, 0);
    MOPS_CHECK_HIP(hipGetLastError());
    MOPS_TRY(gemm_d(s, false, true, n2, n2, n1, -1.0, A21, lda, W, n2, 1.0, A22, lda, GM_LOWER));
    return ldl_nopiv_rec(s, true, n2, A22, lda, W);
  }
  // UPPER: trailing block first
  const int n0 = n2;  // leading part, factored last
  const int nt = n1;  // trailing part
  double* A01 = A + (long)n0 * lda;
  double* A11 = A + n0 + (long)n0 * lda;
  MOPS_TRY(ldl_nopiv_rec(s, false, nt, A11, lda, W));
  MOPS_TRY(trsm_d(s, F_RIGHT | F_UPPER | F_UNIT | F_TRANSA, nt, n0, 1.0, A11, lda, A01, lda));
  MOPS_TRY(copy_tile_d(s, n0, nt, A01, lda, W, n0));
  hipLaunchKernelGGL(k_div_cols_by_diag, dim3(grid_for((long)n0 * nt)), dim3(256), 0, s, n0, nt,
                     A01, lda, A11, lda, 0);
  MOPS_CHECK_HIP(hipGetLastError());
  MOPS_TRY(gemm_d(s, false, true, n0, n0, nt, -1.0, A01, lda, W, n0, 1.0, A, lda, GM_UPPER));
  return ldl_nopiv_rec(s, false, n0, A, lda, W);
}

}  // namespace

size_t getrf_nopiv_ws_bytes(int m, int n) {
  const int k = m < n ? m : n;
  if (k <= 0) return 0;
  const long big = m > n ? m : n;
  return trsm_ws_bytes(F_RIGHT, k, (int)big) + gemm_split_ws_bytes(big, big, k, sizeof(double)) +
         4096;
}

int getrf_nopiv_d(hipStream_t s, int m, int n, double* A, long lda) {
  return lu_nopiv_rec(s, m, n, A, lda);
}

size_t sytrf_nopiv_ws_bytes(int n) {
  if (n <= 0) return 0;
  return (size_t)n * (n / 2 + DB) * sizeof(double) + trsm_ws_bytes(F_RIGHT, n, n) +
         gemm_split_ws_bytes(n, n, n, sizeof(double)) + 4096;
}

int sytrf_nopiv_d(hipStream_t s, bool lower, int n, double* A, long lda) {
  if (n <= 0) return ST_OK;
  double* W = (double*)ws_carve((size_t)n * (n / 2 + DB) * sizeof(double));
  if (!W) {
    set_error("ldl: workspace too small");
    return ST_NOMEM;
  }
  return ldl_nopiv_rec(s, lower, n, A, lda, W);
}

size_t sytrs_nopiv_ws_bytes(int n, int nrhs) { return trsm_ws_bytes(F_LEFT, n, nrhs) + 1024; }

int sytrs_nopiv_d(hipStream_t s, bool lower, int n, int nrhs, const double* A, long lda,
                  double* B, long ldb) {
  if (n <= 0 || nrhs <= 0) return ST_OK;
  // SolveLDLnoPiv (ldlnp.go:251-264): unit-triangle solve, diagonal, then
  // the transposed unit-triangle solve
  const int tri = lower ? F_LOWER : F_UPPER;
  MOPS_TRY(trsm_d(s, F_LEFT | tri | F_UNIT, n, nrhs, 1.0, A, lda, B, ldb));
  hipLaunchKernelGGL(k_div_rows_by_diag, dim3(grid_for((long)n * nrhs)), dim3(256), 0, s, n, nrhs,
                     B, ldb, A, lda);
  MOPS_CHECK_HIP(hipGetLastError());
  return trsm_d(s, F_LEFT | tri | F_UNIT | F_TRANSA, n, nrhs, 1.0, A, lda, B, ldb);
}

size_t inverse_trm_ws_bytes(int n) {
  if (n <= 0) return 0;
  return (size_t)n * n * sizeof(double) + trtri_ws_bytes(n) + 1024;
}

int inverse_trm_d(hipStream_t s, bool lower, bool unit, int n, double* A, long lda) {
  if (n <= 0) return ST_OK;
  double* X = (double*)ws_carve((size_t)n * n * sizeof(double));
  if (!X) {
    set_error("inverse: workspace too small");
    return ST_NOMEM;
  }
  MOPS_TRY(trtri_d(s, lower, unit, n, A, lda, X, n));
  // the inverse's triangle back in place; with UNIT the stored diagonal is
  // neither read nor written (trinv.go:86-143 never touch a11)
  hipLaunchKernelGGL(k_copy_triangle, dim3(grid_for((long)n * n)), dim3(256), 0, s, n,
                     lower ? 1 : 0, unit ? 0 : 1, X, (long)n, A, lda);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

}  // namespace mops
