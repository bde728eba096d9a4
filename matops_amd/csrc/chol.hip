// Blocked Cholesky factorization and solve on the GPU.
//
// Replaces the panel loop of blockedCHOL / unblockedCHOL (chol.go:17-94),
// DecomposeCHOL (chol.go:116-127) and SolveCHOL (chol.go:147-158). The
// reference's step for a diagonal block of width nb is
//   A11 = chol(A11)  (per column: sqrt, a21 /= a11, SYR; chol.go:29-46)
//   A21 = A21 * tril(A11)^-T          (SolveTrm RIGHT|LOWER|TRANSA, :79)
//   A22 = A22 - A21 * A21^T           (RankUpdateSym LOWER, :81)
// and the UPPER variant with A12 / U11^-T (:82-87). The same step runs here
// with the diagonal block factored recursively: leaves of up to 64x64 are
// factored in registers by one workgroup (the reference's column algorithm)
// which also returns the leaf's inverse; larger diagonal blocks are split in two and
// joined with GEMMs, building inv(L11) on the way so that the panel solve is
// one GEMM A21 = A21 * inv(L11)^T.
#include <cmath>

#include "leaf.hpp"
#include "mops_solvers.hpp"

namespace mops {
namespace {

constexpr int LEAF = 64;

// Cholesky of an n x n (n <= NL) diagonal block plus the inverse of the
// factor, one workgroup of NL*LPC threads, the block held in registers
// (layout in leaf.hpp). Per column j: the LPC lanes owning column j take the
// square root (NaN -> info, chol.go:33-36), divide the rest of the column
// by it (InvScale, chol.go:42) and publish it through LDS; after one barrier
// every lane applies the rank-1 update to its part of the trailing
// triangle (the SYR of chol.go:44). The UPPER case factors U^T = L.
template <int NL, int LPC>
__global__ __launch_bounds__(NL* LPC) void k_potf2_leaf(int n, int upper, double* A, long lda,
                                                        double* X, long ldx, int goff, int* info,
                                                        double* badval) {
  constexpr int RPL = NL / LPC;
  constexpr int LDL = NL + 1;
  __shared__ double Ls[NL * LDL];
  __shared__ double col[2][NL];
  const int t = threadIdx.x;
  const int c = t / LPC, sub = t % LPC;
  const int lane = t & 63, base = lane & ~(LPC - 1);
  double v[RPL];
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int r = sub + LPC * q;
    double a = 0.0;
    if (r < n && c < n && r >= c) a = upper ? A[c + (long)r * lda] : A[r + (long)c * lda];
    v[q] = a;
  }
  bool bad = false;
  for (int j = 0; j < n; ++j) {
    double* cj = col[j & 1];
    if (c == j) {
      const double d = __shfl(leaf::pick<RPL>(v, j / LPC), base | (j % LPC), 64);
      const double r = sqrt(d);
      if (isnan(r) && !bad && sub == j % LPC && info) {
        bad = true;
        const int old = atomicMin(info, goff + j);
        if ((old < 0 || old > goff + j) && badval) *badval = d;
      }
#pragma unroll
      for (int q = 0; q < RPL; ++q) {
        const int rr = sub + LPC * q;
        if (rr > j) v[q] = v[q] / r;
        if (rr == j) v[q] = r;
        if (rr < NL) cj[rr] = rr >= j ? v[q] : 0.0;
      }
    }
    __syncthreads();
    if (c > j && c < n) {
      const double lc = cj[c];
#pragma unroll
      for (int q = 0; q < RPL; ++q) {
        const int rr = sub + LPC * q;
        if (rr >= c && rr < n) v[q] -= cj[rr] * lc;
      }
    }
  }
  // publish the factor: to global memory and to LDS for the inverse
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int r = sub + LPC * q;
    if (c < NL) Ls[r * LDL + c] = v[q];
    if (r < n && c < n && r >= c) {
      if (upper)
        A[c + (long)r * lda] = v[q];
      else
        A[r + (long)c * lda] = v[q];
    }
  }
  __syncthreads();
  double x[RPL];
  leaf::inv_lower_column<NL, LPC, LDL>(Ls, n, c, sub, false, x);
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int r = sub + LPC * q;
    if (r < n && c < n) {
      if (upper)
        X[c + (long)r * ldx] = x[q];  // inv(U) = inv(L)^T, zeros below
      else
        X[r + (long)c * ldx] = x[q];  // inv(L), zeros above
    }
  }
}

__global__ void k_init_info(int* info, double* badval) {
  if (info) *info = 0x7fffffff;
  if (badval) *badval = 0.0;
}

__global__ void k_fix_info(int* info) {
  if (info && *info == 0x7fffffff) *info = -1;
}

inline int split_point(int n) {
  // first half: a multiple of the leaf size, about half of n
  int h = ((n / 2 + LEAF - 1) / LEAF) * LEAF;
  if (h >= n) h = n - LEAF;
  return h;
}

// Factor the n x n diagonal block at A (global offset goff) and write the
// dense inverse of its factor into X (n x n, ld ldx). T: scratch of n*n.
int chol_diag(hipStream_t s, bool upper, int n, double* A, long lda, double* X, long ldx, int goff,
              int* info, double* badval, double* T) {
  if (n <= LEAF) {
    hipLaunchKernelGGL((k_potf2_leaf<64, 4>), dim3(1), dim3(256), 0, s, n, upper ? 1 : 0, A, lda,
                       X, ldx, goff, info, badval);
    MOPS_CHECK_HIP(hipGetLastError());
    return ST_OK;
  }
  const int h = split_point(n), m = n - h;
  double* X11 = X;
  double* X22 = X + h + (long)h * ldx;
  MOPS_TRY(chol_diag(s, upper, h, A, lda, X11, ldx, goff, info, badval, T));
  if (!upper) {
    double* A21 = A + h;
    double* A22 = A + h + (long)h * lda;
    // A21 = A21 * X11^T ; A22 -= A21 A21^T
    MOPS_TRY(copy_tile_d(s, m, h, A21, lda, T, m));
    MOPS_TRY(gemm_d(s, false, true, m, h, h, 1.0, T, m, X11, ldx, 0.0, A21, lda, GM_FULL,
                    TRI_B_UPPER));
    MOPS_TRY(gemm_d(s, false, true, m, m, h, -1.0, A21, lda, A21, lda, 1.0, A22, lda, GM_LOWER));
    MOPS_TRY(chol_diag(s, upper, m, A22, lda, X22, ldx, goff + h, info, badval, T));
    // X21 = -X22 * A21 * X11 ; X12 = 0
    MOPS_TRY(gemm_d(s, false, false, m, h, h, 1.0, A21, lda, X11, ldx, 0.0, T, m, GM_FULL,
                    TRI_B_LOWER));
    MOPS_TRY(gemm_d(s, false, false, m, h, m, -1.0, X22, ldx, T, m, 0.0, X + h, ldx, GM_FULL,
                    TRI_A_LOWER));
    MOPS_TRY(scale_tile_d(s, h, m, 0.0, X + (long)h * ldx, ldx));
  } else {
    double* A12 = A + (long)h * lda;
    double* A22 = A + h + (long)h * lda;
    // A12 = X11^T * A12 ; A22 -= A12^T A12
    MOPS_TRY(copy_tile_d(s, h, m, A12, lda, T, h));
    MOPS_TRY(gemm_d(s, true, false, h, m, h, 1.0, X11, ldx, T, h, 0.0, A12, lda, GM_FULL,
                    TRI_A_LOWER));
    MOPS_TRY(gemm_d(s, true, false, m, m, h, -1.0, A12, lda, A12, lda, 1.0, A22, lda, GM_UPPER));
    MOPS_TRY(chol_diag(s, upper, m, A22, lda, X22, ldx, goff + h, info, badval, T));
    // X12 = -X11 * A12 * X22 ; X21 = 0
    MOPS_TRY(gemm_d(s, false, false, h, m, m, 1.0, A12, lda, X22, ldx, 0.0, T, h, GM_FULL,
                    TRI_B_UPPER));
    MOPS_TRY(gemm_d(s, false, false, h, m, h, -1.0, X11, ldx, T, h, 0.0, X + (long)h * ldx, ldx,
                    GM_FULL, TRI_A_UPPER));
    MOPS_TRY(scale_tile_d(s, m, h, 0.0, X + h, ldx));
  }
  return ST_OK;
}

}  // namespace

size_t potrf_ws_bytes(int N, int nb) {
  if (N <= 0) return 0;
  if (nb <= 0 || nb > N) nb = N;
  const size_t x = (size_t)nb * nb * sizeof(double);
  const size_t t = (size_t)N * nb * sizeof(double);  // panel copy, also >= nb*nb
  return x + t + 1024;
}

namespace {

// Diagonal block k (order kb) and the panel below / right of it:
// factor + inverse, then A21 = A21 * inv(L11)^T (or A12 = inv(U11)^T * A12).
int chol_panel(hipStream_t s, bool upper, int N, int k, int kb, double* A, long lda, double* X,
               double* T, int* info, double* badval) {
  double* Akk = A + k + (long)k * lda;
  MOPS_TRY(chol_diag(s, upper, kb, Akk, lda, X, kb, k, info, badval, T));
  const int m = N - k - kb;
  if (m <= 0) return ST_OK;
  if (!upper) {
    double* A21 = A + (k + kb) + (long)k * lda;
    MOPS_TRY(copy_tile_d(s, m, kb, A21, lda, T, m));
    return gemm_d(s, false, true, m, kb, kb, 1.0, T, m, X, kb, 0.0, A21, lda, GM_FULL,
                  TRI_B_UPPER);
  }
  double* A12 = A + k + (long)(k + kb) * lda;
  MOPS_TRY(copy_tile_d(s, kb, m, A12, lda, T, kb));
  return gemm_d(s, true, false, kb, m, kb, 1.0, X, kb, T, kb, 0.0, A12, lda, GM_FULL,
                TRI_A_LOWER);
}

// Update of the trailing block that starts c0 columns after the panel's
// end (rows and columns j0 = k+kb+c0 ... j0+nc of the trailing matrix),
// with the panel's rows/columns r0 = k+kb+c0 ... : trap = only the next
// block column (a trapezoid of nc columns), else the square to the end.
int chol_update(hipStream_t s, bool upper, int N, int k, int kb, int c0, int nc, bool trap,
                double* A, long lda) {
  const int j0 = k + kb + c0;
  const int rows = N - j0;
  if (rows <= 0 || nc <= 0) return ST_OK;
  double* C = A + j0 + (long)j0 * lda;
  if (!upper) {
    const double* P = A + j0 + (long)k * lda;  // A21 rows j0..N
    if (trap)
      return gemm_d(s, false, true, rows, nc, kb, -1.0, P, lda, P, lda, 1.0, C, lda, GM_LOWER_T);
    return gemm_d(s, false, true, rows, rows, kb, -1.0, P, lda, P, lda, 1.0, C, lda, GM_LOWER);
  }
  const double* P = A + k + (long)j0 * lda;  // A12 columns j0..N
  if (trap)
    return gemm_d(s, true, false, nc, rows, kb, -1.0, P, lda, P, lda, 1.0, C, lda, GM_UPPER_T);
  return gemm_d(s, true, false, rows, rows, kb, -1.0, P, lda, P, lda, 1.0, C, lda, GM_UPPER);
}

}  // namespace

int potrf_d(hipStream_t s, int flags, int N, int nb, double* A, long lda, int* info,
            double* badval) {
  if (N <= 0) return ST_OK;
  // blockedCHOL takes the LOWER branch when the bit is set, otherwise UPPER
  // (chol.go:40, :77).
  const bool upper = !(flags & F_LOWER);
  // DecomposeCHOL: unblocked when N < nb or nb == 0 (chol.go:121). The
  // unblocked and blocked algorithms compute the same factor; here both are
  // the blocked one with the outer block clamped to N.
  if (nb <= 0 || nb > N) nb = N;
  double* X = (double*)ws_carve((size_t)nb * nb * sizeof(double));
  double* T = (double*)ws_carve((size_t)N * nb * sizeof(double));
  if (!X || !T) {
    set_error("potrf: workspace too small");
    return ST_NOMEM;
  }
  hipLaunchKernelGGL(k_init_info, dim3(1), dim3(1), 0, s, info, badval);
  MOPS_CHECK_HIP(hipGetLastError());
  const int nsteps = ceil_div(N, nb);
  if (nsteps <= 2) {
    for (int k = 0; k < N; k += nb) {
      const int kb = min(nb, N - k);
      MOPS_TRY(chol_panel(s, upper, N, k, kb, A, lda, X, T, info, badval));
      const int m = N - k - kb;
      if (m <= 0) break;
      trace_mark_begin(s, TR_SYRK_UPDATE);
      MOPS_TRY(chol_update(s, upper, N, k, kb, 0, m, false, A, lda));
      trace_mark_end(s, TR_SYRK_UPDATE, (double)m * (m + 1) * kb);
    }
  } else {
    // Look-ahead: the panel of step k+1 (and the update of its block column
    // by step k) runs on a second stream while the caller's stream applies
    // step k to the rest of the trailing matrix.
    hipStream_t s2 = aux_stream();
    hipEvent_t e_fork = sync_event(0), e_pan = sync_event(1), e_rest = sync_event(2),
               e_join = sync_event(3);
    if (!s2 || !e_fork || !e_pan || !e_rest || !e_join) {
      set_error("potrf: could not create the look-ahead stream");
      return ST_HIP;
    }
    MOPS_CHECK_HIP(hipEventRecord(e_fork, s));
    MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_fork, 0));
    MOPS_TRY(chol_panel(s2, upper, N, 0, min(nb, N), A, lda, X, T, info, badval));
    MOPS_CHECK_HIP(hipEventRecord(e_pan, s2));
    bool have_rest = false;
    for (int k = 0; k < N; k += nb) {
      const int kb = min(nb, N - k);
      const int m = N - k - kb;
      if (m <= 0) break;
      const int kb1 = min(nb, m);  // width of the next block column
      // s2 may touch block column k+1 only after the previous rest update
      if (have_rest) MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_rest, 0));
      // caller's stream: rest of the trailing matrix, after panel k
      MOPS_CHECK_HIP(hipStreamWaitEvent(s, e_pan, 0));
      if (m > kb1) {
        const int mr = m - kb1;
        trace_mark_begin(s, TR_SYRK_UPDATE);
        MOPS_TRY(chol_update(s, upper, N, k, kb, kb1, mr, false, A, lda));
        trace_mark_end(s, TR_SYRK_UPDATE, (double)mr * (mr + 1) * kb);
        MOPS_CHECK_HIP(hipEventRecord(e_rest, s));
        have_rest = true;
      }
      // second stream: next block column, then the next panel
      MOPS_TRY(chol_update(s2, upper, N, k, kb, 0, kb1, true, A, lda));
      MOPS_TRY(chol_panel(s2, upper, N, k + kb, kb1, A, lda, X, T, info, badval));
      MOPS_CHECK_HIP(hipEventRecord(e_pan, s2));
    }
    MOPS_CHECK_HIP(hipEventRecord(e_join, s2));
    MOPS_CHECK_HIP(hipStreamWaitEvent(s, e_join, 0));
  }
  hipLaunchKernelGGL(k_fix_info, dim3(1), dim3(1), 0, s, info);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

size_t potrs_ws_bytes(int flags, int N, int nrhs) {
  return 2 * trsm_ws_bytes(flags | F_LEFT, N, nrhs);
}

int potrs_d(hipStream_t s, int flags, int N, int nrhs, const double* A, long lda, double* B,
            long ldb) {
  // SolveCHOL (chol.go:147-158): UPPER solves U^T then U; LOWER solves L then
  // L^T; neither flag leaves B unchanged.
  if (flags & F_UPPER) {
    MOPS_TRY(trsm_d(s, F_UPPER | F_TRANSA, N, nrhs, 1.0, A, lda, B, ldb));
    MOPS_TRY(trsm_d(s, F_UPPER, N, nrhs, 1.0, A, lda, B, ldb));
  } else if (flags & F_LOWER) {
    MOPS_TRY(trsm_d(s, F_LOWER, N, nrhs, 1.0, A, lda, B, ldb));
    MOPS_TRY(trsm_d(s, F_LOWER | F_TRANSA, N, nrhs, 1.0, A, lda, B, ldb));
  }
  return ST_OK;
}

}  // namespace mops
