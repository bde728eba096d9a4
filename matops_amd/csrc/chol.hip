// Blocked Cholesky factorization and solve on the GPU.
//
// Replaces the panel loop of blockedCHOL / unblockedCHOL (chol.go:17-94),
// DecomposeCHOL (chol.go:116-127) and SolveCHOL (chol.go:147-158). The
// reference's step for a diagonal block of width nb is
//   A11 = chol(A11)  (per column: sqrt, a21 /= a11, SYR; chol.go:29-46)
//   A21 = A21 * tril(A11)^-T          (SolveTrm RIGHT|LOWER|TRANSA, :79)
//   A22 = A22 - A21 * A21^T           (RankUpdateSym LOWER, :81)
// and the UPPER variant with A12 / U11^-T (:82-87). The same step runs here
// with the panel (diagonal block and the rows below it) factored
// recursively: 64-column leaves factor their diagonal block in registers
// (the reference's column algorithm), which also yields the leaf's inverse,
// and solve the rows below with one GEMM; the halves are joined by a
// trapezoidal GEMM update. The trailing update of step k runs on the
// caller's stream while the next panel is factored on a second stream.
#include <cmath>

#include "mops_solvers.hpp"

namespace mops {
namespace {

constexpr int LEAF = 64;

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

constexpr unsigned LEAF_SPIN_LIMIT = 1u << 26;

// One 64-column leaf of the Cholesky panel in one launch (64-thread
// workgroups, one wave each).
//  * workgroup 0 factors the n x n diagonal block (n <= 64): lane i holds row
//    i in registers; column j takes the square root of the diagonal (a NaN
//    records its global index, chol.go:33-36), scales the column by the
//    reciprocal (InvScale, chol.go:42 -- built with -ffast-math, the
//    reference's loop of divisions is a reciprocal multiply) and applies the
//    rank-1 update to the rest of each row (the SYR of chol.go:44), the
//    column entries broadcast with readlane. All loops are unrolled so the
//    registers are addressed statically; the critical chain per column is a
//    square root, a reciprocal and one multiply-add. The factor is stored
//    and released to the other workgroups through a flag.
//  * workgroup b >= 1 solves 64 rows below the block (X L^T = B, the TRSM of
//    chol.go:79 for this leaf): lane = row, right-looking substitution with
//    the factor broadcast from LDS.
// The UPPER case runs the same on U^T (rows of U^T = columns of U).
__global__ __launch_bounds__(64) void k_chol_leaf(int n, int mb, int upper, double* A, long lda,
                                                  double* B, int goff, int* info,
                                                  double* badval, int* flag, int* err) {
  constexpr int NL = 64;
  const int l = threadIdx.x;
  if (blockIdx.x == 0) {
    double a[NL];
#pragma unroll
    for (int c = 0; c < NL; ++c) {
      double v = 0.0;
      if (l < n && c < n && c <= l) v = upper ? A[c + (long)l * lda] : A[l + (long)c * lda];
      a[c] = v;
    }
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      if (j < n) {
        const double d = readlane_d(a[j], j);
        const double r = sqrt(d);
        if (l == 0 && isnan(r) && info) {
          const int old = atomicMin(info, goff + j);
          if ((old < 0 || old > goff + j) && badval) *badval = d;
        }
        const double ri = 1.0 / r;
        a[j] = l > j ? a[j] * ri : (l == j ? r : a[j]);
#pragma unroll
        for (int k = j + 1; k < NL; ++k) {
          if (k < n) {
            const double lk = readlane_d(a[j], k);
            if (l >= k) a[k] -= a[j] * lk;
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NL; ++c)
      if (l < n && c < n && c <= l) {
        if (upper)
          A[c + (long)l * lda] = a[c];
        else
          A[l + (long)c * lda] = a[c];
      }
    if (l == 0) __hip_atomic_store(flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  // ---- rows below: row index of this lane
  __shared__ double Ls[NL][NL + 1];
  __shared__ double dinv[NL];
  const int row = (blockIdx.x - 1) * NL + l;
  double x[NL];
#pragma unroll
  for (int c = 0; c < NL; ++c) {
    double v = 0.0;
    if (row < mb && c < n) v = upper ? B[c + (long)row * lda] : B[row + (long)c * lda];
    x[c] = v;
  }
  if (l == 0) {
    unsigned spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
      if (++spins > LEAF_SPIN_LIMIT) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
  // the factor, row by row: Ls[i][k] = L[i][k] (k <= i)
  for (int idx = l; idx < NL * NL; idx += NL) {
    const int i = idx % NL, k = idx / NL;
    double v = 0.0;
    if (i < n && k <= i)
      v = upper ? __hip_atomic_load(&A[k + (long)i * lda], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT)
                : __hip_atomic_load(&A[i + (long)k * lda], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
    Ls[i][k] = v;
  }
  __syncthreads();
  dinv[l] = l < n ? 1.0 / Ls[l][l] : 0.0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    if (k < n) {
      x[k] *= dinv[k];
#pragma unroll
      for (int c = k + 1; c < NL; ++c)
        if (c < n) x[c] -= x[k] * Ls[c][k];
    }
  }
  if (row < mb) {
#pragma unroll
    for (int c = 0; c < NL; ++c)
      if (c < n) {
        if (upper)
          B[c + (long)row * lda] = x[c];
        else
          B[row + (long)c * lda] = x[c];
      }
  }
}

__global__ void k_zero_flags(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void k_init_info(int* info, double* badval) {
  if (info) *info = 0x7fffffff;
  if (badval) *badval = 0.0;
}

// info: -1 when the factorization succeeded, -2 when a leaf's wait for its
// diagonal block timed out, else the first NaN square root's index.
__global__ void k_fix_info(int* info, const int* err) {
  if (!info) return;
  if (*err) *info = -2;
  else if (*info == 0x7fffffff) *info = -1;
}

// Factor the panel of columns [j0, j0+w) (rows j0..N-1) in place: the
// recursive (left half, trapezoidal update of the right half, right half)
// scheme of a Cholesky panel, with 64-column leaves (k_chol_leaf: the
// diagonal block and the rows below it in one launch). The UPPER case is
// the transpose of each step.
struct LeafFlags {
  int* flags;  // one zeroed flag per leaf launch
  int next;
  int* err;
};

int chol_rec(hipStream_t s, bool upper, int N, int j0, int w, double* A, long lda, LeafFlags& fl,
             int* info, double* badval) {
  if (w <= LEAF) {
    double* D = A + j0 + (long)j0 * lda;
    const int m = N - j0 - w;
    double* B = upper ? A + j0 + (long)(j0 + w) * lda : A + (j0 + w) + (long)j0 * lda;
    int* flag = fl.flags + fl.next;
    fl.next += 1;
    hipLaunchKernelGGL(k_chol_leaf, dim3(1 + ceil_div(m > 0 ? m : 0, LEAF)), dim3(LEAF), 0, s, w,
                       m, upper ? 1 : 0, D, lda, B, j0, info, badval, flag, fl.err);
    MOPS_CHECK_HIP(hipGetLastError());
    return ST_OK;
  }
  int h = ((w / 2 + LEAF - 1) / LEAF) * LEAF;
  if (h >= w) h = w - LEAF;
  MOPS_TRY(chol_rec(s, upper, N, j0, h, A, lda, fl, info, badval));
  const int rows = N - j0 - h;
  double* C = A + (j0 + h) + (long)(j0 + h) * lda;
  if (!upper) {
    const double* P = A + (j0 + h) + (long)j0 * lda;
    MOPS_TRY(gemm_d(s, false, true, rows, w - h, h, -1.0, P, lda, P, lda, 1.0, C, lda,
                    GM_LOWER_T));
  } else {
    const double* P = A + j0 + (long)(j0 + h) * lda;
    MOPS_TRY(gemm_d(s, true, false, w - h, rows, h, -1.0, P, lda, P, lda, 1.0, C, lda,
                    GM_UPPER_T));
  }
  return chol_rec(s, upper, N, j0 + h, w - h, A, lda, fl, info, badval);
}

}  // namespace

size_t potrf_ws_bytes(int N, int nb) {
  if (nb <= 0 || nb > N) nb = N;
  return (size_t)(ceil_div(N, LEAF) + 2 * ceil_div(N, nb > 0 ? nb : 1) + 8) * sizeof(int) + 1024;
}

namespace {

// Panel k (diagonal block of order kb and everything below / right of it).
int chol_panel(hipStream_t s, bool upper, int N, int k, int kb, double* A, long lda,
               LeafFlags& fl, int* info, double* badval) {
  return chol_rec(s, upper, N, k, kb, A, lda, fl, info, badval);
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
  // one flag per 64-column leaf, plus an error word
  const int nleaf = ceil_div(N, LEAF) + 2 * ceil_div(N, nb) + 4;
  LeafFlags fl;
  fl.flags = (int*)ws_carve((size_t)(nleaf + 1) * sizeof(int));
  fl.next = 0;
  if (!fl.flags) {
    set_error("potrf: workspace too small");
    return ST_NOMEM;
  }
  fl.err = fl.flags + nleaf;
  hipLaunchKernelGGL(k_zero_flags, dim3(ceil_div(nleaf + 1, 256)), dim3(256), 0, s, fl.flags,
                     nleaf + 1);
  hipLaunchKernelGGL(k_init_info, dim3(1), dim3(1), 0, s, info, badval);
  MOPS_CHECK_HIP(hipGetLastError());
  const int nsteps = ceil_div(N, nb);
  if (nsteps <= 2) {
    for (int k = 0; k < N; k += nb) {
      const int kb = min(nb, N - k);
      MOPS_TRY(chol_panel(s, upper, N, k, kb, A, lda, fl, info, badval));
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
    MOPS_TRY(chol_panel(s2, upper, N, 0, min(nb, N), A, lda, fl, info, badval));
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
      MOPS_TRY(chol_panel(s2, upper, N, k + kb, kb1, A, lda, fl, info, badval));
      MOPS_CHECK_HIP(hipEventRecord(e_pan, s2));
    }
    MOPS_CHECK_HIP(hipEventRecord(e_join, s2));
    MOPS_CHECK_HIP(hipStreamWaitEvent(s, e_join, 0));
  }
  hipLaunchKernelGGL(k_fix_info, dim3(1), dim3(1), 0, s, info, fl.err);
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
