// Symmetric indefinite factorizations with pivoting, and their solves:
//
//   ldl_piv_d   A = P^T L D L^T P or P^T U D U^T P with 1x1 pivots
//               (DecomposeLDL, ldl.go:25-484): at each step the largest
//               |diagonal| of the part not yet factored
//   bk_d        A = L D L^T or U D U^T with Bunch-Kaufman 1x1 / 2x2 pivots
//               (DecomposeBK, ldlbk.go:107-456, 1114-1140; the algorithm
//               of LAPACK's dsytf2, with its ipiv format)
//   ldl_piv_solve_d, bk_solve_d   SolveLDL (ldl.go:486-511) and SolveBK
//               (ldlbks.go:15-157, 303-316; LAPACK's dsytrs order)
//
// Every pivot step is four launches that keep their state in device
// memory, so the host never waits for a pivot decision:
//   decide  one workgroup: the pivot search (the reference's IAMax, with
//           its pair logic and tie rules: calgo/dvec_util.c:164-195) and,
//           for Bunch-Kaufman, the 1x1 / 2x2 choice; writes the pivot
//           entries and the next step
//   swap    the symmetric interchange in the stored triangle, one thread per
//           swapped pair (pivot.go:145-228 for LDL, which also moves the
//           rows of the factored columns; ldlbk.go:48-105 and the 2x2 extra
//           swap for Bunch-Kaufman, which do not)
//   update  the rank-1 / rank-2 update of the stored triangle of the part
//           not yet factored, 32 x 32 tiles over the whole GPU
//   scale   the pivot column(s): l = a / d, or the rows times E^-1
// A step's size (1 or 2 columns) is known only on the device: the host
// queues n steps, those past the end return at once, and each update grid
// is sized for the largest part the step can leave.
//
// Blocking. The blocked LOWER LDL (ldl.go:90-295) makes the pivot decisions
// of the unblocked one on the same updated values when its workspace has
// nb + 1 columns (its diagonal column then stays apart from the panel
// columns), so the step algorithm serves every nb; only its zero-pivot
// check is limited to the columns the reference factors unblocked (the
// bounded panels divide by zero without an error). Bunch-Kaufman likewise
// (the blocked form, ldlbk.go:459-1111, makes dsytrf's decisions, which are
// dsytf2's). The blocked UPPER LDL pivots only inside each diagonal block
// (ldl.go:356-418), which changes the result: it is kept, with the steps
// on the block, the interchanges on the blocks beside it, one triangular
// solve and one triangle-only GEMM per block.
#include "mops_solvers.hpp"

namespace mops {

namespace {

// The state of the step sequence (device memory).
struct SymStep {
  int next;  // LOWER: first column of the next step; UPPER: order of the part left for it
  int k;     // this step: LOWER its first column; UPPER the order of the part left
  int r;     // pivot: LOWER relative to k; UPPER absolute, -1 none (Bunch-Kaufman)
  int np;    // 1 or 2 columns
  int skip;  // this step does nothing (past the end, or stopped)
  int err;   // 0, or 1 + the index of a zero pivot (LDL)
  double d;  // blocked LDL: the pivot of step k (its updated diagonal)
};

constexpr double BK_ALPHA = 0.6403882032022075;  // (1 + sqrt(17)) / 8, ldlbk.go:16
constexpr int SD_T = 256;                         // threads of the decide kernel

// The reference's IAMax over x[0], x[inc], ... x[(N-1) inc] by one
// workgroup (every thread gets the result): pair (2q, 2q+1) offers the
// larger of its two (the first on a tie; a NaN first hides the second) and
// the first offer strictly above the running maximum (from 0.0) wins; an
// odd last element likewise. That is the smallest index among the offers
// of the largest value when it is above zero, else 0.
__device__ int wg_iamax(const double* __restrict__ x, long inc, int N, double* sv, int* si) {
  if (N <= 1) return 0;
  const int t = threadIdx.x;
  double bv = 0.0;
  int bi = 0x7fffffff;
  const int npair = N / 2;
  for (int q = t; q < npair; q += SD_T) {
    double c0 = fabs(x[(long)(2 * q) * inc]);
    const double c1 = fabs(x[(long)(2 * q + 1) * inc]);
    int i = 2 * q;
    if (c1 > c0) {
      c0 = c1;
      i = 2 * q + 1;
    }
    if (c0 > bv) {
      bv = c0;
      bi = i;
    }
  }
  if ((N & 1) && t == SD_T - 1) {
    // the last offer; it wins only above every pair's (ties go to the lower
    // index in the reduction below)
    const double c0 = fabs(x[(long)(N - 1) * inc]);
    if (c0 > bv) {
      bv = c0;
      bi = N - 1;
    }
  }
  sv[t] = bv;
  si[t] = bi;
  __syncthreads();
  for (int h = SD_T / 2; h > 0; h >>= 1) {
    if (t < h) {
      const double v = sv[t + h];
      const int i = si[t + h];
      if (v > sv[t] || (v == sv[t] && i < si[t])) {
        sv[t] = v;
        si[t] = i;
      }
    }
    __syncthreads();
  }
  const int r = si[0];
  __syncthreads();  // sv / si free for the next search
  return r == 0x7fffffff ? 0 : r;
}

// A[i][j] of a column-major matrix
__device__ __forceinline__ double& at(double* A, long lda, int i, int j) {
  return A[i + (long)j * lda];
}
__device__ __forceinline__ double at(const double* A, long lda, int i, int j) {
  return A[i + (long)j * lda];
}

__device__ __forceinline__ void swap_el(double* a, double* b) {
  const double v = *a;
  *a = *b;
  *b = v;
}

__global__ void k_sym_init(SymStep* st, int next) {
  st->next = next;
  st->skip = 0;
  st->err = 0;
}

__global__ void k_fill_i64(int64_t* p, long n, int64_t v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = v;
}

// info = the 0-based index of the zero pivot, -1 when none
__global__ void k_sym_info(const SymStep* st, int* info) { *info = st->err - 1; }

// The pivot decision of one step on the n x n view A; ipiv: the view's
// pivot entries. base: added to a zero pivot's index; LOWER LDL checks for
// a zero pivot only in the columns from check_from on.
template <bool BK, bool UPPER>
__global__ __launch_bounds__(SD_T) void k_sym_decide(int n, const double* __restrict__ A, long lda,
                                                     SymStep* st, int64_t* ipiv, int base,
                                                     int check_from) {
  __shared__ double sv[SD_T];
  __shared__ int si[SD_T];
  const int t = threadIdx.x;
  const int k = st->next;
  const bool end = st->err != 0 || (UPPER ? k <= 0 : k >= n);
  __syncthreads();  // every thread has read the state before thread 0 writes it
  if (end) {
    if (t == 0) st->skip = 1;
    return;
  }
  int r = 0, np = 1, err = 0;
  if (!BK) {
    if (!UPPER) {
      // ldl.go:47-64: the largest |diagonal| of the trailing part
      r = wg_iamax(A + k + (long)k * lda, lda + 1, n - k, sv, si);
      if (t == 0) ipiv[k] = r > 0 ? r + k + 1 : 0;
      if (k >= check_from && at(A, lda, k + r, k + r) == 0.0) err = 1 + base + k;
    } else {
      // ldl.go:321-333: the largest |diagonal| of the leading part
      r = wg_iamax(A, lda + 1, k, sv, si);
      if (t == 0) ipiv[k - 1] = r < k - 1 ? r + 1 : 0;
      if (at(A, lda, r, r) == 0.0) err = base + k;  // 1 + index k - 1
    }
  } else if (!UPPER) {
    // findBKPivot, LOWER (ldlbk.go:109-159) on T = A[k:, k:]
    const int m = n - k;
    const double* T = A + k + (long)k * lda;
    if (m > 1) {
      const double amax = fabs(T[0]);
      const int rr = wg_iamax(T + 1, 1, m - 1, sv, si) + 1;
      const double rmax = fabs(T[rr]);
      if (!(amax >= BK_ALPHA * rmax)) {
        // the largest off-diagonal |element| of row / column rr
        const int q = wg_iamax(T + rr, lda, rr, sv, si);
        double qmax = fabs(at(T, lda, rr, q));
        if (rr < m - 1) {
          const double* col = T + (rr + 1) + (long)rr * lda;
          const double qmax2 = fabs(col[wg_iamax(col, 1, m - rr - 1, sv, si)]);
          if (qmax2 > qmax) qmax = qmax2;
        }
        if (!(amax >= BK_ALPHA * rmax * (rmax / qmax))) {
          r = rr;
          np = fabs(at(T, lda, rr, rr)) >= BK_ALPHA * qmax ? 1 : 2;
        }
      }
    }
    if (t == 0) {
      ipiv[k] = np == 1 ? r + k + 1 : -(r + k + 1);
      if (np == 2) ipiv[k + 1] = -(r + k + 1);
    }
  } else {
    // findBKPivot, UPPER (ldlbk.go:161-208) on A[0:e, 0:e]
    const int e = k, last = e - 1;
    r = e == 1 ? 0 : -1;
    if (e > 1) {
      const double amax = fabs(at(A, lda, last, last));
      const int rr = wg_iamax(A + (long)last * lda, 1, last, sv, si);
      const double rmax = fabs(at(A, lda, rr, last));
      if (!(amax >= BK_ALPHA * rmax)) {
        double qmax = 0.0;
        if (rr > 0) qmax = fabs(at(A, lda, wg_iamax(A + (long)rr * lda, 1, rr, sv, si), rr));
        const double* row = A + rr + (long)(rr + 1) * lda;
        const double qmax2 = fabs(row[(long)wg_iamax(row, lda, last - rr, sv, si) * lda]);
        if (qmax2 > qmax) qmax = qmax2;
        if (!(amax >= BK_ALPHA * rmax * (rmax / qmax))) {
          r = rr;
          np = fabs(at(A, lda, rr, rr)) >= BK_ALPHA * qmax ? 1 : 2;
        }
      }
    }
    if (t == 0) {
      if (np == 1) {
        ipiv[e - 1] = r == -1 ? e : r + 1;
      } else {
        ipiv[e - 2] = -(r + 1);
        ipiv[e - 1] = -(r + 1);
      }
    }
  }
  if (t == 0) {
    st->k = k;
    st->r = r;
    st->np = np;
    st->err = err;
    st->skip = 0;  // the interchange of a zero-pivot step is still made
    st->next = UPPER ? k - np : k + np;
  }
}

// The symmetric interchange of the step, one thread per swapped pair (the
// pairs of one interchange are disjoint).
template <bool BK, bool UPPER>
__global__ void k_sym_swap(int n, double* A, long lda, const SymStep* st) {
  if (st->skip) return;
  const int k = st->k, r = st->r, np = st->np;
  const long e0 = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  if (!BK && !UPPER) {
    // applyPivotSym LOWER (pivot.go:157-178), AL = A[k:, :k], AR = A[k:, k:]
    if (r == 0) return;
    const int p = k + r;  // the pivot row
    const long n1 = k, n2 = r - 1, n3 = n - p - 1;
    for (long e = e0; e < n1 + n2 + n3 + 1; e += stride) {
      if (e < n1) {  // rows k and p of the factored columns
        swap_el(&at(A, lda, k, (int)e), &at(A, lda, p, (int)e));
      } else if (e < n1 + n2) {  // column k between the two <-> row p
        const int i = k + 1 + (int)(e - n1);
        swap_el(&at(A, lda, i, k), &at(A, lda, p, i));
      } else if (e < n1 + n2 + n3) {  // columns k and p below row p
        const int i = p + 1 + (int)(e - n1 - n2);
        swap_el(&at(A, lda, i, k), &at(A, lda, i, p));
      } else {
        swap_el(&at(A, lda, k, k), &at(A, lda, p, p));
      }
    }
    return;
  }
  if (!BK && UPPER) {
    // applyPivotSym UPPER (pivot.go:179-200), AL = A[0:e, :], AR = A[e-1, e-1:]
    const int e = k, c = e - 1, idx = r;
    if (idx >= c) return;
    const long n1 = n - e, n2 = e - idx - 2, n3 = idx;
    for (long q = e0; q < n1 + n2 + n3 + 1; q += stride) {
      if (q < n1) {  // rows c and idx right of the part left
        const int j = e + (int)q;
        swap_el(&at(A, lda, c, j), &at(A, lda, idx, j));
      } else if (q < n1 + n2) {  // column c between the two <-> row idx
        const int i = idx + 1 + (int)(q - n1);
        swap_el(&at(A, lda, i, c), &at(A, lda, idx, i));
      } else if (q < n1 + n2 + n3) {  // columns idx and c above row idx
        const int i = (int)(q - n1 - n2);
        swap_el(&at(A, lda, i, idx), &at(A, lda, i, c));
      } else {
        swap_el(&at(A, lda, c, c), &at(A, lda, idx, idx));
      }
    }
    return;
  }
  if (BK && !UPPER) {
    // ldlbk.go:228-243: applyBKPivotSym(T = A[k:, k:], np - 1, r) unless r
    // is 0 or np - 1, and for a 2x2 pivot T[1][0] <-> T[r][0]
    if (r == 0 || r == np - 1) return;
    double* T = A + k + (long)k * lda;
    const int m = n - k, src = np - 1, dst = r;
    const long n2 = dst - src - 1, n3 = m - dst - 1, nx = np == 2 ? 1 : 0;
    for (long e = e0; e < n2 + n3 + 1 + nx; e += stride) {
      if (e < n2) {
        const int i = src + 1 + (int)e;
        swap_el(&at(T, lda, i, src), &at(T, lda, dst, i));
      } else if (e < n2 + n3) {
        const int i = dst + 1 + (int)(e - n2);
        swap_el(&at(T, lda, i, src), &at(T, lda, i, dst));
      } else if (e == n2 + n3) {
        swap_el(&at(T, lda, src, src), &at(T, lda, dst, dst));
      } else {
        swap_el(&at(T, lda, 1, 0), &at(T, lda, r, 0));
      }
    }
    return;
  }
  // BK UPPER (ldlbk.go:359-374): applyBKPivotSym(A[0:e, 0:e], e - np, r)
  // unless r is -1, and for a 2x2 pivot A[e-2][e-1] <-> A[r][e-1]
  if (r == -1) return;
  const int e = k, src = e - np, dst = r, nr = e - 1;
  if (src == dst && np == 1) return;
  const long n2 = src > dst ? src - dst - 1 : 0, n3 = src > dst ? dst : 0;
  const long nd = src > dst ? 1 : 0, nx = np == 2 && r != nr - 1 ? 1 : 0;
  for (long q = e0; q < n2 + n3 + nd + nx; q += stride) {
    if (q < n2) {
      const int i = dst + 1 + (int)q;
      swap_el(&at(A, lda, i, src), &at(A, lda, dst, i));
    } else if (q < n2 + n3) {
      const int i = (int)(q - n2);
      swap_el(&at(A, lda, i, src), &at(A, lda, i, dst));
    } else if (q < n2 + n3 + nd) {
      swap_el(&at(A, lda, src, src), &at(A, lda, dst, dst));
    } else {
      swap_el(&at(A, lda, nr - 1, nr), &at(A, lda, r, nr));
    }
  }
}

// The update of the part not yet factored, on its stored triangle: LOWER
// A22 = A[k+np:, k+np:], UPPER A00 = A[0:e-np, 0:e-np]; block (x, y) is the
// 32 x 32 tile of rows 32 y, columns 32 x of that part.
//   np = 1: A22[i][j] += x_i (alpha x_j), x the pivot column, alpha = -1/d
//           (MVUpdateTrm)
//   np = 2: A22[i][j] -= l_i0 c_j0 + l_i1 c_j1, c the two pivot columns and
//           l = c E^-1 their scaled form (ldlbk.go:287-298)
template <bool UPPER>
__global__ __launch_bounds__(256) void k_sym_update(int n, double* A, long lda,
                                                    const SymStep* st) {
  if (st->skip || st->err) return;
  const int k = st->k, np = st->np;
  const int m = UPPER ? k - np : n - k - np;  // order of the updated part
  const int bi = blockIdx.y * 32, bj = blockIdx.x * 32;
  if (bi >= m || bj >= m) return;
  if (!UPPER && bj > bi + 31) return;  // above the diagonal
  if (UPPER && bi > bj + 31) return;   // below the diagonal
  const int o = UPPER ? 0 : k + np;    // first row / column of the part
  const int c0 = UPPER ? k - np : k;   // first pivot column
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int i = bi + tx;
  if (i >= m) return;
  if (np == 1) {
    const double alpha = -1.0 / at(A, lda, c0, c0);
    const double xi = at(A, lda, o + i, c0);
    for (int q = ty; q < 32; q += 8) {
      const int j = bj + q;
      if (j >= m || (!UPPER && i < j) || (UPPER && i > j)) continue;
      at(A, lda, o + i, o + j) += xi * (alpha * at(A, lda, o + j, c0));
    }
    return;
  }
  // the 2 x 2 block E = [a b; b d] at rows / columns c0, c0 + 1; its
  // inverse is scale [d/b -1; -1 a/b], scale = 1 / ((a/b)(d/b) - 1) / b
  const double a = at(A, lda, c0, c0), d = at(A, lda, c0 + 1, c0 + 1);
  const double b = UPPER ? at(A, lda, c0, c0 + 1) : at(A, lda, c0 + 1, c0);
  const double i00 = d / b, i11 = a / b;
  const double scale = 1.0 / ((a / b) * (d / b) - 1.0) / b;
  const double ci0 = at(A, lda, o + i, c0), ci1 = at(A, lda, o + i, c0 + 1);
  const double l0 = scale * (ci0 * i00 + ci1 * -1.0);
  const double l1 = scale * (ci0 * -1.0 + ci1 * i11);
  for (int q = ty; q < 32; q += 8) {
    const int j = bj + q;
    if (j >= m || (!UPPER && i < j) || (UPPER && i > j)) continue;
    at(A, lda, o + i, o + j) -= l0 * at(A, lda, o + j, c0) + l1 * at(A, lda, o + j, c0 + 1);
  }
}

// The pivot column(s) of the step: a / d (InvScale divides), or each row
// of the two columns times E^-1.
template <bool UPPER>
__global__ void k_sym_scale(int n, double* A, long lda, const SymStep* st) {
  if (st->skip || st->err) return;
  const int k = st->k, np = st->np;
  const int m = UPPER ? k - np : n - k - np;
  const int o = UPPER ? 0 : k + np;
  const int c0 = UPPER ? k - np : k;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    if (np == 1) {
      at(A, lda, o + i, c0) /= at(A, lda, c0, c0);
    } else {
      const double a = at(A, lda, c0, c0), d = at(A, lda, c0 + 1, c0 + 1);
      const double b = UPPER ? at(A, lda, c0, c0 + 1) : at(A, lda, c0 + 1, c0);
      const double i00 = d / b, i11 = a / b;
      const double scale = 1.0 / ((a / b) * (d / b) - 1.0) / b;
      const double ci0 = at(A, lda, o + i, c0), ci1 = at(A, lda, o + i, c0 + 1);
      at(A, lda, o + i, c0) = scale * (ci0 * i00 + ci1 * -1.0);
      at(A, lda, o + i, c0 + 1) = scale * (ci0 * -1.0 + ci1 * i11);
    }
  }
}

inline int grid_for(long total) {
  const long b = (total + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

// The n steps of the unblocked algorithm on the n x n view A (state st
// already initialised).
template <bool BK, bool UPPER>
int sym_steps(hipStream_t s, int n, double* A, long lda, int64_t* ipiv, SymStep* st, int base,
              int check_from) {
  for (int i = 0; i < n; ++i) {
    hipLaunchKernelGGL((k_sym_decide<BK, UPPER>), dim3(1), dim3(SD_T), 0, s, n, A, lda, st, ipiv,
                       base, check_from);
    hipLaunchKernelGGL((k_sym_swap<BK, UPPER>), dim3(grid_for(n + 2)), dim3(256), 0, s, n, A, lda,
                       st);
    const int mmax = n - i - 1;  // the order of the part this step can leave
    if (mmax > 0) {
      const int tiles = ceil_div(mmax, 32);
      hipLaunchKernelGGL((k_sym_update<UPPER>), dim3(tiles, tiles), dim3(256), 0, s, n, A, lda,
                         st);
      hipLaunchKernelGGL((k_sym_scale<UPPER>), dim3(grid_for(mmax)), dim3(256), 0, s, n, A, lda,
                         st);
    }
    MOPS_CHECK_HIP(hipGetLastError());
  }
  return ST_OK;
}

// After a zero pivot in the columns from c0 on of the blocked LOWER LDL:
// the reference returns from its unblocked tail (ldl.go:280-285) before the
// tail's interchanges reach the rows of columns 0..c0-1 and before its
// pivots are made global. Undoes the first and the second of the step
// algorithm for the tail's steps up to the zero pivot. One workgroup.
__global__ __launch_bounds__(256) void k_ldl_tail_fixup(const SymStep* st, double* A, long lda,
                                                        int64_t* ipiv, int c0) {
  const int e = st->err - 1;  // the zero pivot's column
  if (e < c0) return;
  for (int c = threadIdx.x; c < c0; c += 256)
    for (int q = e; q >= c0; --q)
      if (ipiv[q] > 0) swap_el(&at(A, lda, q, c), &at(A, lda, (int)ipiv[q] - 1, c));
  __syncthreads();
  for (int q = c0 + threadIdx.x; q <= e; q += 256)
    if (ipiv[q] > 0) ipiv[q] -= c0;
}

// ---- blocked LOWER LDL: panels of PW columns with delayed updates ----
// (blkLowerLDL with its bounded panel, ldl.go:90-295, with the workspace of
// PW + 1 columns that keeps the updated diagonal apart.) Per step k of a
// panel starting at k0 (j = k - k0):
//   column  a = A[k+1:, k] - L[k+1:, k0:k] W[k, 0:j]^T (the updated column;
//           findAndBuildPivot's GEMV), W[:, j] = a, L[:, k] = a / d_k,
//           diag -= a a / d_k (MVUpdateDiag), d_k = diag[k] stored in A[k][k]
//   decide  IAMax of the updated diagonal from k + 1 (each workgroup offers
//           the winner of its rows, the last one to finish reduces them)
//   swap    the symmetric interchange of the next step in the left columns,
//           the (not yet updated) trailing part, diag and W's rows
// so one launch per step; after the panel, A22 -= L21 W21^T as one
// triangle-only GEMM. The pivots are those of the unblocked algorithm on
// the same values; the updated values differ from its rank-1 updates by
// rounding only.
constexpr int PW = 64;     // panel width
constexpr int LC_T = 256;  // rows per workgroup of the column kernels

__device__ __forceinline__ double ld_ag(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_ag(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The reference's IAMax offer of this thread's row: rows e = 0, 1, ... of
// the search pair up as (2q, 2q + 1) on neighbouring lanes; the pair offers
// its larger |value| (the first on a tie; a NaN first hides the second), a
// lone last row its own; an offer counts only when above zero (the running
// maximum starts at 0.0). in: this thread's row is in the search.
__device__ __forceinline__ void pair_offer(double v, bool in, int e, double& ov, int& oi) {
  const double c = in ? fabs(v) : 0.0;
  const double cp = __shfl_xor(c, 1, 64);
  const bool pin = __shfl_xor((int)in, 1, 64) != 0;
  ov = 0.0;
  oi = 0x7fffffff;
  if ((e & 1) == 0 && in) {
    double b = c;
    int bi = e;
    if (pin && cp > c) {
      b = cp;
      bi = e + 1;
    }
    if (b > 0.0) {
      ov = b;
      oi = bi;
    }
  }
}

// max |value| over the workgroup, ties to the lower index (every thread
// gets it); NaN offers never win
__device__ __forceinline__ void wg_best(double& v, int& i, double* sv, int* si) {
  const int t = threadIdx.x;
  sv[t] = v;
  si[t] = i;
  __syncthreads();
  for (int h = LC_T / 2; h > 0; h >>= 1) {
    if (t < h) {
      const double w = sv[t + h];
      const int wi = si[t + h];
      if (w > sv[t] || (w == sv[t] && wi < si[t])) {
        sv[t] = w;
        si[t] = wi;
      }
    }
    __syncthreads();
  }
  v = sv[0];
  i = si[0];
  __syncthreads();
}

// Shared tail of the panel kernels, run by the workgroup that finished last:
// reduce the G partial offers, record the pivot of step kd (ipiv, st->k,
// st->r, st->d, the zero check) and make the part of its interchange that
// the next column kernel does not: rows kd and p of the panel's L columns
// k0..kd-1 and of W's first jw columns, the two diagonal entries of A and
// of diag. The trailing column / row moves are made by the next column
// kernel, the rows of the columns left of the panel after the panel.
__device__ void ldl_decide(int n, double* A, long lda, double* diag, double* W, long ldw, int k0,
                           int jw, SymStep* st, int64_t* ipiv, const double* pv, const int* pi,
                           int G, int kd, int check_from, double* sv, int* si) {
  const int t = threadIdx.x;
  double v = 0.0;
  int vi = 0x7fffffff;
  for (int b = t; b < G; b += LC_T) {
    const double w = ld_ag(pv + b);
    const int wi = __hip_atomic_load(pi + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (w > v || (w == v && wi < vi)) {
      v = w;
      vi = wi;
    }
  }
  wg_best(v, vi, sv, si);
  const int r = vi == 0x7fffffff ? 0 : vi;
  const int p = kd + r;
  if (t == 0) {
    const double dp = ld_ag(diag + p);
    ipiv[kd] = r > 0 ? p + 1 : 0;
    st->k = kd;
    st->r = r;
    st->d = dp;
    if (kd >= check_from && dp == 0.0) st->err = 1 + kd;
  }
  if (r == 0) return;
  __syncthreads();  // diag[p] is read before it is swapped
  const int nl = kd - k0;  // panel columns left of kd
  for (int e = t; e < nl + jw + 2; e += LC_T) {
    double *x, *y;
    if (e < nl) {
      x = &at(A, lda, kd, k0 + e);
      y = &at(A, lda, p, k0 + e);
    } else if (e < nl + jw) {
      x = W + kd + (long)(e - nl) * ldw;
      y = W + p + (long)(e - nl) * ldw;
    } else if (e == nl + jw) {
      x = diag + kd;
      y = diag + p;
    } else {
      x = &at(A, lda, kd, kd);
      y = &at(A, lda, p, p);
    }
    const double a = ld_ag(x), c = ld_ag(y);
    *x = c;
    *y = a;
  }
}

// last-workgroup election: every workgroup's agent-scope stores are
// complete before it arrives
__device__ __forceinline__ bool arrive_last(unsigned* ctr, unsigned base, int* flag) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned prev = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
    *flag = prev == base + gridDim.x - 1;
  }
  __syncthreads();
  return *flag != 0;
}

// Panel start k0: diag = the trailing diagonal, then the decision of step k0.
__global__ __launch_bounds__(LC_T) void k_ldl_pinit(int n, double* A, long lda, int k0,
                                                    double* diag, double* W, long ldw,
                                                    SymStep* st, int64_t* ipiv, double* pv,
                                                    int* pi, unsigned* ctr, unsigned base,
                                                    int check_from) {
  __shared__ double sv[LC_T];
  __shared__ int si[LC_T];
  __shared__ int last;
  const int t = threadIdx.x, b = blockIdx.x;
  const bool halt = st->err != 0;
  const int e = b * LC_T + t, i = k0 + e;
  const bool in = !halt && i < n;
  double d = 0.0;
  if (in) {
    d = at(A, lda, i, i);
    st_ag(diag + i, d);
  }
  double ov;
  int oi;
  pair_offer(d, in, e, ov, oi);
  wg_best(ov, oi, sv, si);
  if (t == 0) {
    st_ag(pv + b, ov);
    __hip_atomic_store(pi + b, oi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (!arrive_last(ctr, base, &last) || halt) return;
  ldl_decide(n, A, lda, diag, W, ldw, k0, 0, st, ipiv, pv, pi, gridDim.x, k0, check_from, sv, si);
}

// Step k = k0 + j of the panel (pivot st->r, st->d decided before): the
// trailing moves of its interchange fused with the column (row i's new
// column value is read where the interchange puts it, the value it
// displaces is moved there), then (decide) the decision of step k + 1.
__global__ __launch_bounds__(LC_T) void k_ldl_pcol(int n, double* A, long lda, int k0, int j,
                                                   double* diag, double* W, long ldw,
                                                   SymStep* st, int64_t* ipiv, double* pv,
                                                   int* pi, unsigned* ctr, unsigned base,
                                                   int check_from, int decide) {
  __shared__ double sv[LC_T];
  __shared__ int si[LC_T];
  __shared__ int last;
  const int t = threadIdx.x, b = blockIdx.x;
  const int k = k0 + j;
  const bool halt = st->err != 0;
  const int e = b * LC_T + t, i = k + 1 + e;
  const bool in = !halt && i < n;
  double nd = 0.0;
  if (!halt) {
    const double dk = st->d;
    const int p = k + st->r;
    if (b == 0 && t == 0) at(A, lda, k, k) = dk;
    if (in) {
      double a = at(A, lda, i, k);
      if (p != k) {  // applyPivotSym LOWER, pivot.go:167-172
        if (i < p) {
          const double x = at(A, lda, p, i);
          at(A, lda, p, i) = a;
          a = x;
        } else if (i > p) {
          const double x = at(A, lda, i, p);
          at(A, lda, i, p) = a;
          a = x;
        }
      }
      const double* Li = A + i + (long)k0 * lda;
      const double* Wk = W + k;
#pragma unroll 8
      for (int q = 0; q < j; ++q) a -= Li[(long)q * lda] * Wk[(long)q * ldw];
      st_ag(W + i + (long)j * ldw, a);
      st_ag(&at(A, lda, i, k), a / dk);
      nd = diag[i] + a * a * (-1.0 / dk);
      st_ag(diag + i, nd);
    }
  }
  double ov;
  int oi;
  pair_offer(nd, in, e, ov, oi);
  wg_best(ov, oi, sv, si);
  if (t == 0) {
    st_ag(pv + b, ov);
    __hip_atomic_store(pi + b, oi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (!arrive_last(ctr, base, &last) || halt || !decide || k + 1 >= n) return;
  ldl_decide(n, A, lda, diag, W, ldw, k0, j + 1, st, ipiv, pv, pi, gridDim.x, k + 1, check_from,
             sv, si);
}

// The trailing moves of step st->k's interchange (for a step whose column
// kernel does not run: the one with a zero pivot).
__global__ void k_ldl_swap_trailing(int n, double* A, long lda, const SymStep* st) {
  const int kd = st->k, r = st->r, p = kd + r;
  if (r == 0) return;
  for (int i = kd + 1 + blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    if (i < p)
      swap_el(&at(A, lda, i, kd), &at(A, lda, p, i));
    else if (i > p)
      swap_el(&at(A, lda, i, kd), &at(A, lda, i, p));
  }
}

// The interchanges of steps k0 .. k0 + w - 1 (ipiv, 1-based, 0 = none, in
// order) on the rows of columns 0 .. ncol - 1 (the columns left of the
// panel; applyRowPivots FORWARD, ldl.go:288). Each workgroup composes the
// row moves in LDS (rows k0 .. k0 + w - 1 and the distinct pivot rows
// beyond), then moves the affected rows of its RP_C columns.
constexpr int RP_C = 32;  // columns per workgroup of k_ldl_panel_rows
__global__ __launch_bounds__(256) void k_ldl_panel_rows(int ncol, double* A, long lda, int k0,
                                                        int w, const int64_t* __restrict__ ipiv,
                                                        int n) {
  __shared__ double vals[2 * PW * RP_C];
  __shared__ int row[2 * PW];  // row of slot s
  __shared__ int src[2 * PW];  // the row whose value ends in row[s]
  __shared__ long piv[PW];     // the panel's pivots, loaded by all threads at once
  __shared__ int ns;
  for (int j = threadIdx.x; j < w; j += blockDim.x) piv[j] = ipiv[k0 + j];
  __syncthreads();
  if (threadIdx.x < 64) {
    // one wave composes the moves in order; a pivot row beyond the panel is
    // looked up among the slots added so far (at most w of them, one per
    // lane) with a ballot
    const int l = threadIdx.x;
    for (int q = l; q < w; q += 64) {
      row[q] = k0 + q;
      src[q] = k0 + q;
    }
    int m = w;
    for (int j = 0; j < w; ++j) {
      const long pk = piv[j];
      if (pk <= 0 || pk > n) continue;
      const int pr = (int)pk - 1;
      int sp;
      if (pr < k0 + w) {
        sp = pr - k0;
      } else {
        const unsigned long long hit = __ballot(w + l < m && row[w + l] == pr);
        if (hit) {
          sp = w + __ffsll((long long)hit) - 1;
        } else {
          sp = m++;
          if (l == 0) {
            row[sp] = pr;
            src[sp] = pr;
          }
        }
      }
      if (l == 0) {
        const int x = src[j];
        src[j] = src[sp];
        src[sp] = x;
      }
    }
    if (l == 0) ns = m;
  }
  __syncthreads();
  // all sources of this workgroup's RP_C columns are read into LDS before
  // any destination is written (the moves permute the slot rows)
  const int m = ns, c0 = blockIdx.x * RP_C;
  for (int e = threadIdx.x; e < m * RP_C; e += 256) {
    const int q = e / RP_C, c = c0 + e % RP_C;
    if (c < ncol) vals[e] = A[src[q] + (long)c * lda];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < m * RP_C; e += 256) {
    const int q = e / RP_C, c = c0 + e % RP_C;
    if (c < ncol) A[row[q] + (long)c * lda] = vals[e];
  }
}

__global__ void k_zero_u32(unsigned* p) { *p = 0u; }

// The interchanges of a block's pivots p[0..kb) (1-based within the block,
// 0 = none), last to first (BACKWARD, pivot.go:63-97), on the columns
// (on_cols) or the rows of the rows x cols matrix X; one thread per row
// (on_cols) or column of X.
__global__ void k_block_pivots(int rows, int cols, double* X, long ldx, const int64_t* p, int kb,
                               int on_cols) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (on_cols) {
    if (t >= rows) return;
    for (int q = kb - 1; q >= 0; --q)
      if (p[q] > 0) swap_el(&X[t + (long)q * ldx], &X[t + (long)(p[q] - 1) * ldx]);
  } else {
    if (t >= cols) return;
    double* x = X + (long)t * ldx;
    for (int q = kb - 1; q >= 0; --q)
      if (p[q] > 0) swap_el(&x[q], &x[p[q] - 1]);
  }
}

__global__ void k_shift_pivots(int64_t* p, int kb, int off) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < kb && p[t] > 0) p[t] += off;
}

// X (rows x kb) columns times 1 / d_j (SolveDiag RIGHT), and W = X D
// (MultDiag RIGHT), d_j = A[j][j]
__global__ void k_diag_right(int rows, int kb, double* X, long ldx, const double* A, long lda,
                             double* W, long ldw) {
  const long total = (long)rows * kb;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % rows), j = (int)(e / rows);
    const double d = A[(long)j * (lda + 1)];
    const double x = X[i + (long)j * ldx] * (1.0 / d);
    X[i + (long)j * ldx] = x;
    W[i + (long)j * ldw] = x * d;
  }
}

// perm = the row order after the interchanges rows k <-> p[k] - 1 for the
// entries p[k] > 0 of p[0..npv), in order (forward) or last to first:
// B'[i] = B[perm[i]]. Entries beyond n are ignored. One workgroup: the
// pivots are staged in LDS by all threads, then one thread runs the
// sequence of interchanges on an index array in LDS (in global memory
// beyond PERM_LDS rows).
constexpr int PERM_LDS = 16384;
__global__ __launch_bounds__(256) void k_pivots_to_perm(int n, int npv,
                                                        const int64_t* __restrict__ p,
                                                        int forward, int* perm) {
  __shared__ int q[PERM_LDS];
  __shared__ int pv[PERM_LDS];
  const bool lds = n <= PERM_LDS;
  int* v = lds ? q : perm;
  for (int i = threadIdx.x; i < n; i += 256) v[i] = i;
  if (lds)
    for (int i = threadIdx.x; i < npv; i += 256) {
      const long pk = p[i];
      pv[i] = pk > 0 && pk <= n ? (int)pk - 1 : -1;
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int s = 0; s < npv; ++s) {
      const int k = forward ? s : npv - 1 - s;
      int o;
      if (lds) {
        o = pv[k];
      } else {
        const long pk = p[k];
        o = pk > 0 && pk <= n ? (int)pk - 1 : -1;
      }
      if (o >= 0) {
        const int t = v[k];
        v[k] = v[o];
        v[o] = t;
      }
    }
  }
  __syncthreads();
  if (lds)
    for (int i = threadIdx.x; i < n; i += 256) perm[i] = q[i];
}

// gather: Y[i][j] = X[perm[i]][j]; else scatter: Y[perm[i]][j] = X[i][j]
__global__ void k_permute_rows(int n, int nrhs, const int* __restrict__ perm, const double* X,
                               long ldx, double* Y, long ldy, int gather) {
  const long total = (long)n * nrhs;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (gather)
      Y[i + (long)j * ldy] = X[perm[i] + (long)j * ldx];
    else
      Y[perm[i] + (long)j * ldy] = X[i + (long)j * ldx];
  }
}

// SolveDiag(B, A, LEFT) (diag.go:85-139): rows times 1 / d_k; a 1 x 1 A
// divides
__global__ void k_solve_diag_left(int n, int nrhs, double* B, long ldb, const double* A,
                                  long lda) {
  const long total = (long)n * nrhs;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (n == 1)
      B[(long)j * ldb] /= A[0];
    else
      B[i + (long)j * ldb] *= 1.0 / A[(long)i * (lda + 1)];
  }
}

// Sum over the 256 threads; every thread gets it
__device__ double wg_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// The 2 x 2 block solve of SolveBK (ldlbks.go:62-80): [b0 b1] = E^-1 [b0 b1]
__device__ __forceinline__ void solve_e(double a, double b, double d, double* b0, double* b1) {
  const double apb = a / b, dpb = d / b;
  const double scale = (apb * dpb - 1.0) * b;
  const double s0 = *b0, s1 = *b1;
  *b0 = (dpb * s0 - s1) / scale;
  *b1 = (apb * s1 - s0) / scale;
}

// SolveBK on one right-hand side column per workgroup, in LAPACK dsytrs
// order. LOWER (ldlbks.go:15-157): blocks first to last, the interchange,
// the rows below -= the block's columns times b1, b1 = D^-1 b1; then last
// to first, b1 -= the block's columns^T times the rows below, the
// interchange. UPPER: the mirror image (blocks last to first, then first to
// last, the rows above).
// b[i] <-> b[j]; a pivot entry out of range is ignored
__device__ __forceinline__ void swap_rows(double* b, int i, long j, int n) {
  if (j >= 0 && j < n) swap_el(&b[i], &b[j]);
}

template <bool UPPER>
__global__ __launch_bounds__(256) void k_bk_solve(int n, const double* __restrict__ A, long lda,
                                                  const int64_t* __restrict__ p, double* B,
                                                  long ldb) {
  __shared__ double red[4];
  double* b = B + (long)blockIdx.x * ldb;
  const int t = threadIdx.x;
  if (!UPPER) {
    for (int k = 0; k < n;) {
      const int np = p[k] < 0 && k + 1 < n ? 2 : 1;
      if (t == 0) {
        if (np == 1) {
          if (p[k] != k + 1) swap_rows(b, k, p[k] - 1, n);
        } else if (-p[k] != k + 2) {
          swap_rows(b, k + 1, -p[k] - 1, n);
        }
      }
      __syncthreads();
      const double b0 = b[k], b1 = np == 2 ? b[k + 1] : 0.0;
      for (int i = k + np + t; i < n; i += 256) {
        double v = b[i] - at(A, lda, i, k) * b0;
        if (np == 2) v -= at(A, lda, i, k + 1) * b1;
        b[i] = v;
      }
      if (t == 0) {
        if (np == 1) {
          b[k] = b0 / at(A, lda, k, k);
        } else {
          double x0 = b0, x1 = b1;
          solve_e(at(A, lda, k, k), at(A, lda, k + 1, k), at(A, lda, k + 1, k + 1), &x0, &x1);
          b[k] = x0;
          b[k + 1] = x1;
        }
      }
      __syncthreads();
      k += np;
    }
    for (int nc = n; nc > 0;) {
      const int np = p[nc - 1] < 0 && nc > 1 ? 2 : 1, k = nc - np;
      double s0 = 0.0, s1 = 0.0;
      for (int i = nc + t; i < n; i += 256) {
        s0 += at(A, lda, i, k) * b[i];
        if (np == 2) s1 += at(A, lda, i, k + 1) * b[i];
      }
      s0 = wg_sum(s0, red);
      if (np == 2) s1 = wg_sum(s1, red);
      if (t == 0) {
        b[k] -= s0;
        if (np == 2) b[k + 1] -= s1;
        if (np == 1) {
          if (p[k] != nc) swap_rows(b, k, p[k] - 1, n);
        } else if (-p[k] != nc) {
          swap_rows(b, k + 1, -p[k] - 1, n);
        }
      }
      __syncthreads();
      nc -= np;
    }
    return;
  }
  for (int kk = n; kk > 0;) {  // kk: 1-based index of the block's last row
    const int np = p[kk - 1] < 0 && kk > 1 ? 2 : 1, k = kk - np;
    if (t == 0) {
      if (np == 1) {
        if (p[k] != kk) swap_rows(b, k, p[k] - 1, n);
      } else if (-p[k] != kk - 1) {
        swap_rows(b, k, -p[k] - 1, n);
      }
    }
    __syncthreads();
    const double b0 = b[k], b1 = np == 2 ? b[k + 1] : 0.0;
    for (int i = t; i < k; i += 256) {
      double v = b[i] - at(A, lda, i, k) * b0;
      if (np == 2) v -= at(A, lda, i, k + 1) * b1;
      b[i] = v;
    }
    if (t == 0) {
      if (np == 1) {
        b[k] = b0 / at(A, lda, k, k);
      } else {
        double x0 = b0, x1 = b1;
        solve_e(at(A, lda, k, k), at(A, lda, k, k + 1), at(A, lda, k + 1, k + 1), &x0, &x1);
        b[k] = x0;
        b[k + 1] = x1;
      }
    }
    __syncthreads();
    kk -= np;
  }
  for (int k = 0; k < n;) {
    const int np = p[k] < 0 && k + 1 < n ? 2 : 1;
    double s0 = 0.0, s1 = 0.0;
    for (int i = t; i < k; i += 256) {
      s0 += at(A, lda, i, k) * b[i];
      if (np == 2) s1 += at(A, lda, i, k + 1) * b[i];
    }
    s0 = wg_sum(s0, red);
    if (np == 2) s1 = wg_sum(s1, red);
    if (t == 0) {
      b[k] -= s0;
      if (np == 2) b[k + 1] -= s1;
      if (np == 1) {
        if (p[k] != k + 1) swap_rows(b, k, p[k] - 1, n);
      } else if (-p[k] != k + 1) {
        swap_rows(b, k, -p[k] - 1, n);
      }
    }
    __syncthreads();
    k += np;
  }
}

// SolveBK with the right-hand side and the pivots in LDS, for n <= BKS_MAX:
// one workgroup of BKS_T threads per right-hand side column, thread t owns
// rows t + BKS_T m (m < BKS_R); the L column(s) of the next block are
// loaded into registers while the current block is applied, so the
// sequential sweep waits on LDS and barriers instead of global memory.
// Same order of operations per block as k_bk_solve.
constexpr int BKS_T = 1024, BKS_MAX = 8 * BKS_T;

template <int R>
__device__ __forceinline__ void bks_load(const double* __restrict__ A, long lda, int col, bool two,
                                         int lo, int hi, double (&c0)[R], double (&c1)[R]) {
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int i = threadIdx.x + BKS_T * m;
    const bool in = i >= lo && i < hi;
    c0[m] = in ? A[i + (long)col * lda] : 0.0;
    c1[m] = in && two ? A[i + (long)(col + 1) * lda] : 0.0;
  }
}

// the sum over the workgroup of v, in every thread (red: 16 doubles)
__device__ __forceinline__ double bks_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int w = 0; w < BKS_T / 64; ++w) r += red[w];
  return r;
}

template <bool UPPER, int R>
__global__ __launch_bounds__(BKS_T) void k_bk_solve_lds(int n, const double* __restrict__ A,
                                                        long lda, const int64_t* __restrict__ p,
                                                        double* B, long ldb) {
  __shared__ double lb[BKS_MAX];
  __shared__ int lp[BKS_MAX];
  __shared__ double red[2][BKS_T / 64];
  double* b = B + (long)blockIdx.x * ldb;
  const int t = threadIdx.x;
  for (int i = t; i < n; i += BKS_T) {
    lb[i] = b[i];
    const long v = p[i];
    lp[i] = v >= -n && v <= n ? (int)v : 0;  // 0: no interchange, a 1x1 block
  }
  __syncthreads();
  // the 2 x 2 block solve of rows k, k + 1 (ldlbks.go:62-80), by thread 0
  auto dsolve = [&](int k, int np) {
    if (np == 1) {
      lb[k] = lb[k] / at(A, lda, k, k);
    } else {
      double x0 = lb[k], x1 = lb[k + 1];
      solve_e(at(A, lda, k, k), UPPER ? at(A, lda, k, k + 1) : at(A, lda, k + 1, k),
              at(A, lda, k + 1, k + 1), &x0, &x1);
      lb[k] = x0;
      lb[k + 1] = x1;
    }
  };
  auto swp = [&](int i, int j) {
    if (j >= 0 && j < n) {
      const double v = lb[i];
      lb[i] = lb[j];
      lb[j] = v;
    }
  };
  double c0[R], c1[R], x0[R], x1[R];
  // ---- phase 1: interchange, rows beyond the block -= its columns times
  // b1, b1 = E^-1 b1; blocks first to last (LOWER) or last to first
  {
    int k, np;
    if (!UPPER) {
      k = 0;
      np = lp[0] < 0 && n > 1 ? 2 : 1;
    } else {
      np = lp[n - 1] < 0 && n > 1 ? 2 : 1;
      k = n - np;
    }
    bks_load(A, lda, k, np == 2, UPPER ? 0 : k + np, UPPER ? k : n, c0, c1);
    int pk = -1, pnp = 0;  // the block whose E^-1 is pending
    while (true) {
      int nk, nnp;  // the next block
      if (!UPPER) {
        nk = k + np;
        nnp = nk < n ? (lp[nk] < 0 && nk + 1 < n ? 2 : 1) : 0;
      } else {
        nnp = k > 0 ? (lp[k - 1] < 0 && k > 1 ? 2 : 1) : 0;
        nk = k - nnp;
      }
      if (t == 0) {
        if (pnp) dsolve(pk, pnp);
        const int v = lp[k];
        if (!UPPER) {
          if (np == 1) {
            if (v != k + 1) swp(k, v - 1);
          } else if (-v != k + 2) {
            swp(k + 1, -v - 1);
          }
        } else {
          const int kk = k + np;  // 1-based last row of the block
          if (np == 1) {
            if (v != kk) swp(k, v - 1);
          } else if (-v != kk - 1) {
            swp(k, -v - 1);
          }
        }
      }
      __syncthreads();
      const double b0 = lb[k], b1 = np == 2 ? lb[k + 1] : 0.0;
      if (nnp) bks_load(A, lda, nk, nnp == 2, UPPER ? 0 : nk + nnp, UPPER ? nk : n, x0, x1);
      const int lo = UPPER ? 0 : k + np, hi = UPPER ? k : n;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int i = t + BKS_T * m;
        if (i >= lo && i < hi) {
          double v = lb[i] - c0[m] * b0;
          if (np == 2) v -= c1[m] * b1;
          lb[i] = v;
        }
      }
      __syncthreads();
      pk = k;
      pnp = np;
      if (!nnp) break;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        c0[m] = x0[m];
        c1[m] = x1[m];
      }
      k = nk;
      np = nnp;
    }
    if (t == 0 && pnp) dsolve(pk, pnp);
    __syncthreads();
  }
  // ---- phase 2: b1 -= the block's columns^T times the rows beyond it,
  // then the interchange; blocks last to first (LOWER) or first to last
  {
    int k, np;
    if (!UPPER) {
      np = lp[n - 1] < 0 && n > 1 ? 2 : 1;
      k = n - np;
    } else {
      k = 0;
      np = lp[0] < 0 && n > 1 ? 2 : 1;
    }
    bks_load(A, lda, k, np == 2, UPPER ? 0 : k + np, UPPER ? k : n, c0, c1);
    int par = 0;
    while (true) {
      int nk, nnp;
      if (!UPPER) {
        nnp = k > 0 ? (lp[k - 1] < 0 && k > 1 ? 2 : 1) : 0;
        nk = k - nnp;
      } else {
        nk = k + np;
        nnp = nk < n ? (lp[nk] < 0 && nk + 1 < n ? 2 : 1) : 0;
      }
      if (nnp) bks_load(A, lda, nk, nnp == 2, UPPER ? 0 : nk + nnp, UPPER ? nk : n, x0, x1);
      const int lo = UPPER ? 0 : k + np, hi = UPPER ? k : n;
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int i = t + BKS_T * m;
        if (i >= lo && i < hi) {
          const double v = lb[i];
          s0 += c0[m] * v;
          s1 += c1[m] * v;
        }
      }
      s0 = bks_sum(s0, red[par]);
      if (np == 2) s1 = bks_sum(s1, red[par ^ 1]);
      if (t == 0) {
        lb[k] -= s0;
        if (np == 2) lb[k + 1] -= s1;
        const int v = lp[k];
        if (!UPPER) {
          const int nc = k + np;
          if (np == 1) {
            if (v != nc) swp(k, v - 1);
          } else if (-v != nc) {
            swp(k + 1, -v - 1);
          }
        } else {
          if (np == 1) {
            if (v != k + 1) swp(k, v - 1);
          } else if (-v != k + 1) {
            swp(k, -v - 1);
          }
        }
      }
      __syncthreads();
      par ^= 1;
      if (!nnp) break;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        c0[m] = x0[m];
        c1[m] = x1[m];
      }
      k = nk;
      np = nnp;
    }
  }
  for (int i = t; i < n; i += BKS_T) b[i] = lb[i];
}

}  // namespace

namespace {

// The blocked LOWER LDL on the n x n matrix A (state st initialised, ipiv
// cleared): panels of PW columns; zero pivots are checked from check_from
// on. After a panel with checks the host reads the state: on a zero pivot
// at step e the pending updates of the panel's earlier steps are applied
// to A[e:, e:] (the unblocked algorithm's state at its error) and the
// panels stop.
int ldl_lower_panels(hipStream_t s, int n, double* A, long lda, int64_t* ipiv, SymStep* st,
                     int check_from) {
  const int G = ceil_div(n, LC_T);
  double* diag = (double*)ws_carve((size_t)n * sizeof(double));
  double* W = (double*)ws_carve((size_t)n * PW * sizeof(double));
  double* pv = (double*)ws_carve((size_t)G * sizeof(double));
  int* pi = (int*)ws_carve((size_t)G * sizeof(int));
  unsigned* ctr = (unsigned*)ws_carve(sizeof(unsigned));
  if (!diag || !W || !pv || !pi || !ctr) {
    set_error("ldl: workspace too small");
    return ST_NOMEM;
  }
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, ctr);
  unsigned base = 0;
  for (int k0 = 0; k0 < n; k0 += PW) {
    const int w = min(PW, n - k0);
    int g = ceil_div(n - k0, LC_T);
    hipLaunchKernelGGL(k_ldl_pinit, dim3(g), dim3(LC_T), 0, s, n, A, lda, k0, diag, W, (long)n,
                       st, ipiv, pv, pi, ctr, base, check_from);
    base += g;
    for (int j = 0; j < w; ++j) {
      const int rows = n - (k0 + j) - 1;
      g = rows > 0 ? ceil_div(rows, LC_T) : 1;
      hipLaunchKernelGGL(k_ldl_pcol, dim3(g), dim3(LC_T), 0, s, n, A, lda, k0, j, diag, W,
                         (long)n, st, ipiv, pv, pi, ctr, base, check_from, j + 1 < w ? 1 : 0);
      base += g;
    }
    MOPS_CHECK_HIP(hipGetLastError());
    if (k0 + w - 1 >= check_from) {
      int err = 0;
      MOPS_CHECK_HIP(hipMemcpyAsync(&err, &st->err, sizeof(int), hipMemcpyDeviceToHost, s));
      MOPS_CHECK_HIP(hipStreamSynchronize(s));
      if (err) {
        // complete the interchange of the zero-pivot step e, move the rows
        // of the columns left of the panel for steps k0 .. e, and apply the
        // pending updates of steps k0 .. e - 1 to A[e:, e:]
        const int e = err - 1, kp = e - k0;
        hipLaunchKernelGGL(k_ldl_swap_trailing, dim3(grid_for(n - e)), dim3(256), 0, s, n, A, lda,
                           st);
        if (k0 > 0)
          hipLaunchKernelGGL(k_ldl_panel_rows, dim3(ceil_div(k0, RP_C)), dim3(256), 0, s, k0, A,
                             lda, k0, kp + 1, ipiv, n);
        MOPS_CHECK_HIP(hipGetLastError());
        if (kp > 0 && e < n)
          MOPS_TRY(gemm_d(s, false, true, n - e, n - e, kp, -1.0, A + e + (long)k0 * lda, lda,
                          W + e, (long)n, 1.0, A + e + (long)e * lda, lda, GM_LOWER));
        return ST_OK;
      }
    }
    if (k0 > 0) {
      hipLaunchKernelGGL(k_ldl_panel_rows, dim3(ceil_div(k0, RP_C)), dim3(256), 0, s, k0, A, lda,
                         k0, w, ipiv, n);
      MOPS_CHECK_HIP(hipGetLastError());
    }
    if (k0 + w < n) {
      const int m2 = n - k0 - w;
      MOPS_TRY(gemm_d(s, false, true, m2, m2, w, -1.0, A + (k0 + w) + (long)k0 * lda, lda,
                      W + (k0 + w), (long)n, 1.0, A + (k0 + w) + (long)(k0 + w) * lda, lda,
                      GM_LOWER));
    }
  }
  return ST_OK;
}

}  // namespace

size_t ldl_piv_ws_bytes(int n, int nb, bool upper) {
  if (n <= 0) return 0;
  size_t b = 1024;
  if (!upper)
    b += (size_t)n * (PW + 1) * sizeof(double) + (size_t)ceil_div(n, LC_T) * 12 + 2048 +
         gemm_split_ws_bytes(n, n, PW, sizeof(double));
  if (upper && nb > 0 && n >= nb)
    b += (size_t)n * nb * sizeof(double) + trsm_ws_bytes(F_RIGHT, nb, n) +
         gemm_split_ws_bytes(n, n, nb, sizeof(double)) + 1024;
  return b;
}

int ldl_piv_d(hipStream_t s, bool upper, int n, int nb, double* A, long lda, int64_t* ipiv,
              long npiv, int* info, int* err_host) {
  if (err_host) *err_host = 0;
  if (npiv > 0) {
    hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(npiv)), dim3(256), 0, s, ipiv, npiv, (int64_t)0);
    MOPS_CHECK_HIP(hipGetLastError());
  }
  SymStep* st = (SymStep*)ws_carve(sizeof(SymStep));
  if (!st) {
    set_error("ldl: workspace too small");
    return ST_NOMEM;
  }
  const bool blocked = !(n < nb || nb == 0);
  if (n <= 0) {
    hipLaunchKernelGGL(k_sym_init, dim3(1), dim3(1), 0, s, st, 0);
  } else if (!upper) {
    // the columns the blocked form factors unblocked: from the first
    // multiple of nb that leaves at most nb columns
    int check_from = 0;
    if (blocked) {
      int k = 0;
      while (n - k > nb) k += nb;
      check_from = k;
    }
    hipLaunchKernelGGL(k_sym_init, dim3(1), dim3(1), 0, s, st, 0);
    MOPS_TRY(ldl_lower_panels(s, n, A, lda, ipiv, st, check_from));
    if (check_from > 0) {
      hipLaunchKernelGGL(k_ldl_tail_fixup, dim3(1), dim3(256), 0, s, st, A, lda, ipiv, check_from);
      MOPS_CHECK_HIP(hipGetLastError());
    }
  } else if (!blocked) {
    hipLaunchKernelGGL(k_sym_init, dim3(1), dim3(1), 0, s, st, n);
    MOPS_TRY((sym_steps<false, true>(s, n, A, lda, ipiv, st, 0, 0)));
  } else {
    // blkUpperLDL (ldl.go:356-418): pivots inside each diagonal block; the
    // host reads the state after each block to stop where the reference
    // returns on a zero pivot
    double* W = (double*)ws_carve((size_t)n * nb * sizeof(double));
    if (!W) {
      set_error("ldl: workspace too small");
      return ST_NOMEM;
    }
    for (int e = n; e > 0;) {
      const int kb = e < nb ? e : nb, k = e - kb;
      double* A11 = A + k + (long)k * lda;
      hipLaunchKernelGGL(k_sym_init, dim3(1), dim3(1), 0, s, st, kb);
      MOPS_TRY((sym_steps<false, true>(s, kb, A11, lda, ipiv + k, st, k, 0)));
      SymStep h;
      MOPS_CHECK_HIP(hipMemcpyAsync(&h, st, sizeof(SymStep), hipMemcpyDeviceToHost, s));
      MOPS_CHECK_HIP(hipStreamSynchronize(s));
      if (h.err) break;
      if (k > 0)
        hipLaunchKernelGGL(k_block_pivots, dim3(ceil_div(k, 256)), dim3(256), 0, s, k, kb,
                           A + (long)k * lda, lda, ipiv + k, kb, 1);
      if (n - e > 0)
        hipLaunchKernelGGL(k_block_pivots, dim3(ceil_div(n - e, 256)), dim3(256), 0, s, kb, n - e,
                           A + k + (long)e * lda, lda, ipiv + k, kb, 0);
      hipLaunchKernelGGL(k_shift_pivots, dim3(ceil_div(kb, 256)), dim3(256), 0, s, ipiv + k, kb, k);
      MOPS_CHECK_HIP(hipGetLastError());
      if (k > 0) {
        double* A01 = A + (long)k * lda;
        MOPS_TRY(trsm_d(s, F_RIGHT | F_UPPER | F_UNIT | F_TRANSA, kb, k, 1.0, A11, lda, A01, lda));
        hipLaunchKernelGGL(k_diag_right, dim3(grid_for((long)k * kb)), dim3(256), 0, s, k, kb, A01,
                           lda, A11, lda, W, (long)k);
        MOPS_CHECK_HIP(hipGetLastError());
        MOPS_TRY(gemm_d(s, false, true, k, k, kb, -1.0, A01, lda, W, k, 1.0, A, lda, GM_UPPER));
      }
      e = k;
    }
  }
  if (info) hipLaunchKernelGGL(k_sym_info, dim3(1), dim3(1), 0, s, st, info);
  if (err_host) {
    SymStep h;
    MOPS_CHECK_HIP(hipMemcpyAsync(&h, st, sizeof(SymStep), hipMemcpyDeviceToHost, s));
    MOPS_CHECK_HIP(hipStreamSynchronize(s));
    *err_host = h.err;
  }
  return ST_OK;
}

size_t bk_ws_bytes(int n) { return n > 0 ? 1024 : 0; }

int bk_d(hipStream_t s, bool upper, int n, double* A, long lda, int64_t* ipiv, long npiv) {
  if (npiv > 0) {
    hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(npiv)), dim3(256), 0, s, ipiv, npiv, (int64_t)0);
    MOPS_CHECK_HIP(hipGetLastError());
  }
  if (n <= 0) return ST_OK;
  SymStep* st = (SymStep*)ws_carve(sizeof(SymStep));
  if (!st) {
    set_error("bk: workspace too small");
    return ST_NOMEM;
  }
  hipLaunchKernelGGL(k_sym_init, dim3(1), dim3(1), 0, s, st, upper ? n : 0);
  if (upper) return sym_steps<true, true>(s, n, A, lda, ipiv, st, 0, 0);
  return sym_steps<true, false>(s, n, A, lda, ipiv, st, 0, 0);
}

size_t ldl_piv_solve_ws_bytes(int n, int nrhs) {
  if (n <= 0 || nrhs <= 0) return 0;
  return (size_t)n * nrhs * sizeof(double) + (size_t)n * sizeof(int) +
         trsm_ws_bytes(F_LEFT, n, nrhs) + 1024;
}

int ldl_piv_solve_d(hipStream_t s, bool upper, int n, int nrhs, const double* A, long lda,
                    const int64_t* ipiv, int npiv, double* B, long ldb) {
  if (n <= 0 || nrhs <= 0) return ST_OK;
  // SolveLDL (ldl.go:486-511): B = P B (the interchanges in order for LOWER,
  // last to first for UPPER), the unit-triangle solve, the diagonal, the
  // transposed unit-triangle solve, then B = P^T B. P is applied as one
  // gather into a workspace copy and undone as one scatter back into B.
  double* W = (double*)ws_carve((size_t)n * nrhs * sizeof(double));
  int* perm = (int*)ws_carve((size_t)n * sizeof(int));
  if (!W || !perm) {
    set_error("ldl solve: workspace too small");
    return ST_NOMEM;
  }
  const int np = npiv < n ? npiv : n;
  hipLaunchKernelGGL(k_pivots_to_perm, dim3(1), dim3(256), 0, s, n, np, ipiv, upper ? 0 : 1, perm);
  const int g = grid_for((long)n * nrhs);
  hipLaunchKernelGGL(k_permute_rows, dim3(g), dim3(256), 0, s, n, nrhs, perm, B, ldb, W, (long)n, 1);
  MOPS_CHECK_HIP(hipGetLastError());
  const int tri = upper ? F_UPPER : F_LOWER;
  MOPS_TRY(trsm_d(s, F_LEFT | tri | F_UNIT, n, nrhs, 1.0, A, lda, W, n));
  hipLaunchKernelGGL(k_solve_diag_left, dim3(g), dim3(256), 0, s, n, nrhs, W, (long)n, A, lda);
  MOPS_CHECK_HIP(hipGetLastError());
  MOPS_TRY(trsm_d(s, F_LEFT | tri | F_UNIT | F_TRANSA, n, nrhs, 1.0, A, lda, W, n));
  hipLaunchKernelGGL(k_permute_rows, dim3(g), dim3(256), 0, s, n, nrhs, perm, W, (long)n, B, ldb, 0);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

int bk_solve_d(hipStream_t s, bool upper, int n, int nrhs, const double* A, long lda,
               const int64_t* ipiv, double* B, long ldb) {
  if (n <= 0 || nrhs <= 0) return ST_OK;
  if (n <= 4 * BKS_T) {
    if (upper)
      hipLaunchKernelGGL((k_bk_solve_lds<true, 4>), dim3(nrhs), dim3(BKS_T), 0, s, n, A, lda, ipiv,
                         B, ldb);
    else
      hipLaunchKernelGGL((k_bk_solve_lds<false, 4>), dim3(nrhs), dim3(BKS_T), 0, s, n, A, lda, ipiv,
                         B, ldb);
  } else if (n <= BKS_MAX) {
    if (upper)
      hipLaunchKernelGGL((k_bk_solve_lds<true, 8>), dim3(nrhs), dim3(BKS_T), 0, s, n, A, lda, ipiv,
                         B, ldb);
    else
      hipLaunchKernelGGL((k_bk_solve_lds<false, 8>), dim3(nrhs), dim3(BKS_T), 0, s, n, A, lda, ipiv,
                         B, ldb);
  } else if (upper) {
    hipLaunchKernelGGL(k_bk_solve<true>, dim3(nrhs), dim3(256), 0, s, n, A, lda, ipiv, B, ldb);
  } else {
    hipLaunchKernelGGL(k_bk_solve<false>, dim3(nrhs), dim3(256), 0, s, n, A, lda, ipiv, B, ldb);
  }
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

}  // namespace mops
