// Solve A X = B by LU in reduced precision plus FP64 iterative refinement
// (no counterpart in matops; it reuses the LU path's kernels).
//
//   factor:  prec FP32 -> getrf_mixed_d (FP32 trailing updates, FP64 panels)
//            prec BF16 -> getrf_mixed_d with the trailing products on BF16
//                         operands (FP32 accumulation, FP32 storage)
//            prec FP64 -> getrf_d (the plain FP64 path; one solve suffices)
//   loop:    D = solve(LU, R)   (FP64 triangular solves on the factors)
//            X += D
//            R = B - A X        (FP64 GEMM with the original A)
//            eta = max_j ||r_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)
//   until eta <= tol or maxit solves.
#include <cstring>

#include "../../include/mops.h"
#include "mops_lu.hpp"
#include "mops_solvers.hpp"

namespace mops {
namespace {

// non-negative doubles order like their bit patterns
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* p, double v) {
  atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// Infinity norm max_i sum_j |A[i][j]| in two passes: part[c][i] = sum of
// |A[i][j]| over column chunk c (thread = row, workgroups over rows x
// chunks), then the chunks of each row are added in chunk order (so the
// value does not depend on timing) and the maximum is taken.
constexpr int NORM_CHUNKS = 64;

__global__ void k_row_abs_partial(int n, const double* __restrict__ A, long lda, int chunk,
                                  double* part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j0 = blockIdx.y * chunk, j1 = min(n, j0 + chunk);
  double s = 0.0;
  for (int j = j0; j < j1; ++j) s += fabs(A[i + (long)j * lda]);
  part[(long)blockIdx.y * n + i] = s;
}

__global__ void k_inf_norm_finish(int n, int nch, const double* __restrict__ part,
                                  unsigned long long* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int c = 0; c < nch; ++c) s += part[(long)c * n + i];
  // a NaN row sum has larger bits than +Inf, so the maximum keeps it
  atomic_max_nonneg(out, s != s ? __longlong_as_double(0x7ff8000000000000ll) : s);
}

// NaN-propagating maximum (fmax drops a NaN operand): a NaN anywhere makes
// the result NaN, so a non-finite residual or solution is never hidden.
__device__ __forceinline__ double max_nan(double m, double v) { return (v > m || v != v) ? v : m; }

// out[j] = max_i |M[i][j]| (NaN if any entry is NaN); workgroup = column
__global__ __launch_bounds__(256) void k_col_absmax(int n, const double* __restrict__ M, long ld,
                                                    double* out) {
  __shared__ double red[256];
  const int j = blockIdx.x;
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) m = max_nan(m, fabs(M[i + (long)j * ld]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = max_nan(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

// bits[z * nrhs + j] = max over the rows of |M[i][j]| for M = R (z = 0)
// and X (z = 1), as the bits of a non-negative double (a NaN is stored as
// the quiet NaN, whose bits exceed those of +Inf, so the maximum keeps it):
// 2048 rows per workgroup, the bits cleared beforehand
constexpr int ABSMAX_ROWS = 2048;
__global__ __launch_bounds__(256) void k_absmax_rx(int n, int nrhs, const double* __restrict__ R,
                                                   long ldr, const double* __restrict__ X, long ldx,
                                                   unsigned long long* bits) {
  __shared__ double red[256];
  const int j = blockIdx.y;
  const double* M = blockIdx.z ? X + (long)j * ldx : R + (long)j * ldr;
  const int i0 = blockIdx.x * ABSMAX_ROWS, i1 = min(n, i0 + ABSMAX_ROWS);
  double m = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) m = max_nan(m, fabs(M[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = max_nan(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double v = red[0];
    atomic_max_nonneg(&bits[blockIdx.z * nrhs + j],
                      v != v ? __longlong_as_double(0x7ff8000000000000ll) : v);
  }
}

__global__ void k_zero_u64n(unsigned long long* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

__global__ void k_add_tile(int n, int nrhs, const double* __restrict__ D, long ldd, double* X,
                           long ldx) {
  const long total = (long)n * nrhs;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % n), j = (int)(t / n);
    X[i + (long)j * ldx] += D[i + (long)j * ldd];
  }
}

__global__ void k_zero_tile(int n, int nrhs, double* X, long ldx) {
  const long total = (long)n * nrhs;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x)
    X[(t % n) + (t / n) * ldx] = 0.0;
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

// eta = max_j rmax_j / (anorm xmax_j + bmax_j); +Inf when any of the norms
// is not finite (a singular A gives Inf/NaN in X, which must never read as
// converged). rmax and xmax come as k_absmax_rx's bits, which are cleared
// here for the next iteration.
__global__ void k_eta(int nrhs, unsigned long long* bits, const double* bmax,
                      const unsigned long long* anorm_bits, double* eta) {
  if (threadIdx.x != 0) return;
  const double an = __longlong_as_double((long long)*anorm_bits);
  double e = isfinite(an) ? 0.0 : INFINITY;
  for (int j = 0; j < nrhs; ++j) {
    const double rmax = __longlong_as_double((long long)bits[j]);
    const double xmax = __longlong_as_double((long long)bits[nrhs + j]);
    bits[j] = 0ull;
    bits[nrhs + j] = 0ull;
    if (!isfinite(rmax) || !isfinite(xmax) || !isfinite(bmax[j])) {
      e = INFINITY;
      continue;
    }
    const double den = an * xmax + bmax[j];
    const double q = den > 0.0 ? rmax / den : (rmax > 0.0 ? INFINITY : 0.0);
    e = max_nan(e, q);
  }
  *eta = e;
}

inline int grid_for(long total) {
  long b = (total + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

}  // namespace

size_t gesv_ir_ws_bytes(int n, int nrhs, int prec) {
  if (n <= 0 || nrhs <= 0) return 0;
  size_t b = (size_t)n * n * sizeof(double);          // LU
  b += 2 * (size_t)n * nrhs * sizeof(double);         // R, D
  b += (size_t)n * sizeof(int) + 6 * nrhs * sizeof(double) + 4096;
  b += (size_t)NORM_CHUNKS * n * sizeof(double);     // row-sum chunks of the norm
  const size_t f = prec != 0 ? getrf_mixed_ws_bytes(n, 256) : getrf_ws_bytes(n, n, 256);
  const size_t r = getrs_prep_ws_bytes(n, nrhs) + gemm_split_ws_bytes(n, nrhs, n, sizeof(double));
  return b + (f > r ? f : r);
}

int gesv_ir_d(hipStream_t s, int n, int nrhs, const double* A, long lda, const double* B,
              long ldb, double* X, long ldx, int prec, double tol, int maxit,
              mops_ir_report* res) {
  std::memset(res, 0, sizeof(*res));
  if (n <= 0 || nrhs <= 0) {
    res->converged = 1;
    return ST_OK;
  }
  if (prec != 0 && prec != 1 && prec != 2) {
    set_error("gesv_ir: precision %d is not available (0 = FP64, 1 = FP32, 2 = BF16)", prec);
    return ST_BADARG;
  }
  if (maxit < 1) maxit = 1;
  double* LU = (double*)ws_carve((size_t)n * n * sizeof(double));
  double* R = (double*)ws_carve((size_t)n * nrhs * sizeof(double));
  double* D = (double*)ws_carve((size_t)n * nrhs * sizeof(double));
  int* perm = (int*)ws_carve((size_t)n * sizeof(int));
  double* cm = (double*)ws_carve(4 * (size_t)nrhs * sizeof(double) + 64);
  unsigned long long* bits = (unsigned long long*)ws_carve(2 * (size_t)nrhs * sizeof(double) + 64);
  unsigned long long* anorm = (unsigned long long*)ws_carve(64);
  int* err = (int*)ws_carve(64);
  double* part = (double*)ws_carve((size_t)NORM_CHUNKS * n * sizeof(double));
  if (!LU || !R || !D || !perm || !cm || !bits || !anorm || !err || !part) {
    set_error("gesv_ir: workspace too small");
    return ST_NOMEM;
  }
  double *bmax = cm + 2 * nrhs, *eta = cm + 3 * nrhs;
  // own events: the factorizations use the shared synchronization events
  struct Events {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  for (hipEvent_t& x : ev.e) MOPS_CHECK_HIP(hipEventCreate(&x));
  hipEvent_t e0 = ev.e[0], e1 = ev.e[1], e2 = ev.e[2];
  MOPS_CHECK_HIP(hipEventRecord(e0, s));
  {
    const size_t mark = ws_mark();
    if (prec == 1 || prec == 2) {
      MOPS_TRY(getrf_mixed_d(s, n, 256, A, lda, LU, n, nullptr, perm, err, prec == 2));
    } else {
      MOPS_TRY(copy_tile_d(s, n, n, A, lda, LU, n));
      MOPS_TRY(getrf_d(s, n, n, 256, LU, n, nullptr, perm, err));
    }
    ws_release(mark);
  }
  MOPS_CHECK_HIP(hipEventRecord(e1, s));
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, anorm);
  {
    const int chunk = ceil_div(n, NORM_CHUNKS), nch = ceil_div(n, chunk);
    hipLaunchKernelGGL(k_row_abs_partial, dim3(ceil_div(n, 256), nch), dim3(256), 0, s, n, A, lda,
                       chunk, part);
    hipLaunchKernelGGL(k_inf_norm_finish, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, nch, part,
                       anorm);
  }
  hipLaunchKernelGGL(k_col_absmax, dim3(nrhs), dim3(256), 0, s, n, B, ldb, bmax);
  hipLaunchKernelGGL(k_zero_tile, dim3(grid_for((long)n * nrhs)), dim3(256), 0, s, n, nrhs, X, ldx);
  MOPS_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_zero_u64n, dim3(1), dim3(256), 0, s, bits, 2 * nrhs);
  MOPS_TRY(copy_tile_d(s, n, nrhs, B, ldb, R, n));
  // the inverses of the diagonal blocks of L and U, once for every solve
  const size_t pmark = ws_mark();
  GetrsPrep gp;
  MOPS_TRY(getrs_prep_d(s, n, nrhs, LU, n, gp));
  const dim3 amax_grid(ceil_div(n, ABSMAX_ROWS), nrhs, 2);
  for (int it = 1; it <= maxit; ++it) {
    MOPS_TRY(getrs_prep_solve_d(s, n, nrhs, LU, n, perm, R, n, D, n, gp));
    hipLaunchKernelGGL(k_add_tile, dim3(grid_for((long)n * nrhs)), dim3(256), 0, s, n, nrhs, D,
                       (long)n, X, ldx);
    MOPS_TRY(copy_tile_d(s, n, nrhs, B, ldb, R, n));
    MOPS_TRY(gemm_d(s, false, false, n, nrhs, n, -1.0, A, lda, X, ldx, 1.0, R, n));
    hipLaunchKernelGGL(k_absmax_rx, amax_grid, dim3(256), 0, s, n, nrhs, R, (long)n, X, ldx, bits);
    hipLaunchKernelGGL(k_eta, dim3(1), dim3(1), 0, s, nrhs, bits, bmax, anorm, eta);
    MOPS_CHECK_HIP(hipGetLastError());
    double e = 0.0;
    MOPS_CHECK_HIP(hipMemcpyAsync(&e, eta, sizeof(double), hipMemcpyDeviceToHost, s));
    MOPS_CHECK_HIP(hipStreamSynchronize(s));
    res->iterations = it;
    res->backward_error = e;
    if (it == 1) res->first_error = e;
    if (e <= tol) {
      res->converged = 1;
      break;
    }
  }
  ws_release(pmark);
  MOPS_CHECK_HIP(hipEventRecord(e2, s));
  MOPS_CHECK_HIP(hipEventSynchronize(e2));
  float f_ms = 0.f, r_ms = 0.f;
  MOPS_CHECK_HIP(hipEventElapsedTime(&f_ms, e0, e1));
  MOPS_CHECK_HIP(hipEventElapsedTime(&r_ms, e1, e2));
  res->factor_ms = f_ms;
  res->refine_ms = r_ms;
  int eflag = 0;
  MOPS_CHECK_HIP(hipMemcpy(&eflag, err, sizeof(int), hipMemcpyDeviceToHost));
  if (eflag) {
    set_error("gesv_ir: LU panel grid barrier timed out");
    return ST_HIP;
  }
  return ST_OK;
}

}  // namespace mops
