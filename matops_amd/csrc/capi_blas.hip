// C ABI: runtime queries and the BLAS-3 entry points (GEMM, TRSM, SYRK).
#include <cstring>

#include "../../include/mops.h"
#include "mops_kernels.hpp"
#include "mops_solvers.hpp"

using namespace mops;

namespace {

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// Copies a rows x cols column-major tile between host and device memory.
int h2d(double* d, long ldd, const double* h, long ldh, int rows, int cols, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return ST_OK;
  MOPS_CHECK_HIP(hipMemcpy2DAsync(d, ldd * sizeof(double), h, ldh * sizeof(double),
                                  rows * sizeof(double), cols, hipMemcpyHostToDevice, s));
  return ST_OK;
}

int d2h(double* h, long ldh, const double* d, long ldd, int rows, int cols, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return ST_OK;
  MOPS_CHECK_HIP(hipMemcpy2DAsync(h, ldh * sizeof(double), d, ldd * sizeof(double),
                                  rows * sizeof(double), cols, hipMemcpyDeviceToHost, s));
  return ST_OK;
}

// B[:, S:E] (LEFT) or B[S:E, :] (RIGHT) = alpha op(A)^-1 B, op(A) N x N:
// the host-memory solve behind dmmat_solve_blk and dmmat_solve_unb.
int solve_host(mops_mdata_t* B, const mops_mdata_t* A, double alpha, int flags, int N, int S,
               int E) {
  if (E - S <= 0 || N <= 0) return ST_OK;  // calgo/cmops.go:203, dmmat_solve.c:875
  if (!B || !B->md) return -1;
  if (!A || !A->md) return -2;
  const bool right = flags & F_RIGHT;
  const int nrhs = E - S;
  hipStream_t s = host_stream();
  if (!s) {
    set_error("could not create a HIP stream");
    return ST_HIP;
  }
  const size_t na = (size_t)N * N, nb = (size_t)N * nrhs;
  WsUse wu(s);
  MOPS_TRY(ws_reserve((na + nb) * sizeof(double) + trsm_ws_bytes(flags, N, nrhs) + 1024));
  double* dA = (double*)ws_carve(na * sizeof(double));
  double* dB = (double*)ws_carve(nb * sizeof(double));
  // LEFT: columns S:E of B (dmmat_solve.c:425); RIGHT: rows S:E (:405)
  double* hB = right ? B->md + S : B->md + (long)S * B->step;
  const int brows = right ? nrhs : N, bcols = right ? N : nrhs;
  MOPS_TRY(h2d(dA, N, A->md, A->step, N, N, s));
  MOPS_TRY(h2d(dB, brows, hB, B->step, brows, bcols, s));
  MOPS_TRY(trsm_d(s, flags, N, nrhs, alpha, dA, N, dB, brows));
  MOPS_TRY(d2h(hB, B->step, dB, brows, brows, bcols, s));
  return take_status(s);
}

}  // namespace

extern "C" {

const char* mops_last_error(void) { return last_error(); }

int mops_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

void mops_thread_release(void) { thread_release(); }

int mops_d_status(void* stream) {
  clear_error();
  int st = take_status(as_stream(stream));
  // the no-allocation GEMM keeps its status word in its own scratch
  if (int* ns = noalloc_status_word()) {
    StatusOverride so(ns);
    const int st2 = take_status(as_stream(stream));
    if (st == ST_OK) st = st2;
  }
  return st;
}

void mops_trace_enable(int on) { trace_enable(on); }

int mops_trace_collect(int cls, int* count, double* ms, double* flops) {
  if (cls < 0 || cls >= TR_NCLASS) return -1;
  if (!count || !ms || !flops) return -2;
  return trace_collect(cls, count, ms, flops);
}

int mops_d_dgemm(int flags, int M, int N, int P, double alpha, const double* A, int lda,
                 const double* B, int ldb, double beta, double* C, int ldc, void* stream) {
  clear_error();
  const bool ta = flags & F_TRANSA, tb = flags & F_TRANSB;
  if (M < 0) return -2;
  if (N < 0) return -3;
  if (P < 0) return -4;
  const int arows = ta ? P : M, brows = tb ? N : P;
  if (lda < (arows > 1 ? arows : 1)) return -7;
  if (ldb < (brows > 1 ? brows : 1)) return -9;
  if (ldc < (M > 1 ? M : 1)) return -12;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(gemm_split_ws_bytes(M, N, P, sizeof(double))));
  return gemm_d(as_stream(stream), ta, tb, M, N, P, alpha, A, lda, B, ldb, beta, C, ldc);
}

int mops_dmult_mm_blocked3(mops_mdata_t* C, const mops_mdata_t* A, const mops_mdata_t* B,
                           double alpha, double beta, int flags, int P, int S, int L, int R,
                           int E, int vlen, int NB, int MB) {
  (void)vlen;
  (void)NB;
  (void)MB;
  clear_error();
  if (!C || !C->md) return -1;
  if (L - S <= 0 || E - R <= 0) return ST_OK;  // calgo/dmmat_gemm.c:376-379
  if (P > 0 && (!A || !A->md)) return -2;
  if (P > 0 && (!B || !B->md)) return -3;
  if (P < 0) return -7;
  const bool ta = flags & F_TRANSA, tb = flags & F_TRANSB;
  const int rows = E - R, cols = L - S;
  hipStream_t s = host_stream();
  if (!s) {
    set_error("could not create a HIP stream");
    return ST_HIP;
  }
  const size_t na = (size_t)rows * P, nbb = (size_t)P * cols, nc = (size_t)rows * cols;
  WsUse wu(s);
  MOPS_TRY(ws_reserve((na + nbb + nc) * sizeof(double) + 1024 +
                      gemm_split_ws_bytes(rows, cols, P, sizeof(double))));
  double* dA = (double*)ws_carve(na * sizeof(double) + 8);
  double* dB = (double*)ws_carve(nbb * sizeof(double) + 8);
  double* dC = (double*)ws_carve(nc * sizeof(double) + 8);
  // The block of C and the rows/columns of op(A)/op(B) it needs
  // (calgo/dmmat_gemm.c:399-411 index the same way).
  if (P > 0) {
    if (!ta)
      MOPS_TRY(h2d(dA, rows, A->md + R, A->step, rows, P, s));
    else
      MOPS_TRY(h2d(dA, P, A->md + (long)R * A->step, A->step, P, rows, s));
    if (!tb)
      MOPS_TRY(h2d(dB, P, B->md + (long)S * B->step, B->step, P, cols, s));
    else
      MOPS_TRY(h2d(dB, cols, B->md + S, B->step, cols, P, s));
  }
  double* hC = C->md + R + (long)S * C->step;
  if (beta != 0.0) MOPS_TRY(h2d(dC, rows, hC, C->step, rows, cols, s));
  MOPS_TRY(gemm_d(s, ta, tb, rows, cols, P, alpha, dA, ta ? P : rows, dB, tb ? cols : P, beta,
                  dC, rows));
  MOPS_TRY(d2h(hC, C->step, dC, rows, rows, cols, s));
  return take_status(s);
}

int mops_d_dtrsm(int flags, int N, int nrhs, double alpha, const double* A, int lda, double* B,
                 int ldb, void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < (N > 1 ? N : 1)) return -6;
  const int brows = (flags & F_RIGHT) ? nrhs : N;
  if (ldb < (brows > 1 ? brows : 1)) return -8;
  if (N == 0 || nrhs == 0) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(trsm_ws_bytes(flags, N, nrhs)));
  return trsm_d(as_stream(stream), flags, N, nrhs, alpha, A, lda, B, ldb);
}

int mops_d_dsyrk(int flags, int N, int P, double alpha, const double* A, int lda, double beta,
                 double* C, int ldc, void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (P < 0) return -3;
  const bool ta = flags & F_TRANSA;
  const int arows = ta ? P : N;
  if (lda < (arows > 1 ? arows : 1)) return -6;
  if (ldc < (N > 1 ? N : 1)) return -9;
  const int mode = (flags & F_LOWER) ? GM_LOWER : GM_UPPER;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(gemm_split_ws_bytes(N, N, P, sizeof(double))));
  return gemm_d(as_stream(stream), ta, !ta, N, N, P, alpha, A, lda, A, lda, beta, C, ldc, mode);
}

int mops_dmmat_solve_blk(mops_mdata_t* B, const mops_mdata_t* A, double alpha, int flags, int N,
                         int S, int E, int NB) {
  (void)NB;
  clear_error();
  return solve_host(B, A, alpha, flags, N, S, E);
}

int mops_dmmat_solve_unb(mops_mdata_t* B, const mops_mdata_t* A, double alpha, int flags, int N,
                         int S, int E) {
  clear_error();
  // dmmat_solve_unb picks the lower triangle iff LOWER is set
  // (calgo/dmmat_solve.c:409, 423); the blocked kernels test UPPER first
  const int f = (flags & F_LOWER) ? (flags & ~F_UPPER) : ((flags & ~F_LOWER) | F_UPPER);
  return solve_host(B, A, alpha, f, N, S, E);
}

int mops_dmmat_rank_blk(mops_mdata_t* C, const mops_mdata_t* A, double alpha, double beta,
                        int flags, int P, int S, int E, int vlen, int NB) {
  (void)vlen;
  (void)NB;
  clear_error();
  if (E - S <= 0 || P <= 0) return ST_OK;  // calgo/dmmat_rank.c:73-75
  if (!C || !C->md) return -1;
  if (!A || !A->md) return -2;
  const bool ta = flags & F_TRANSA;
  const int n = E - S;
  hipStream_t s = host_stream();
  if (!s) {
    set_error("could not create a HIP stream");
    return ST_HIP;
  }
  const size_t na = (size_t)n * P, nc = (size_t)n * n;
  WsUse wu(s);
  MOPS_TRY(ws_reserve((na + nc) * sizeof(double) + 1024 +
                      gemm_split_ws_bytes(n, n, P, sizeof(double))));
  double* dA = (double*)ws_carve(na * sizeof(double));
  double* dC = (double*)ws_carve(nc * sizeof(double));
  // rows S:E of A (NOTRANS) or columns S:E (TRANSA), as dmmat_rank.c:83-124
  if (!ta)
    MOPS_TRY(h2d(dA, n, A->md + S, A->step, n, P, s));
  else
    MOPS_TRY(h2d(dA, P, A->md + (long)S * A->step, A->step, P, n, s));
  double* hC = C->md + S + (long)S * C->step;
  MOPS_TRY(h2d(dC, n, hC, C->step, n, n, s));
  const int mode = (flags & F_LOWER) ? GM_LOWER : GM_UPPER;
  MOPS_TRY(gemm_d(s, ta, !ta, n, n, P, alpha, dA, ta ? P : n, dA, ta ? P : n, beta, dC, n, mode));
  MOPS_TRY(d2h(hC, C->step, dC, n, n, n, s));
  return take_status(s);
}

}  // extern "C"
