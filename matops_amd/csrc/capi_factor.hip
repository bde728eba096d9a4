// C ABI: factorizations and solves (Cholesky here; LU, mixed precision and
// the Cholesky adjoint in their own files).
#include <cmath>
#include <cstdio>

#include "../../include/mops.h"
#include "mops_solvers.hpp"

using namespace mops;

namespace {

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

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

hipStream_t need_stream() {
  hipStream_t s = host_stream();
  if (!s) set_error("could not create a HIP stream");
  return s;
}

}  // namespace

extern "C" {

int mops_d_dpotrf(int flags, int N, int nb, double* A, int lda, int* info, void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (nb < 0) return -3;
  if (lda < (N > 1 ? N : 1)) return -5;
  if (N == 0) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(potrf_ws_bytes(N, nb)));
  return potrf_d(as_stream(stream), flags, N, nb, A, lda, info, nullptr);
}

int mops_d_dpotrs(int flags, int N, int nrhs, const double* A, int lda, double* B, int ldb,
                  void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < (N > 1 ? N : 1)) return -5;
  if (ldb < (N > 1 ? N : 1)) return -7;
  if (N == 0 || nrhs == 0) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(potrs_ws_bytes(flags, N, nrhs)));
  return potrs_d(as_stream(stream), flags, N, nrhs, A, lda, B, ldb);
}

int mops_dpotrf(mops_mdata_t* A, int flags, int N, int nb) {
  clear_error();
  if (!A || (N > 0 && !A->md)) return -1;
  if (N < 0) return -3;
  if (nb < 0) return -4;
  if (N == 0) return ST_OK;
  if (A->step < N) return -1;
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)N * N;
  WsUse wu(s);
  MOPS_TRY(ws_reserve(na * sizeof(double) + potrf_ws_bytes(N, nb) + 2048));
  double* dA = (double*)ws_carve(na * sizeof(double));
  int* dinfo = (int*)ws_carve(sizeof(int));
  double* dbad = (double*)ws_carve(sizeof(double));
  MOPS_TRY(h2d(dA, N, A->md, A->step, N, N, s));
  MOPS_TRY(potrf_d(s, flags, N, nb, dA, N, dinfo, dbad));
  MOPS_TRY(d2h(A->md, A->step, dA, N, N, N, s));
  int info = -1;
  double bad = 0.0;
  MOPS_CHECK_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  MOPS_CHECK_HIP(hipMemcpyAsync(&bad, dbad, sizeof(double), hipMemcpyDeviceToHost, s));
  MOPS_CHECK_HIP(hipStreamSynchronize(s));
  MOPS_TRY(take_status(s));
  if (info == -2) {
    set_error("Cholesky leaf: wait for the diagonal block timed out");
    return ST_HIP;
  }
  if (info >= 0) {
    // the panic text of unblockedCHOL (chol.go:35)
    set_error("illegal value at %d: %e", info, bad);
    return info + 1;
  }
  return ST_OK;
}

int mops_d_dpotrf_adjoint(int N, int nb, const double* L, int ldl, double* Abar, int lda,
                          void* stream) {
  clear_error();
  if (N < 0) return -1;
  if (nb < 0) return -2;
  if (ldl < (N > 1 ? N : 1)) return -4;
  if (lda < (N > 1 ? N : 1)) return -6;
  if (N == 0) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(chol_adjoint_ws_bytes(N, nb)));
  return chol_adjoint_d(as_stream(stream), N, nb, L, ldl, Abar, lda);
}

int mops_dpotrf_adjoint(const mops_mdata_t* L, mops_mdata_t* Lbar, int N, int nb) {
  clear_error();
  if (!L || (N > 0 && !L->md)) return -1;
  if (!Lbar || (N > 0 && !Lbar->md)) return -2;
  if (N < 0) return -3;
  if (nb < 0) return -4;
  if (N == 0) return ST_OK;
  if (L->step < N) return -1;
  if (Lbar->step < N) return -2;
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)N * N;
  WsUse wu(s);
  MOPS_TRY(ws_reserve(2 * na * sizeof(double) + chol_adjoint_ws_bytes(N, nb) + 2048));
  double* dL = (double*)ws_carve(na * sizeof(double));
  double* dB = (double*)ws_carve(na * sizeof(double));
  MOPS_TRY(h2d(dL, N, L->md, L->step, N, N, s));
  MOPS_TRY(h2d(dB, N, Lbar->md, Lbar->step, N, N, s));
  MOPS_TRY(chol_adjoint_d(s, N, nb, dL, N, dB, N));
  MOPS_TRY(d2h(Lbar->md, Lbar->step, dB, N, N, N, s));
  return take_status(s);
}

int mops_dpotrs(mops_mdata_t* B, const mops_mdata_t* A, int flags, int N, int nrhs) {
  clear_error();
  if (!B || (N > 0 && nrhs > 0 && !B->md)) return -1;
  if (!A || (N > 0 && !A->md)) return -2;
  if (N < 0) return -4;
  if (nrhs < 0) return -5;
  if (N == 0 || nrhs == 0) return ST_OK;
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)N * N, nb = (size_t)N * nrhs;
  WsUse wu(s);
  MOPS_TRY(ws_reserve((na + nb) * sizeof(double) + potrs_ws_bytes(flags, N, nrhs) + 1024));
  double* dA = (double*)ws_carve(na * sizeof(double));
  double* dB = (double*)ws_carve(nb * sizeof(double));
  MOPS_TRY(h2d(dA, N, A->md, A->step, N, N, s));
  MOPS_TRY(h2d(dB, N, B->md, B->step, N, nrhs, s));
  MOPS_TRY(potrs_d(s, flags, N, nrhs, dA, N, dB, N));
  MOPS_TRY(d2h(B->md, B->step, dB, N, N, nrhs, s));
  return take_status(s);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// LU with partial pivoting

#include "mops_lu.hpp"

extern "C" {

int mops_d_dgetrf(int M, int N, int nb, double* A, int lda, int64_t* ipiv, int* perm,
                  int* status, void* stream) {
  clear_error();
  if (M < 0) return -1;
  if (N < 0) return -2;
  if (nb < 0) return -3;
  if (lda < (M > 1 ? M : 1)) return -5;
  if (M == 0 || N == 0) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(getrf_ws_bytes(M, N, nb)));
  return getrf_d(as_stream(stream), M, N, nb, A, lda, reinterpret_cast<long long*>(ipiv), perm,
                 status);
}

int mops_d_dgetrs(int flags, int N, int nrhs, const double* A, int lda, const int64_t* ipiv,
                  int npiv, const int* perm, double* B, int ldb, void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < (N > 1 ? N : 1)) return -5;
  if (!perm && !ipiv) return -6;
  if (ldb < (N > 1 ? N : 1)) return -10;
  if (N == 0 || nrhs == 0) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(getrs_ws_bytes(N, nrhs)));
  return getrs_d(as_stream(stream), flags, N, nrhs, A, lda,
                 reinterpret_cast<const long long*>(ipiv), npiv, perm, B, ldb);
}

int mops_dgetrf(mops_mdata_t* A, int64_t* pivots, int npiv, int M, int N, int nb) {
  clear_error();
  if (!A || (M > 0 && N > 0 && !A->md)) return -1;
  const int mn = M < N ? M : N;
  if (npiv < mn) {
    set_error("pivot array < min(A.Rows(),A.Cols())");  // lu.go:260-262
    return -3;
  }
  if (M < 0) return -4;
  if (N < 0) return -5;
  if (nb < 0) return -6;
  for (int k = 0; k < npiv; ++k) pivots[k] = 0;  // lu.go:264-266
  if (M == 0 || N == 0) return ST_OK;
  if (A->step < M) return -1;
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)M * N;
  WsUse wu(s);
  MOPS_TRY(ws_reserve(na * sizeof(double) + (size_t)mn * 8 + getrf_ws_bytes(M, N, nb) + 4096));
  double* dA = (double*)ws_carve(na * sizeof(double));
  long long* dp = (long long*)ws_carve((size_t)mn * sizeof(long long));
  int* dst = (int*)ws_carve(sizeof(int));
  MOPS_TRY(h2d(dA, M, A->md, A->step, M, N, s));
  MOPS_TRY(getrf_d(s, M, N, nb, dA, M, dp, nullptr, dst));
  MOPS_TRY(d2h(A->md, A->step, dA, M, M, N, s));
  int st = 0;
  MOPS_CHECK_HIP(hipMemcpyAsync(pivots, dp, (size_t)mn * sizeof(long long),
                                hipMemcpyDeviceToHost, s));
  MOPS_CHECK_HIP(hipMemcpyAsync(&st, dst, sizeof(int), hipMemcpyDeviceToHost, s));
  MOPS_CHECK_HIP(hipStreamSynchronize(s));
  MOPS_TRY(take_status(s));
  if (st != 0) {
    set_error("LU panel: grid barrier timed out");
    return ST_HIP;
  }
  return ST_OK;
}

int mops_dgetrs(mops_mdata_t* B, const mops_mdata_t* A, const int64_t* pivots, int npiv,
                int flags, int N, int nrhs) {
  clear_error();
  if (!B || (N > 0 && nrhs > 0 && !B->md)) return -1;
  if (!A || (N > 0 && !A->md)) return -2;
  if (N < 0) return -6;
  if (nrhs < 0) return -7;
  if (N == 0 || nrhs == 0) return ST_OK;
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)N * N, nb = (size_t)N * nrhs;
  const int np = npiv < N ? npiv : N;
  WsUse wu(s);
  MOPS_TRY(ws_reserve((na + nb) * sizeof(double) + (size_t)np * 8 + getrs_ws_bytes(N, nrhs) +
                      4096));
  double* dA = (double*)ws_carve(na * sizeof(double));
  double* dB = (double*)ws_carve(nb * sizeof(double));
  long long* dp = (long long*)ws_carve((size_t)(np > 0 ? np : 1) * sizeof(long long));
  MOPS_TRY(h2d(dA, N, A->md, A->step, N, N, s));
  MOPS_TRY(h2d(dB, N, B->md, B->step, N, nrhs, s));
  if (np > 0)
    MOPS_CHECK_HIP(hipMemcpyAsync(dp, pivots, (size_t)np * sizeof(long long),
                                  hipMemcpyHostToDevice, s));
  MOPS_TRY(getrs_d(s, flags, N, nrhs, dA, N, dp, np, nullptr, dB, N));
  MOPS_TRY(d2h(B->md, B->step, dB, N, N, nrhs, s));
  return take_status(s);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Mixed precision LU with iterative refinement

extern "C" {

int mops_d_dgesv_ir(int N, int nrhs, const double* A, int lda, const double* B, int ldb, double* X,
                    int ldx, int prec, double tol, int maxit, mops_ir_report* rep, void* stream) {
  clear_error();
  if (N < 0) return -1;
  if (nrhs < 0) return -2;
  if (lda < (N > 1 ? N : 1)) return -4;
  if (ldb < (N > 1 ? N : 1)) return -6;
  if (ldx < (N > 1 ? N : 1)) return -8;
  if (prec != MOPS_PREC_FP64 && prec != MOPS_PREC_FP32 && prec != MOPS_PREC_BF16) {
    set_error("precision %d is not available (MOPS_PREC_FP64, MOPS_PREC_FP32, MOPS_PREC_BF16)",
              prec);
    return -9;
  }
  if (!rep) return -12;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(gesv_ir_ws_bytes(N, nrhs, prec)));
  return gesv_ir_d(as_stream(stream), N, nrhs, A, lda, B, ldb, X, ldx, prec, tol, maxit, rep);
}

int mops_dgesv_ir(const mops_mdata_t* A, const mops_mdata_t* B, mops_mdata_t* X, int N, int nrhs,
                  int prec, double tol, int maxit, mops_ir_report* rep) {
  clear_error();
  if (!A || (N > 0 && !A->md)) return -1;
  if (!B || (N > 0 && nrhs > 0 && !B->md)) return -2;
  if (!X || (N > 0 && nrhs > 0 && !X->md)) return -3;
  if (N < 0) return -4;
  if (nrhs < 0) return -5;
  if (prec != MOPS_PREC_FP64 && prec != MOPS_PREC_FP32 && prec != MOPS_PREC_BF16) {
    set_error("precision %d is not available (MOPS_PREC_FP64, MOPS_PREC_FP32, MOPS_PREC_BF16)",
              prec);
    return -6;
  }
  if (!rep) return -9;
  if (N == 0 || nrhs == 0) {
    *rep = mops_ir_report{0, 1, 0.0, 0.0, 0.0, 0.0};
    return ST_OK;
  }
  if (A->step < N || B->step < N || X->step < N) return -1;
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)N * N, nb = (size_t)N * nrhs;
  WsUse wu(s);
  MOPS_TRY(ws_reserve((na + 2 * nb) * sizeof(double) + gesv_ir_ws_bytes(N, nrhs, prec) + 4096));
  double* dA = (double*)ws_carve(na * sizeof(double));
  double* dB = (double*)ws_carve(nb * sizeof(double));
  double* dX = (double*)ws_carve(nb * sizeof(double));
  MOPS_TRY(h2d(dA, N, A->md, A->step, N, N, s));
  MOPS_TRY(h2d(dB, N, B->md, B->step, N, nrhs, s));
  MOPS_TRY(gesv_ir_d(s, N, nrhs, dA, N, dB, N, dX, N, prec, tol, maxit, rep));
  MOPS_TRY(d2h(X->md, X->step, dX, N, N, nrhs, s));
  return take_status(s);
}

}  // extern "C"
