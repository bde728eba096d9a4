// C ABI: LU and LDL^T without pivoting, the LDL^T solve and the triangular
// inverse (host-memory drop-ins for the Go functions, device twins).
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

// Runs f(device copy of A, its leading dimension, stream) on an M x N host
// matrix: copy in, compute, copy out, check the status word.
template <typename F>
int on_device_copy(mops_mdata_t* A, int M, int N, size_t ws, F f) {
  if (!A || !A->md) return -1;
  if (A->step < M) return -1;
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)M * N;
  WsUse wu(s);
  MOPS_TRY(ws_reserve(na * sizeof(double) + ws + 1024));
  double* dA = (double*)ws_carve(na * sizeof(double));
  MOPS_TRY(h2d(dA, M, A->md, A->step, M, N, s));
  MOPS_TRY(f(dA, (long)M, s));
  MOPS_TRY(d2h(A->md, A->step, dA, M, M, N, s));
  return take_status(s);
}

}  // namespace

extern "C" {

int mops_dgetrf_nopiv(mops_mdata_t* A, int M, int N, int nb) {
  clear_error();
  if (M < 0) return -2;
  if (N < 0) return -3;
  if (nb < 0) return -4;
  if (M == 0 || N == 0) return ST_OK;
  return on_device_copy(A, M, N, getrf_nopiv_ws_bytes(M, N), [&](double* dA, long ld,
                                                                 hipStream_t s) {
    return getrf_nopiv_d(s, M, N, dA, ld);
  });
}

int mops_d_dgetrf_nopiv(int M, int N, int nb, double* A, int lda, void* stream) {
  clear_error();
  if (M < 0) return -1;
  if (N < 0) return -2;
  if (nb < 0) return -3;
  if (lda < (M > 1 ? M : 1)) return -5;
  if (M == 0 || N == 0) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(getrf_nopiv_ws_bytes(M, N)));
  return getrf_nopiv_d(as_stream(stream), M, N, A, lda);
}

int mops_dsytrf_nopiv(mops_mdata_t* A, int flags, int N, int nb) {
  clear_error();
  if (N < 0) return -3;
  if (nb < 0) return -4;
  if (N == 0) return ST_OK;
  const bool lower = flags & F_LOWER;
  return on_device_copy(A, N, N, sytrf_nopiv_ws_bytes(N), [&](double* dA, long ld,
                                                              hipStream_t s) {
    return sytrf_nopiv_d(s, lower, N, dA, ld);
  });
}

int mops_d_dsytrf_nopiv(int flags, int N, int nb, double* A, int lda, void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (nb < 0) return -3;
  if (lda < (N > 1 ? N : 1)) return -5;
  if (N == 0) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(sytrf_nopiv_ws_bytes(N)));
  return sytrf_nopiv_d(as_stream(stream), (flags & F_LOWER) != 0, N, A, lda);
}

int mops_dsytrs_nopiv(mops_mdata_t* B, const mops_mdata_t* A, int flags, int N, int nrhs) {
  clear_error();
  if (!B || (N > 0 && nrhs > 0 && !B->md)) return -1;
  if (!A || (N > 0 && !A->md)) return -2;
  if (N < 0) return -4;
  if (nrhs < 0) return -5;
  if (N == 0 || nrhs == 0) return ST_OK;
  // SolveLDLnoPiv does nothing unless a triangle bit is set (ldlnp.go:252-264)
  if (!(flags & (F_LOWER | F_UPPER))) return ST_OK;
  if (A->step < N || B->step < N) return -1;
  const bool lower = !(flags & F_UPPER);
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)N * N, nb = (size_t)N * nrhs;
  WsUse wu(s);
  MOPS_TRY(ws_reserve((na + nb) * sizeof(double) + sytrs_nopiv_ws_bytes(N, nrhs) + 1024));
  double* dA = (double*)ws_carve(na * sizeof(double));
  double* dB = (double*)ws_carve(nb * sizeof(double));
  MOPS_TRY(h2d(dA, N, A->md, A->step, N, N, s));
  MOPS_TRY(h2d(dB, N, B->md, B->step, N, nrhs, s));
  MOPS_TRY(sytrs_nopiv_d(s, lower, N, nrhs, dA, N, dB, N));
  MOPS_TRY(d2h(B->md, B->step, dB, N, N, nrhs, s));
  return take_status(s);
}

int mops_d_dsytrs_nopiv(int flags, int N, int nrhs, const double* A, int lda, double* B, int ldb,
                        void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < (N > 1 ? N : 1)) return -5;
  if (ldb < (N > 1 ? N : 1)) return -7;
  if (N == 0 || nrhs == 0 || !(flags & (F_LOWER | F_UPPER))) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(sytrs_nopiv_ws_bytes(N, nrhs)));
  return sytrs_nopiv_d(as_stream(stream), !(flags & F_UPPER), N, nrhs, A, lda, B, ldb);
}

int mops_dinverse_trm(mops_mdata_t* A, int flags, int N, int nb) {
  clear_error();
  if (N < 0) return -3;
  if (nb < 0) return -4;
  if (N == 0) return ST_OK;
  const bool lower = flags & F_LOWER, unit = flags & F_UNIT;
  return on_device_copy(A, N, N, inverse_trm_ws_bytes(N), [&](double* dA, long ld,
                                                              hipStream_t s) {
    return inverse_trm_d(s, lower, unit, N, dA, ld);
  });
}

int mops_d_dinverse_trm(int flags, int N, int nb, double* A, int lda, void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (nb < 0) return -3;
  if (lda < (N > 1 ? N : 1)) return -5;
  if (N == 0) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(inverse_trm_ws_bytes(N)));
  return inverse_trm_d(as_stream(stream), (flags & F_LOWER) != 0, (flags & F_UNIT) != 0, N, A,
                       lda);
}

}  // extern "C"
