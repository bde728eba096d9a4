// C ABI: the symmetric indefinite factorizations with pivoting (LDL with 1x1
// pivots, Bunch-Kaufman) and their solves; host-memory drop-ins for the Go
// functions, device twins.
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

// The zero-pivot texts of unblkLowerLDL (ldl.go:60) and unblkUpperLDL
// (ldl.go:335)
const char* zero_pivot_text(bool upper) {
  return upper ? "zero on diagonal." : "zero value on diagonal";
}

// Host-memory factorization: A and the pivots through the device, f runs
// on the device copies; with_info: f sets the device int to the 0-based
// index of a zero pivot or -1, and 1 + that index is returned.
template <typename F>
int factor_on_device(mops_mdata_t* A, int64_t* ipiv, int npiv, int N, size_t ws, bool with_info,
                     F f) {
  if (A->step < N) return -1;
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)N * N;
  WsUse wu(s);
  MOPS_TRY(ws_reserve(na * sizeof(double) + (size_t)npiv * sizeof(int64_t) + ws + 2048));
  double* dA = (double*)ws_carve(na * sizeof(double));
  int64_t* dp = (int64_t*)ws_carve((size_t)npiv * sizeof(int64_t));
  int* dinfo = (int*)ws_carve(sizeof(int));
  MOPS_TRY(h2d(dA, N, A->md, A->step, N, N, s));
  MOPS_TRY(f(s, dA, dp, dinfo));
  MOPS_TRY(d2h(A->md, A->step, dA, N, N, N, s));
  int info = -1;
  MOPS_CHECK_HIP(
      hipMemcpyAsync(ipiv, dp, (size_t)npiv * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  if (with_info)
    MOPS_CHECK_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  MOPS_CHECK_HIP(hipStreamSynchronize(s));
  MOPS_TRY(take_status(s));
  return info >= 0 ? info + 1 : ST_OK;
}

// Host-memory solve: B (N x nrhs) = A^-1 B through the device.
template <typename F>
int solve_on_device(mops_mdata_t* B, const mops_mdata_t* A, const int64_t* ipiv, int npiv, int N,
                    int nrhs, size_t ws, F f) {
  if (A->step < N || B->step < N) return -1;
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)N * N, nb = (size_t)N * nrhs;
  WsUse wu(s);
  MOPS_TRY(ws_reserve((na + nb) * sizeof(double) + (size_t)npiv * sizeof(int64_t) + ws + 2048));
  double* dA = (double*)ws_carve(na * sizeof(double));
  double* dB = (double*)ws_carve(nb * sizeof(double));
  int64_t* dp = (int64_t*)ws_carve((size_t)(npiv > 0 ? npiv : 1) * sizeof(int64_t));
  MOPS_TRY(h2d(dA, N, A->md, A->step, N, N, s));
  MOPS_TRY(h2d(dB, N, B->md, B->step, N, nrhs, s));
  if (npiv > 0)
    MOPS_CHECK_HIP(
        hipMemcpyAsync(dp, ipiv, (size_t)npiv * sizeof(int64_t), hipMemcpyHostToDevice, s));
  MOPS_TRY(f(s, dA, dp, dB));
  MOPS_TRY(d2h(B->md, B->step, dB, N, N, nrhs, s));
  return take_status(s);
}

}  // namespace

extern "C" {

int mops_dsytrf_ldl(mops_mdata_t* A, int64_t* ipiv, int npiv, int flags, int N, int nb) {
  clear_error();
  if (!A || (N > 0 && !A->md)) return -1;
  if (npiv < 0 || (npiv > 0 && !ipiv)) return -2;
  if (N < 0) return -5;
  if (nb < 0) return -6;
  if (npiv < N) {
    set_error("pivot array shorter than A.Rows()");
    return -3;
  }
  for (int k = 0; k < npiv; ++k) ipiv[k] = 0;  // ldl.go:444-446
  if (N == 0) return ST_OK;
  const bool upper = !(flags & F_LOWER);
  const int rc = factor_on_device(A, ipiv, npiv, N, ldl_piv_ws_bytes(N, nb, upper), true,
                                  [&](hipStream_t s, double* dA, int64_t* dp, int* dinfo) {
                                    return ldl_piv_d(s, upper, N, nb, dA, N, dp, npiv, dinfo,
                                                     nullptr);
                                  });
  if (rc > 0) set_error("%s", zero_pivot_text(upper));
  return rc;
}

int mops_d_dsytrf_ldl(int flags, int N, int nb, double* A, int lda, int64_t* ipiv, int npiv,
                      int* info, void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (nb < 0) return -3;
  if (lda < (N > 1 ? N : 1)) return -5;
  if (npiv < N || (npiv > 0 && !ipiv)) {
    set_error("pivot array shorter than A.Rows()");
    return -7;
  }
  const bool upper = !(flags & F_LOWER);
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(ldl_piv_ws_bytes(N, nb, upper) + 1024));
  return ldl_piv_d(as_stream(stream), upper, N, nb, A, lda, ipiv, npiv, info, nullptr);
}

int mops_dsytrs_ldl(mops_mdata_t* B, const mops_mdata_t* A, const int64_t* ipiv, int npiv,
                    int flags, int N, int nrhs) {
  clear_error();
  if (!B || (N > 0 && nrhs > 0 && !B->md)) return -1;
  if (!A || (N > 0 && !A->md)) return -2;
  if (npiv < 0 || (npiv > 0 && !ipiv)) return -3;
  if (N < 0) return -6;
  if (nrhs < 0) return -7;
  if (N == 0 || nrhs == 0 || !(flags & (F_LOWER | F_UPPER))) return ST_OK;
  const bool upper = flags & F_UPPER;  // ldl.go:487: UPPER is tested first
  return solve_on_device(B, A, ipiv, npiv, N, nrhs, ldl_piv_solve_ws_bytes(N, nrhs),
                         [&](hipStream_t s, const double* dA, const int64_t* dp, double* dB) {
                           return ldl_piv_solve_d(s, upper, N, nrhs, dA, N, dp, npiv, dB, N);
                         });
}

int mops_d_dsytrs_ldl(int flags, int N, int nrhs, const double* A, int lda, const int64_t* ipiv,
                      int npiv, double* B, int ldb, void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < (N > 1 ? N : 1)) return -5;
  if (npiv < 0 || (npiv > 0 && !ipiv)) return -6;
  if (ldb < (N > 1 ? N : 1)) return -9;
  if (N == 0 || nrhs == 0 || !(flags & (F_LOWER | F_UPPER))) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(ldl_piv_solve_ws_bytes(N, nrhs)));
  return ldl_piv_solve_d(as_stream(stream), (flags & F_UPPER) != 0, N, nrhs, A, lda, ipiv, npiv, B,
                         ldb);
}

int mops_dsytrf_bk(mops_mdata_t* A, int64_t* ipiv, int npiv, int flags, int N, int nb) {
  clear_error();
  if (!A || (N > 0 && !A->md)) return -1;
  if (npiv < 0 || (npiv > 0 && !ipiv)) return -2;
  if (N < 0) return -5;
  if (nb < 0) return -6;
  for (int k = 0; k < npiv; ++k) ipiv[k] = 0;  // ldlbk.go:1116-1118
  // ldlbk.go:1124-1136: LOWER, else UPPER, else nothing
  if (N == 0 || !(flags & (F_LOWER | F_UPPER))) return ST_OK;
  if (npiv < N) {
    set_error("pivot array shorter than A.Rows()");
    return -3;
  }
  const bool upper = !(flags & F_LOWER);
  return factor_on_device(A, ipiv, npiv, N, bk_ws_bytes(N), false,
                          [&](hipStream_t s, double* dA, int64_t* dp, int*) {
                            return bk_d(s, upper, N, dA, N, dp, npiv);
                          });
}

int mops_d_dsytrf_bk(int flags, int N, int nb, double* A, int lda, int64_t* ipiv, int npiv,
                     void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (nb < 0) return -3;
  if (lda < (N > 1 ? N : 1)) return -5;
  if (npiv < 0 || (npiv > 0 && !ipiv)) return -6;
  const bool any = flags & (F_LOWER | F_UPPER);
  if (any && npiv < N) {
    set_error("pivot array shorter than A.Rows()");
    return -7;
  }
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(bk_ws_bytes(N) + 1024));
  return bk_d(as_stream(stream), !(flags & F_LOWER), any ? N : 0, A, lda, ipiv, npiv);
}

int mops_dsytrs_bk(mops_mdata_t* B, const mops_mdata_t* A, const int64_t* ipiv, int npiv,
                   int flags, int N, int nrhs) {
  clear_error();
  if (!B || (N > 0 && nrhs > 0 && !B->md)) return -1;
  if (!A || (N > 0 && !A->md)) return -2;
  if (npiv < 0 || (npiv > 0 && !ipiv)) return -3;
  if (N < 0) return -6;
  if (nrhs < 0) return -7;
  if (N == 0 || nrhs == 0 || !(flags & (F_LOWER | F_UPPER))) return ST_OK;
  if (npiv < N) {
    set_error("pivot array shorter than A.Rows()");
    return -3;
  }
  const bool upper = !(flags & F_LOWER);  // ldlbks.go:304: LOWER is tested first
  return solve_on_device(B, A, ipiv, npiv, N, nrhs, 1024,
                         [&](hipStream_t s, const double* dA, const int64_t* dp, double* dB) {
                           return bk_solve_d(s, upper, N, nrhs, dA, N, dp, dB, N);
                         });
}

int mops_d_dsytrs_bk(int flags, int N, int nrhs, const double* A, int lda, const int64_t* ipiv,
                     int npiv, double* B, int ldb, void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < (N > 1 ? N : 1)) return -5;
  if (npiv < 0 || (npiv > 0 && !ipiv)) return -6;
  if (ldb < (N > 1 ? N : 1)) return -9;
  if (N == 0 || nrhs == 0 || !(flags & (F_LOWER | F_UPPER))) return ST_OK;
  if (npiv < N) {
    set_error("pivot array shorter than A.Rows()");
    return -7;
  }
  return bk_solve_d(as_stream(stream), !(flags & F_LOWER), N, nrhs, A, lda, ipiv, B, ldb);
}

}  // extern "C"
