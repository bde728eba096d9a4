// C ABI: the other BLAS-3 kernels of calgo (SYMM, TRMM, SYR2K, triangular
// update), host-memory drop-ins with calgo's parameter lists and
// device-resident twins.
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

// ---- SYMM ------------------------------------------------------------------

int mops_d_dsymm(int flags, int M, int N, double alpha, const double* A, int lda, const double* B,
                 int ldb, double beta, double* C, int ldc, void* stream) {
  clear_error();
  if (M < 0) return -2;
  if (N < 0) return -3;
  const bool left = flags & F_LEFT;
  const int na = left ? M : N;
  if (lda < (na > 1 ? na : 1)) return -6;
  if (ldb < (M > 1 ? M : 1)) return -8;
  if (ldc < (M > 1 ? M : 1)) return -11;
  if (M == 0 || N == 0) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(symm_ws_bytes(na) + gemm_split_ws_bytes(M, N, na, sizeof(double))));
  return symm_d(as_stream(stream), left, !(flags & F_UPPER), M, N, alpha, A, lda, B, ldb, beta, C,
                ldc);
}

int mops_dmult_symm_blocked2(mops_mdata_t* C, const mops_mdata_t* A, const mops_mdata_t* B,
                             double alpha, double beta, int flags, int P, int S, int L, int R,
                             int E, int vlen, int NB, int MB) {
  (void)vlen;
  (void)NB;
  (void)MB;
  clear_error();
  if (L - S <= 0 || E - R <= 0) return ST_OK;  // calgo/dmmat_symm.c:263-265
  if (!C || !C->md) return -1;
  if (P > 0 && (!A || !A->md)) return -2;
  if (P > 0 && (!B || !B->md)) return -3;
  if (P < 0) return -7;
  const bool left = flags & F_LEFT;
  const bool lower = !(flags & F_UPPER);
  const int rows = E - R, cols = L - S;
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)P * P, nbb = left ? (size_t)P * cols : (size_t)rows * P,
               nc = (size_t)rows * cols;
  WsUse wu(s);
  MOPS_TRY(ws_reserve((2 * na + nbb + nc) * sizeof(double) + 2048 +
                      gemm_split_ws_bytes(rows, cols, P, sizeof(double))));
  double* dA = (double*)ws_carve(na * sizeof(double) + 8);
  double* dS = (double*)ws_carve(na * sizeof(double) + 8);
  double* dB = (double*)ws_carve(nbb * sizeof(double) + 8);
  double* dC = (double*)ws_carve(nc * sizeof(double) + 8);
  double* hC = C->md + R + (long)S * C->step;
  if (beta != 0.0) MOPS_TRY(h2d(dC, rows, hC, C->step, rows, cols, s));
  if (P > 0) {
    MOPS_TRY(h2d(dA, P, A->md, A->step, P, P, s));
    MOPS_TRY(sym_expand_d(s, lower, P, dA, P, dS, P));
    if (left) {
      // C[R:E, S:L] += alpha sym(A)[R:E, :] B[:, S:L]
      MOPS_TRY(h2d(dB, P, B->md + (long)S * B->step, B->step, P, cols, s));
      MOPS_TRY(gemm_d(s, false, false, rows, cols, P, alpha, dS + R, P, dB, P, beta, dC, rows));
    } else {
      // C[R:E, S:L] += alpha B[R:E, :] sym(A)[:, S:L]
      MOPS_TRY(h2d(dB, rows, B->md + R, B->step, rows, P, s));
      MOPS_TRY(gemm_d(s, false, false, rows, cols, P, alpha, dB, rows, dS + (long)S * P, P, beta,
                      dC, rows));
    }
  } else {
    MOPS_TRY(scale_tile_d(s, rows, cols, beta, dC, rows));
  }
  MOPS_TRY(d2h(hC, C->step, dC, rows, rows, cols, s));
  return take_status(s);
}

// ---- TRMM ------------------------------------------------------------------

int mops_d_dtrmm(int flags, int N, int nrhs, double alpha, const double* A, int lda, double* B,
                 int ldb, void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < (N > 1 ? N : 1)) return -6;
  const int brows = (flags & F_RIGHT) ? nrhs : N;
  if (ldb < (brows > 1 ? brows : 1)) return -8;
  if (N == 0 || nrhs == 0) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(trmm_ws_bytes(N, nrhs)));
  return trmm_d(as_stream(stream), flags, N, nrhs, alpha, A, lda, B, ldb);
}

int mops_dmmat_trmm_blk(mops_mdata_t* B, const mops_mdata_t* A, double alpha, int flags, int N,
                        int S, int E, int NB) {
  (void)NB;
  clear_error();
  if (E - S <= 0 || N <= 0) return ST_OK;  // calgo/dmmat_trid.c:790-791
  if (!B || !B->md) return -1;
  if (!A || !A->md) return -2;
  const bool right = flags & F_RIGHT;
  const int nrhs = E - S;
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)N * N, nb = (size_t)N * nrhs;
  WsUse wu(s);
  MOPS_TRY(ws_reserve((na + nb) * sizeof(double) + trmm_ws_bytes(N, nrhs) + 1024));
  double* dA = (double*)ws_carve(na * sizeof(double));
  double* dB = (double*)ws_carve(nb * sizeof(double));
  // LEFT: columns S:E of B; RIGHT: rows S:E, as dmmat_solve_blk indexes them
  double* hB = right ? B->md + S : B->md + (long)S * B->step;
  const int brows = right ? nrhs : N, bcols = right ? N : nrhs;
  MOPS_TRY(h2d(dA, N, A->md, A->step, N, N, s));
  MOPS_TRY(h2d(dB, brows, hB, B->step, brows, bcols, s));
  MOPS_TRY(trmm_d(s, flags, N, nrhs, alpha, dA, N, dB, brows));
  MOPS_TRY(d2h(hB, B->step, dB, brows, brows, bcols, s));
  return take_status(s);
}

// ---- SYR2K and the triangular update ----------------------------------------

int mops_d_dsyr2k(int flags, int N, int P, double alpha, const double* A, int lda,
                  const double* B, int ldb, double beta, double* C, int ldc, void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (P < 0) return -3;
  const int arows = (flags & F_TRANSA) ? P : N;
  if (lda < (arows > 1 ? arows : 1)) return -6;
  if (ldb < (arows > 1 ? arows : 1)) return -8;
  if (ldc < (N > 1 ? N : 1)) return -11;
  if (N == 0) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(gemm_split_ws_bytes(N, N, P, sizeof(double))));
  return syr2k_d(as_stream(stream), flags, N, P, alpha, A, lda, B, ldb, beta, C, ldc);
}

int mops_d_dtrmupd(int flags, int N, int P, double alpha, const double* A, int lda,
                   const double* B, int ldb, double beta, double* C, int ldc, void* stream) {
  clear_error();
  if (N < 0) return -2;
  if (P < 0) return -3;
  const int arows = (flags & F_TRANSA) ? P : N, brows = (flags & F_TRANSB) ? N : P;
  if (lda < (arows > 1 ? arows : 1)) return -6;
  if (ldb < (brows > 1 ? brows : 1)) return -8;
  if (ldc < (N > 1 ? N : 1)) return -11;
  if (N == 0) return ST_OK;
  WsUse wu(as_stream(stream));
  MOPS_TRY(ws_reserve(gemm_split_ws_bytes(N, N, P, sizeof(double))));
  return trmupd_d(as_stream(stream), flags, N, P, alpha, A, lda, B, ldb, beta, C, ldc);
}

int mops_dmmat_rank2_blk(mops_mdata_t* C, const mops_mdata_t* A, const mops_mdata_t* B,
                         double alpha, double beta, int flags, int P, int S, int E, int vlen,
                         int NB) {
  (void)vlen;
  (void)NB;
  clear_error();
  if (E - S <= 0 || P <= 0) return ST_OK;  // calgo/dmmat_rank.c:142-144
  if (!C || !C->md) return -1;
  if (!A || !A->md) return -2;
  if (!B || !B->md) return -3;
  const bool ta = flags & F_TRANSA;
  const int n = E - S;
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)n * P, nc = (size_t)n * n;
  WsUse wu(s);
  MOPS_TRY(ws_reserve((2 * na + nc) * sizeof(double) + 1024 +
                      gemm_split_ws_bytes(n, n, P, sizeof(double))));
  double* dA = (double*)ws_carve(na * sizeof(double));
  double* dB = (double*)ws_carve(na * sizeof(double));
  double* dC = (double*)ws_carve(nc * sizeof(double));
  // rows S:E of A and B (NOTRANS) or columns S:E (TRANSA)
  if (!ta) {
    MOPS_TRY(h2d(dA, n, A->md + S, A->step, n, P, s));
    MOPS_TRY(h2d(dB, n, B->md + S, B->step, n, P, s));
  } else {
    MOPS_TRY(h2d(dA, P, A->md + (long)S * A->step, A->step, P, n, s));
    MOPS_TRY(h2d(dB, P, B->md + (long)S * B->step, B->step, P, n, s));
  }
  double* hC = C->md + S + (long)S * C->step;
  MOPS_TRY(h2d(dC, n, hC, C->step, n, n, s));
  const long ld = ta ? P : n;
  MOPS_TRY(syr2k_d(s, flags, n, P, alpha, dA, ld, dB, ld, beta, dC, n));
  // the other triangle went in and comes back unchanged
  MOPS_TRY(d2h(hC, C->step, dC, n, n, n, s));
  return take_status(s);
}

int mops_dmmat_trmupd_blk(mops_mdata_t* C, const mops_mdata_t* A, const mops_mdata_t* B,
                          double alpha, double beta, int flags, int P, int S, int E, int vlen,
                          int NB) {
  (void)vlen;
  (void)NB;
  clear_error();
  if (E - S <= 0 || P <= 0) return ST_OK;  // calgo/dmmat_rank.c:305-307
  if (!C || !C->md) return -1;
  if (!A || !A->md) return -2;
  if (!B || !B->md) return -3;
  const bool ta = flags & F_TRANSA, tb = flags & F_TRANSB;
  const int n = E - S;
  hipStream_t s = need_stream();
  if (!s) return ST_HIP;
  const size_t na = (size_t)n * P, nc = (size_t)n * n;
  WsUse wu(s);
  MOPS_TRY(ws_reserve((2 * na + nc) * sizeof(double) + 1024 +
                      gemm_split_ws_bytes(n, n, P, sizeof(double))));
  double* dA = (double*)ws_carve(na * sizeof(double));
  double* dB = (double*)ws_carve(na * sizeof(double));
  double* dC = (double*)ws_carve(nc * sizeof(double));
  // rows S:E of op(A), columns S:E of op(B)
  if (!ta)
    MOPS_TRY(h2d(dA, n, A->md + S, A->step, n, P, s));
  else
    MOPS_TRY(h2d(dA, P, A->md + (long)S * A->step, A->step, P, n, s));
  if (!tb)
    MOPS_TRY(h2d(dB, P, B->md + (long)S * B->step, B->step, P, n, s));
  else
    MOPS_TRY(h2d(dB, n, B->md + S, B->step, n, P, s));
  double* hC = C->md + S + (long)S * C->step;
  MOPS_TRY(h2d(dC, n, hC, C->step, n, n, s));
  MOPS_TRY(trmupd_d(s, flags, n, P, alpha, dA, ta ? P : n, dB, tb ? n : P, beta, dC, n));
  MOPS_TRY(d2h(hC, C->step, dC, n, n, n, s));
  return take_status(s);
}

}  // extern "C"
