/*
 * CPU oracle of matops' dense FP64 path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this code, and only as the checker / the CPU baseline. The GPU product
 * (matops_amd/libmops.so) never links or calls it.
 *
 * Two layers, mirroring the reference:
 *  - kernels with calgo's C interface (calgo/cmops.h): GEMM, TRSM, SYRK and
 *    the level-1/2 helpers the unblocked panels call. They are reached
 *    through a table, so the same glue runs either on this file set's own
 *    restatement ("port", calgo_port.c) or on the reference's calgo C
 *    compiled by oracle/Makefile into oracle/_ref/ ("reference").
 *  - the Go glue restated in C line by line (matops_port.c): Mult,
 *    SolveTrm, RankUpdateSym, the MV* helpers, pivots, unblocked/blocked
 *    Cholesky and LU, and the solves.
 */
#ifndef MOPS_ORACLE_H
#define MOPS_ORACLE_H

#include <stdint.h>

typedef struct { double *md; int step; } o_mdata_t; /* calgo/cmops.h:50-53 */
typedef struct { double *md; int inc; } o_mvec_t;   /* calgo/cmops.h:55-58 */

/* A FloatMatrix view: pointer at (0,0), rows, cols, leading index. */
typedef struct { double *p; int rows, cols, ld; } oview;

enum {
  O_TRANSA = 0x1, O_TRANSB = 0x2, O_LOWER = 0x4, O_UPPER = 0x8, O_LEFT = 0x10,
  O_RIGHT = 0x20, O_UNIT = 0x40, O_TRANS = 0x80
};

typedef struct {
  void (*dmult_mm_blocked3)(o_mdata_t *C, const o_mdata_t *A, const o_mdata_t *B, double alpha,
                            double beta, int flags, int P, int S, int L, int R, int E, int vlen,
                            int NB, int MB);
  void (*dmmat_solve_blk)(o_mdata_t *B, const o_mdata_t *A, double alpha, int flags, int N, int S,
                          int E, int NB);
  void (*dmmat_rank_blk)(o_mdata_t *C, const o_mdata_t *A, double alpha, double beta, int flags,
                         int P, int S, int E, int vlen, int NB);
  void (*dmvec_symv_rank)(o_mdata_t *A, const o_mvec_t *X, double alpha, int flags, int S, int L,
                          int NB);
  void (*dmvec_solve_blocked)(o_mvec_t *X, const o_mdata_t *A, int flags, int N, int NB);
  void (*dmult_gemv_blocked)(o_mvec_t *Y, const o_mdata_t *A, const o_mvec_t *X, double alpha,
                             double beta, int flags, int S, int L, int R, int E, int vlen, int MB);
  double (*dvec_dot)(const o_mvec_t *X, const o_mvec_t *Y, double alpha, int N);
  void (*dvec_swap)(o_mvec_t *X, o_mvec_t *Y, int N);
  void (*dvec_invscal)(o_mvec_t *X, double alpha, int N);
  void (*dmmat_solve_unb)(o_mdata_t *B, const o_mdata_t *A, double alpha, int flags, int N, int S,
                          int E);
  void (*dmult_symm_blocked2)(o_mdata_t *C, const o_mdata_t *A, const o_mdata_t *B, double alpha,
                              double beta, int flags, int P, int S, int L, int R, int E, int vlen,
                              int NB, int MB);
  void (*dmmat_trmm_blk)(o_mdata_t *B, const o_mdata_t *A, double alpha, int flags, int N, int S,
                         int E, int NB);
  void (*dmmat_rank2_blk)(o_mdata_t *C, const o_mdata_t *A, const o_mdata_t *B, double alpha,
                          double beta, int flags, int P, int S, int E, int vlen, int NB);
  void (*dmmat_trmupd_blk)(o_mdata_t *C, const o_mdata_t *A, const o_mdata_t *B, double alpha,
                           double beta, int flags, int P, int S, int E, int vlen, int NB);
  int (*dvec_iamax)(const o_mvec_t *X, int N);
} okern_t;

/* calgo_port.c: the restated kernels */
extern const okern_t o_port_kernels;

/* kernel selection: 0 = port, 1 = reference calgo loaded from `path`.
 * Returns 0 on success, -1 when the shared object or a symbol is missing. */
int o_use_kernels(int which, const char *path);
int o_kernels_in_use(void);
int o_use_kernel_mask(int mask, const char *path);

/* tuning globals of mmult.go:33-44 */
void o_BlockingParams(int mb, int nb, int kb);
int o_NumWorkers(int w);

/* Go API restated (matops_port.c). Return 0 = nil error, otherwise an error
 * code: 1 size mismatch / not square / short pivot array, 2 panic. */
int o_Mult(oview C, oview A, oview B, double alpha, double beta, int flags);
int o_SolveTrm(oview B, oview A, double alpha, int flags);
int o_RankUpdateSym(oview C, oview A, double alpha, double beta, int flags);
int o_DecomposeCHOL(oview A, int flags, int nb, int *panic_index, double *panic_value);
void o_SolveCHOL(oview B, oview A, int flags);
int o_DecomposeLU(oview A, int64_t *pivots, int npiv, int nb);
int o_SolveLU(oview B, oview A, const int64_t *pivots, int npiv, int flags);
int o_MultSym(oview C, oview A, oview B, double alpha, double beta, int flags);
int o_MultTrm(oview B, oview A, double alpha, int flags);
int o_RankUpdate2Sym(oview C, oview A, oview B, double alpha, double beta, int flags);
int o_UpdateTrm(oview C, oview A, oview B, double alpha, double beta, int flags);
int o_DecomposeLUnoPiv(oview A, int nb);
int o_DecomposeLDLnoPiv(oview A, int flags, int nb);
void o_SolveLDLnoPiv(oview B, oview A, int flags);
int o_InverseTrm(oview A, int flags, int nb);
/* DecomposeLDL / SolveLDL (ldl.go:25-511) and DecomposeBK / SolveBK
 * (ldlbk.go:48-458, 1114-1140; ldlbks.go). Pivots are Go ints (int64).
 * Returns 0, 1 for an error with its text in *msg ("A not a square matrix",
 * "zero value on diagonal", "zero on diagonal.", "Workspace too small"). */
int o_DecomposeLDL(oview A, oview W, int64_t *ipiv, int npiv, int flags, int nb, const char **msg);
void o_SolveLDL(oview B, oview A, const int64_t *ipiv, int npiv, int flags);
int o_DecomposeBK(oview A, oview W, int64_t *ipiv, int npiv, int flags, int nb, const char **msg);
void o_SolveBK(oview B, oview A, const int64_t *ipiv, int npiv, int flags);
/* IAMax of the reference (mvec.go:405-420 -> dvec_iamax) on n elements */
int o_iamax(const double *x, int inc, int n);

/* Reference adjoint of the Cholesky factorization (closed form, built on
 * the oracle's TRSM/GEMM): given the lower factor L and Lbar (lower),
 * writes the symmetric Abar = 1/2 * (S + S^T) with
 * S = L^-T * Phi(L^T * Lbar) * L^-1, Phi = lower triangle, halved diagonal. */
void o_chol_adjoint(int n, const double *L, int ldl, const double *Lbar, int ldb, double *Abar,
                    int lda);

/* The kernel-level boundary itself (calgo/cmops.h:133-138, 203-214): the
 * selected kernel set called with explicit block indices, as calgo.DMult /
 * DSolveUnblk / DSolveBlk / DSymmRankBlk pass them (calgo/cmops.go:68-240). */
void o_kern_gemm(double *C, int ldc, const double *A, int lda, const double *B, int ldb,
                 double alpha, double beta, int flags, int P, int S, int L, int R, int E);
void o_kern_trsm(double *B, int ldb, const double *A, int lda, double alpha, int flags, int N,
                 int S, int E);
void o_kern_trsm_unb(double *B, int ldb, const double *A, int lda, double alpha, int flags, int N,
                     int S, int E);
void o_kern_syrk(double *C, int ldc, const double *A, int lda, double alpha, double beta,
                 int flags, int P, int S, int E);

/* wall-clock helper (CLOCK_MONOTONIC, seconds) */
double o_now(void);

#endif
