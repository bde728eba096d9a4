/*
 * The Go half of matops' dense FP64 path restated in C -- TEST
 * INFRASTRUCTURE (the oracle). Go is not available to build the reference,
 * so its glue (views, partitioning, pivots, the blocked/unblocked
 * algorithms and the cgo calls) is restated here statement by statement;
 * each function cites what it follows. The kernels are reached through the
 * okern_t table (the port in calgo_port.c, or the reference's calgo C
 * loaded from oracle/_ref).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

static okern_t K;
static int K_init = 0;
static int K_which = 0;

static const okern_t *kern(void) {
  if (!K_init) {
    K = o_port_kernels;
    K_init = 1;
  }
  return &K;
}

int o_use_kernels(int which, const char *path) {
  if (which == 0) {
    K = o_port_kernels;
    K_init = 1;
    K_which = 0;
    return 0;
  }
  void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return -1;
  okern_t k;
#define SYM(name)                                    \
  *(void **)(&k.name) = dlsym(h, #name);             \
  if (!k.name) return -1;
  SYM(dmult_mm_blocked3)
  SYM(dmmat_solve_blk)
  SYM(dmmat_rank_blk)
  SYM(dmvec_symv_rank)
  SYM(dmvec_solve_blocked)
  SYM(dmult_gemv_blocked)
  SYM(dvec_dot)
  SYM(dvec_swap)
  SYM(dvec_invscal)
  SYM(dmmat_solve_unb)
  SYM(dmult_symm_blocked2)
  SYM(dmmat_trmm_blk)
  SYM(dmmat_rank2_blk)
  SYM(dmmat_trmupd_blk)
  SYM(dvec_iamax)
#undef SYM
  K = k;
  K_init = 1;
  K_which = 1;
  return 0;
}

int o_kernels_in_use(void) { return K_which; }

/* Mixes the two kernel sets: bit i of `mask` takes the i-th okern_t entry
 * from the reference (0 dmult_mm_blocked3, 1 dmmat_solve_blk, 2
 * dmmat_rank_blk, 3 dmvec_symv_rank, 4 dmvec_solve_blocked, 5
 * dmult_gemv_blocked, 6 dvec_dot, 7 dvec_swap, 8 dvec_invscal, 9
 * dmmat_solve_unb, 10 dmult_symm_blocked2, 11 dmmat_trmm_blk, 12
 * dmmat_rank2_blk, 13 dmmat_trmupd_blk, 14 dvec_iamax). Used to
 * localise differences between the restatement and the reference. */
int o_use_kernel_mask(int mask, const char *path) {
  if (o_use_kernels(1, path) != 0) return -1;
  okern_t ref = K;
  okern_t mix = o_port_kernels;
  void **dst = (void **)&mix;
  void **src = (void **)&ref;
  for (int i = 0; i < (int)(sizeof(okern_t) / sizeof(void *)); i++)
    if (mask & (1 << i)) dst[i] = src[i];
  K = mix;
  K_which = mask ? 2 : 0;
  return 0;
}

double o_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ---- tuning globals, mmult.go:33-44, 63-73 ---- */
static int vpLen = 196, nB = 68, mB = 68;
static int nWorker = 1;
static const int64_t limitOne = 200 * 200 * 200;

void o_BlockingParams(int mb, int nb, int kb) {
  vpLen = kb;
  nB = nb;
  mB = mb;
}

int o_NumWorkers(int w) {
  int old = nWorker;
  nWorker = w;
  return old;
}

/* ---- views (hrautila/matrix FloatMatrix semantics) ---- */
static oview sub(oview A, int r, int c, int rows, int cols) {
  oview v = {A.p + r + (ptrdiff_t)c * A.ld, rows, cols, A.ld};
  return v;
}
static int numel(oview A) { return A.rows * A.cols; }
static o_mdata_t md(oview A) {
  o_mdata_t m = {A.p, A.ld};
  return m;
}
/* row vectors step by the leading index, column vectors by 1 */
static o_mvec_t mv_rows(oview X) { /* MVMult & co: X.Rows() == 1 -> inc = ld (mvec.go:45-48) */
  o_mvec_t v = {X.p, X.rows == 1 ? X.ld : 1};
  return v;
}
static o_mvec_t mv_cols(oview X) { /* Dot, InvScale, Swap: X.Cols() != 1 -> inc = ld (mvec.go:286) */
  o_mvec_t v = {X.p, X.cols != 1 ? X.ld : 1};
  return v;
}

/* ---- Mult, mmult.go:198-256 with divideWork/scheduleWork (:137-188) ---- */
static int blockIndex4(int i, int r, int sz) { /* mmult.go:105-110 */
  if (i == r) return sz;
  return i * sz / r - ((i * sz / r) & 0x3);
}

static void divideWork(int rows, int cols, int workers, int *colWorkers, int *rowWorkers) {
  /* mmult.go:137-169 */
  int nwsqrt = (int)sqrt((double)workers);
  int issquare = nwsqrt * nwsqrt == workers;
  if (workers == 2 || (workers & 0x1) != 0) {
    if (cols > rows) {
      *colWorkers = workers;
      *rowWorkers = 1;
    } else {
      *rowWorkers = workers;
      *colWorkers = 1;
    }
  } else if (issquare) {
    *colWorkers = nwsqrt;
    *rowWorkers = nwsqrt;
  } else {
    if (cols > rows) {
      *rowWorkers = 2;
      *colWorkers = workers / 2;
    } else {
      *colWorkers = 2;
      *rowWorkers = workers / 2;
    }
  }
}

typedef struct {
  o_mdata_t C, A, B;
  double alpha, beta;
  int flags, P, S, L, R, E;
} mult_task;

static void *mult_worker(void *arg) {
  mult_task *t = (mult_task *)arg;
  kern()->dmult_mm_blocked3(&t->C, &t->A, &t->B, t->alpha, t->beta, t->flags, t->P, t->S, t->L,
                            t->R, t->E, vpLen, nB, mB);
  return NULL;
}

int o_Mult(oview C, oview A, oview B, double alpha, double beta, int flags) {
  int ar = A.rows, ac = A.cols, br = B.rows, bc = B.cols, cr = C.rows, cc = C.cols;
  int empty, ok;
  switch (flags & (O_TRANSA | O_TRANSB)) {
    case O_TRANSA | O_TRANSB:
      empty = ac == 0 || br == 0;
      ok = cr == ac && cc == br && ar == bc;
      break;
    case O_TRANSA:
      empty = ac == 0 || bc == 0;
      ok = cr == ac && cc == bc && ar == br;
      break;
    case O_TRANSB:
      empty = ar == 0 || br == 0;
      ok = cr == ar && cc == br && ac == bc;
      break;
    default:
      empty = ar == 0 || bc == 0;
      ok = cr == ar && cc == bc && ac == br;
  }
  if (empty) return 0;
  if (!ok) return 1; /* "Mult: size mismatch" */
  int64_t psize = (int64_t)numel(C) * (int64_t)A.cols;
  int P = flags & O_TRANSA ? A.rows : A.cols;
  o_mdata_t Cm = md(C), Am = md(A), Bm = md(B);
  if (nWorker <= 1 || psize <= limitOne) {
    kern()->dmult_mm_blocked3(&Cm, &Am, &Bm, alpha, beta, flags, P, 0, C.cols, 0, C.rows, vpLen,
                              nB, mB);
    return 0;
  }
  int colworks, rowworks;
  divideWork(C.rows, C.cols, nWorker, &colworks, &rowworks);
  int ntask = colworks * rowworks;
  pthread_t *th = malloc(sizeof(pthread_t) * ntask);
  mult_task *tk = malloc(sizeof(mult_task) * ntask);
  int n = 0;
  for (int k = 0; k < colworks; k++) {
    int colstart = blockIndex4(k, colworks, C.cols), colend = blockIndex4(k + 1, colworks, C.cols);
    for (int l = 0; l < rowworks; l++) {
      int rowstart = blockIndex4(l, rowworks, C.rows), rowend = blockIndex4(l + 1, rowworks, C.rows);
      mult_task t = {Cm, Am, Bm, alpha, beta, flags, P, colstart, colend, rowstart, rowend};
      tk[n] = t;
      pthread_create(&th[n], NULL, mult_worker, &tk[n]);
      n++;
    }
  }
  for (int i = 0; i < n; i++) pthread_join(th[i], NULL);
  free(th);
  free(tk);
  return 0;
}

/* ---- MultSym, mmult.go:266-326 -> calgo.DMultSymm, calgo/cmops.go:95-118 ---- */
typedef struct {
  o_mdata_t C, A, B;
  double alpha, beta;
  int flags, P, S, L, R, E;
} symm_task;

static void *symm_worker(void *arg) {
  symm_task *t = (symm_task *)arg;
  kern()->dmult_symm_blocked2(&t->C, &t->A, &t->B, t->alpha, t->beta, t->flags, t->P, t->S, t->L,
                              t->R, t->E, vpLen, nB, mB);
  return NULL;
}

int o_MultSym(oview C, oview A, oview B, double alpha, double beta, int flags) {
  int ar = A.rows, ac = A.cols, br = B.rows, bc = B.cols, cr = C.rows, cc = C.cols;
  int empty, ok;
  switch (flags & (O_TRANSA | O_TRANSB)) {
    case O_TRANSA | O_TRANSB:
      empty = ac == 0 || br == 0;
      ok = ar == ac && cr == ac && cc == br && ar == bc;
      break;
    case O_TRANSA:
      empty = ac == 0 || bc == 0;
      ok = ar == ac && cr == ac && cc == bc && ar == br;
      break;
    case O_TRANSB:
      empty = ar == 0 || br == 0;
      ok = ar == ac && cr == ar && cc == br && ac == bc;
      break;
    default:
      empty = ar == 0 || bc == 0;
      ok = ar == ac && cr == ar && cc == bc && ac == br;
  }
  if (empty) return 0;
  if (!ok) return 1; /* "MultSym: size mismatch" */
  int64_t psize = (int64_t)numel(C) * (int64_t)A.cols;
  o_mdata_t Cm = md(C), Am = md(A), Bm = md(B);
  if (nWorker <= 1 || psize <= limitOne) {
    kern()->dmult_symm_blocked2(&Cm, &Am, &Bm, alpha, beta, flags, A.cols, 0, C.cols, 0, C.rows,
                                vpLen, nB, mB);
    return 0;
  }
  int colworks, rowworks;
  divideWork(C.rows, C.cols, nWorker, &colworks, &rowworks);
  int ntask = colworks * rowworks;
  pthread_t *th = malloc(sizeof(pthread_t) * ntask);
  symm_task *tk = malloc(sizeof(symm_task) * ntask);
  int n = 0;
  for (int k = 0; k < colworks; k++) {
    int cs = blockIndex4(k, colworks, C.cols), ce = blockIndex4(k + 1, colworks, C.cols);
    for (int l = 0; l < rowworks; l++) {
      int rs = blockIndex4(l, rowworks, C.rows), re = blockIndex4(l + 1, rowworks, C.rows);
      symm_task t = {Cm, Am, Bm, alpha, beta, flags, A.cols, cs, ce, rs, re};
      tk[n] = t;
      pthread_create(&th[n], NULL, symm_worker, &tk[n]);
      n++;
    }
  }
  for (int i = 0; i < n; i++) pthread_join(th[i], NULL);
  free(th);
  free(tk);
  return 0;
}

/* ---- MultTrm, mmult.go:338-372 -> calgo.DTrmmBlk, calgo/cmops.go:146-166 ---- */
int o_MultTrm(oview B, oview A, double alpha, int flags) {
  int ok = 1, empty = 0;
  int br = B.rows, bc = B.cols, ar = A.rows, ac = A.cols;
  switch (flags & (O_LEFT | O_RIGHT)) {
    case O_LEFT:
      empty = br == 0 || bc == 0;
      ok = br == ac && ac == ar;
      break;
    case O_RIGHT:
      empty = bc == 0 || br == 0;
      ok = bc == ar && ac == ar;
      break;
  }
  if (empty) return 0;
  if (!ok) return 1; /* "A, B size mismatch" */
  int E = flags & O_RIGHT ? br : bc;
  o_mdata_t Bm = md(B), Am = md(A);
  kern()->dmmat_trmm_blk(&Bm, &Am, alpha, flags, ac, 0, E, nB);
  return 0;
}

/* ---- RankUpdate2Sym, mmult.go:453-476 -> DSymmRank2Blk, cmops.go:244-268 ---- */
int o_RankUpdate2Sym(oview C, oview A, oview B, double alpha, double beta, int flags) {
  if (C.rows != C.cols) return 1; /* "C not a square matrix" */
  int P = flags & O_TRANSA ? A.rows : A.cols;
  o_mdata_t Cm = md(C), Am = md(A), Bm = md(B);
  kern()->dmmat_rank2_blk(&Cm, &Am, &Bm, alpha, beta, flags, P, 0, C.rows, vpLen, nB);
  return 0;
}

/* ---- UpdateTrm, mmult.go:501-520 -> DTrmUpdBlk, cmops.go:273-297 ---- */
int o_UpdateTrm(oview C, oview A, oview B, double alpha, double beta, int flags) {
  if (C.rows != C.cols) return 1; /* "C not a square matrix" */
  int P = flags & O_TRANSA ? A.rows : A.cols;
  o_mdata_t Cm = md(C), Am = md(A), Bm = md(B);
  kern()->dmmat_trmupd_blk(&Cm, &Am, &Bm, alpha, beta, flags, P, 0, C.rows, vpLen, nB);
  return 0;
}

/* ---- SolveTrm, mmult.go:388-422 -> calgo.DSolveBlk, calgo/cmops.go:195-214 ---- */
int o_SolveTrm(oview B, oview A, double alpha, int flags) {
  int ok = 1, empty = 0;
  int br = B.rows, bc = B.cols, ar = A.rows, ac = A.cols;
  switch (flags & (O_LEFT | O_RIGHT)) {
    case O_LEFT:
      empty = br == 0;
      ok = br == ac && ac == ar;
      break;
    case O_RIGHT:
      empty = bc == 0;
      ok = bc == ar && ac == ar;
      break;
  }
  if (empty) return 0;
  if (!ok) return 1; /* "A, B size mismatch" */
  int E = flags & O_RIGHT ? br : bc;
  if (ac == 0 || E <= 0) return 0; /* cmops.go:203 */
  o_mdata_t Bm = md(B), Am = md(A);
  kern()->dmmat_solve_blk(&Bm, &Am, alpha, flags, ac, 0, E, nB);
  return 0;
}

/* ---- RankUpdateSym, mmult.go:426-445 -> DSymmRankBlk cmops.go:219-240 ---- */
int o_RankUpdateSym(oview C, oview A, double alpha, double beta, int flags) {
  if (C.rows != C.cols) return 1; /* "C not a square matrix" */
  int P = flags & O_TRANSA ? A.rows : A.cols;
  int E = C.rows;
  if (P == 0 || E <= 0) return 0; /* cmops.go:227 */
  o_mdata_t Cm = md(C), Am = md(A);
  kern()->dmmat_rank_blk(&Cm, &Am, alpha, beta, flags, P, 0, E, vpLen, nB);
  return 0;
}

/* ---- vector helpers (mvec.go) ---- */
static void MVMult(oview Y, oview A, oview X, double alpha, double beta, int flags) {
  /* mvec.go:28-60 */
  if (A.rows == 0 || A.cols == 0) return;
  o_mvec_t Yv = mv_rows(Y), Xv = mv_rows(X);
  o_mdata_t Am = md(A);
  kern()->dmult_gemv_blocked(&Yv, &Am, &Xv, alpha, beta, flags, 0, numel(X), 0, numel(Y), vpLen,
                             mB);
}

static void MVRankUpdateSym(oview A, oview X, double alpha, int flags) {
  /* mvec.go:101-121 -> DSymmRankMV(..., S=0, L=A.Cols(), NB=0) */
  if (A.rows == 0 || A.cols == 0) return;
  o_mvec_t Xv = mv_rows(X);
  o_mdata_t Am = md(A);
  kern()->dmvec_symv_rank(&Am, &Xv, alpha, flags, 0, A.cols, 0);
}

static void MVSolveTrm(oview X, oview A, int flags) {
  /* mvec.go:203-222 -> DSolveBlkMV(X, A, flags, incX, ldA, N=A.Cols(), nB) */
  if (A.rows == 0 || A.cols == 0) return;
  o_mvec_t Xv = mv_rows(X);
  o_mdata_t Am = md(A);
  kern()->dmvec_solve_blocked(&Xv, &Am, flags, A.cols, nB);
}

static double Dot(oview X, oview Y, double alpha) {
  /* mvec.go:274-297 -> DDot: N <= 0 returns 0 (cmops.go:595-597) */
  int N = numel(X);
  if (N <= 0) return 0.0;
  o_mvec_t Xv = mv_cols(X), Yv = mv_cols(Y);
  return kern()->dvec_dot(&Xv, &Yv, alpha, N);
}

static void InvScale(oview X, double alpha) {
  /* mvec.go:422-437 */
  if (numel(X) == 0) return;
  if (X.rows != 1 && X.cols != 1) return;
  o_mvec_t Xv = mv_cols(X);
  kern()->dvec_invscal(&Xv, alpha, numel(X));
}

static void Swap(oview X, oview Y) {
  /* mvec.go:456-482 */
  if (numel(X) == 0 || numel(Y) == 0) return;
  o_mvec_t Xv = mv_cols(X), Yv = mv_cols(Y);
  kern()->dvec_swap(&Xv, &Yv, numel(X));
}

/* ---- pivots, pivot.go ---- */
static void swapRows(oview A, int src, int dst) { /* pivot.go:27-35 */
  if (src == dst || A.rows == 0) return;
  Swap(sub(A, src, 0, 1, A.cols), sub(A, dst, 0, 1, A.cols));
}

static void applyPivots(oview A, const int64_t *p, int np) { /* pivot.go:55-61 */
  for (int k = 0; k < np; k++)
    if (p[k] > 0) swapRows(A, (int)p[k], k);
}

static void pivotIndex(oview A, int64_t *p) { /* pivot.go:98-107 */
  double max = fabs(A.p[0]);
  for (int k = 1; k < A.rows; k++) {
    double v = fabs(A.p[k]);
    if (v > max) {
      p[0] = k;
      max = v;
    }
  }
}

/* ---- Cholesky, chol.go ---- */
static int unblockedCHOL(oview A, int flags, int nr, int *pidx, double *pval) {
  /* chol.go:17-56: per column a11 = sqrt(a11) (NaN -> panic), a21 /= a11,
   * A22 -= a21 a21^T (LOWER), or the a12 row version (UPPER) */
  for (int k = 0; k < A.rows; k++) {
    oview a11 = sub(A, k, k, 1, 1);
    double aval = sqrt(a11.p[0]);
    if (isnan(aval)) {
      *pidx = nr + k;
      *pval = a11.p[0];
      return 2;
    }
    a11.p[0] = aval;
    if (flags & O_LOWER) {
      oview a21 = sub(A, k + 1, k, A.rows - k - 1, 1);
      InvScale(a21, aval);
      MVRankUpdateSym(sub(A, k + 1, k + 1, A.rows - k - 1, A.cols - k - 1), a21, -1.0, flags);
    } else {
      oview a12 = sub(A, k, k + 1, 1, A.cols - k - 1);
      InvScale(a12, aval);
      MVRankUpdateSym(sub(A, k + 1, k + 1, A.rows - k - 1, A.cols - k - 1), a12, -1.0, flags);
    }
  }
  return 0;
}

static int blockedCHOL(oview A, int flags, int nb, int *pidx, double *pval) {
  /* chol.go:58-94 */
  for (int k = 0; k < A.rows && k < A.cols;) {
    int kb = k + nb > A.cols ? A.cols - k : nb; /* partition.go:252-254 */
    oview A11 = sub(A, k, k, kb, kb);
    int e = unblockedCHOL(A11, flags, k, pidx, pval);
    if (e) return e;
    int m = A.rows - k - kb;
    if (flags & O_LOWER) {
      oview A21 = sub(A, k + kb, k, m, kb);
      o_SolveTrm(A21, A11, 1.0, O_RIGHT | O_LOWER | O_TRANSA);
      o_RankUpdateSym(sub(A, k + kb, k + kb, m, A.cols - k - kb), A21, -1.0, 1.0, O_LOWER);
    } else {
      oview A12 = sub(A, k, k + kb, kb, A.cols - k - kb);
      o_SolveTrm(A12, A11, 1.0, O_UPPER | O_TRANSA);
      o_RankUpdateSym(sub(A, k + kb, k + kb, m, A.cols - k - kb), A12, -1.0, 1.0,
                      O_UPPER | O_TRANSA);
    }
    k += kb;
  }
  return 0;
}

int o_DecomposeCHOL(oview A, int flags, int nb, int *pidx, double *pval) {
  /* chol.go:116-127 */
  if (A.cols != A.rows) return 1; /* "A not a square matrix" */
  if (A.cols < nb || nb == 0) return unblockedCHOL(A, flags, 0, pidx, pval);
  return blockedCHOL(A, flags, nb, pidx, pval);
}

void o_SolveCHOL(oview B, oview A, int flags) {
  /* chol.go:147-158 */
  if (flags & O_UPPER) {
    o_SolveTrm(B, A, 1.0, O_UPPER | O_TRANSA);
    o_SolveTrm(B, A, 1.0, O_UPPER);
  } else if (flags & O_LOWER) {
    o_SolveTrm(B, A, 1.0, O_LOWER);
    o_SolveTrm(B, A, 1.0, O_LOWER | O_TRANSA);
  }
}

/* ---- LU with partial pivoting, lu.go ---- */
static void unblockedLUpiv(oview A, int64_t *p, int np) {
  /* lu.go:103-170, FLAME LU variant 3 (left-looking); p relative to A's top */
  int k = 0;
  for (; k < A.rows && k < A.cols; k++) {
    oview a1 = sub(A, 0, k, A.rows, 1);             /* current column */
    oview A00 = sub(A, 0, 0, k, k);
    oview a01 = sub(A, 0, k, k, 1);
    oview a10 = sub(A, k, 0, 1, k);
    oview a11 = sub(A, k, k, 1, 1);
    oview A20 = sub(A, k + 1, 0, A.rows - k - 1, k);
    oview a21 = sub(A, k + 1, k, A.rows - k - 1, 1);
    applyPivots(a1, p, k < np ? k : np);            /* :131 */
    MVSolveTrm(a01, A00, O_LOWER | O_UNIT);          /* :134 */
    a11.p[0] += Dot(a10, a01, -1.0);                /* :136 */
    MVMult(a21, A20, a01, -1.0, 1.0, 0);            /* :138 */
    oview aB1 = sub(A, k, k, A.rows - k, 1);
    pivotIndex(aB1, &p[k]);                         /* :142 */
    applyPivots(aB1, &p[k], 1);                     /* :145 */
    InvScale(a21, a11.p[0]);                        /* :148 */
    applyPivots(sub(A, k, 0, A.rows - k, k), &p[k], 1); /* :151-152 */
    p[k] += k;                                      /* :154 */
  }
  if (k < A.cols) { /* :165-168 */
    oview ATR = sub(A, 0, k, k, A.cols - k);
    applyPivots(ATR, p, np);
    o_SolveTrm(ATR, sub(A, 0, 0, k, k), 1.0, O_LEFT | O_UNIT | O_LOWER);
  }
}

static void blockedLUpiv(oview A, int64_t *p, int np, int nb) {
  /* lu.go:173-236 */
  int k = 0;
  while (k < A.rows && k < A.cols) {
    int kb = k + nb > A.cols ? A.cols - k : nb;    /* partition.go:252-254 */
    int pkb = k + kb > np ? np - k : kb;           /* partition.go:405-407 */
    oview A1 = sub(A, 0, k, A.rows, kb);
    applyPivots(A1, p, k);                          /* :201 */
    oview A00 = sub(A, 0, 0, k, k);
    oview A01 = sub(A, 0, k, k, kb);
    o_SolveTrm(A01, A00, 1.0, O_LOWER | O_UNIT);    /* :204 */
    o_Mult(sub(A, k, k, kb, kb), sub(A, k, 0, kb, k), A01, -1.0, 1.0, 0);        /* :206 */
    int m2 = A.rows - k - kb;
    o_Mult(sub(A, k + kb, k, m2 > 0 ? m2 : 0, kb), sub(A, k + kb, 0, m2 > 0 ? m2 : 0, k), A01,
           -1.0, 1.0, 0);                           /* :208 */
    oview AB1 = sub(A, k, k, A.rows - k, kb);
    unblockedLUpiv(AB1, &p[k], pkb);                /* :212 */
    applyPivots(sub(A, k, 0, A.rows - k, k), &p[k], pkb); /* :216 */
    for (int i = 0; i < pkb; i++) p[k + i] += k;    /* :218-220 */
    k += kb;
  }
  if (k < A.cols) { /* :231-234 */
    oview ATR = sub(A, 0, k, k, A.cols - k);
    applyPivots(ATR, p, np);
    o_SolveTrm(ATR, sub(A, 0, 0, k, k), 1.0, O_LEFT | O_UNIT | O_LOWER);
  }
}

int o_DecomposeLU(oview A, int64_t *pivots, int npiv, int nb) {
  /* lu.go:257-273 */
  int mlen = A.rows < A.cols ? A.rows : A.cols;
  if (npiv < mlen) return 1; /* "pivot array < min(A.Rows(),A.Cols())" */
  memset(pivots, 0, sizeof(int64_t) * (size_t)npiv);
  if (mlen <= nb || nb == 0)
    unblockedLUpiv(A, pivots, npiv);
  else
    blockedLUpiv(A, pivots, npiv, nb);
  return 0;
}

int o_SolveLU(oview B, oview A, const int64_t *pivots, int npiv, int flags) {
  /* lu.go:320-334, including its TRANSA branch as written (pivots first,
   * then L^T and U^T) */
  applyPivots(B, pivots, npiv);
  if (flags & O_TRANSA) {
    o_SolveTrm(B, A, 1.0, O_LOWER | O_UNIT | O_TRANSA);
    o_SolveTrm(B, A, 1.0, O_UPPER | O_TRANSA);
  } else {
    o_SolveTrm(B, A, 1.0, O_LOWER | O_UNIT);
    o_SolveTrm(B, A, 1.0, O_UPPER);
  }
  return 0;
}

/* ---- rank-1 helpers of the no-pivot factorizations and the inverse ----
 * MVRankUpdate (mvec.go:67-99 -> dmvec_rank) and MVUpdateTrm (mvec.go:163-
 * 192 -> dmvec_trmv_upd) restated as plain loops: A += alpha x y^T on the
 * whole block, or on its LOWER / UPPER triangle. Scale multiplies. */
static double velem(oview X, int k) { /* element k of a row or column vector view */
  return X.rows == 1 ? X.p[(ptrdiff_t)k * X.ld] : X.p[k];
}

static void MVRankUpdate(oview A, oview X, oview Y, double alpha) {
  for (int j = 0; j < A.cols; j++) {
    const double yj = alpha * velem(Y, j);
    for (int i = 0; i < A.rows; i++) A.p[i + (ptrdiff_t)j * A.ld] += velem(X, i) * yj;
  }
}

static void MVUpdateTrm(oview A, oview X, oview Y, double alpha, int flags) {
  for (int j = 0; j < A.cols; j++) {
    const double yj = alpha * velem(Y, j);
    const int i0 = flags & O_UPPER ? 0 : j, i1 = flags & O_UPPER ? j + 1 : A.rows;
    for (int i = i0; i < i1 && i < A.rows; i++) A.p[i + (ptrdiff_t)j * A.ld] += velem(X, i) * yj;
  }
}

static void Scale(oview X, double alpha) { /* matrix.FloatMatrix.Scale on a vector view */
  for (int k = 0; k < numel(X); k++) {
    double *p = X.rows == 1 ? &X.p[(ptrdiff_t)k * X.ld] : &X.p[k];
    *p *= alpha;
  }
}

/* ---- DecomposeLUnoPiv, lu.go:38-96, 291-299 ---- */
static void unblockedLUnoPiv(oview A) {
  /* FLAME LU variant 5 (lu.go:38-61): a21 /= a11; A22 -= a21 a12 */
  for (int k = 0; k < A.rows && k < A.cols; k++) {
    oview a11 = sub(A, k, k, 1, 1);
    oview a21 = sub(A, k + 1, k, A.rows - k - 1, 1);
    oview a12 = sub(A, k, k + 1, 1, A.cols - k - 1);
    InvScale(a21, a11.p[0]);
    MVRankUpdate(sub(A, k + 1, k + 1, A.rows - k - 1, A.cols - k - 1), a21, a12, -1.0);
  }
}

static void blockedLUnoPiv(oview A, int nb) {
  /* lu.go:64-96 */
  for (int k = 0; k < A.rows && k < A.cols;) {
    int kb = k + nb > A.rows ? A.rows - k : nb; /* partition.go:252-254 (rows and cols) */
    if (k + kb > A.cols) kb = A.cols - k;
    oview A11 = sub(A, k, k, kb, kb);
    oview A12 = sub(A, k, k + kb, kb, A.cols - k - kb);
    oview A21 = sub(A, k + kb, k, A.rows - k - kb, kb);
    oview A22 = sub(A, k + kb, k + kb, A.rows - k - kb, A.cols - k - kb);
    unblockedLUnoPiv(A11);
    o_SolveTrm(A12, A11, 1.0, O_LEFT | O_LOWER | O_UNIT);
    o_SolveTrm(A21, A11, 1.0, O_RIGHT | O_UPPER);
    o_Mult(A22, A21, A12, -1.0, 1.0, 0);
    k += kb;
  }
}

int o_DecomposeLUnoPiv(oview A, int nb) {
  int mlen = A.rows < A.cols ? A.rows : A.cols;
  if (mlen <= nb || nb == 0)
    unblockedLUnoPiv(A);
  else
    blockedLUnoPiv(A, nb);
  return 0;
}

/* ---- DecomposeLDLnoPiv / SolveLDLnoPiv, ldlnp.go:23-264 ---- */
static void unblkLowerLDLnoPiv(oview A) {
  /* ldlnp.go:23-61: A22 -= a21 a21^T / a11 (lower), then l21 = a21 / a11 */
  for (int k = 0; k < A.rows; k++) {
    oview a11 = sub(A, k, k, 1, 1);
    oview a21 = sub(A, k + 1, k, A.rows - k - 1, 1);
    MVUpdateTrm(sub(A, k + 1, k + 1, A.rows - k - 1, A.cols - k - 1), a21, a21, -1.0 / a11.p[0],
                O_LOWER);
    InvScale(a21, a11.p[0]);
  }
}

static void unblkUpperLDLnoPiv(oview A) {
  /* ldlnp.go:64-101: from the last column: A00 -= a01 a01^T / a11 (upper),
   * then u01 = a01 / a11 */
  for (int k = A.rows - 1; k >= 0; k--) {
    oview a11 = sub(A, k, k, 1, 1);
    oview a01 = sub(A, 0, k, k, 1);
    MVUpdateTrm(sub(A, 0, 0, k, k), a01, a01, -1.0 / a11.p[0], O_UPPER);
    InvScale(a01, a11.p[0]);
  }
}

/* SolveDiag(X, D, RIGHT) with D the diagonal of A11: column k times 1/d_k
 * (diag.go:96-101, 117-122) */
static void solveDiagRight(oview X, oview A11) {
  for (int k = 0; k < X.cols; k++) Scale(sub(X, 0, k, X.rows, 1), 1.0 / A11.p[(ptrdiff_t)k * (A11.ld + 1)]);
}

static void blkLowerLDLnoPiv(oview A, int nb) {
  /* ldlnp.go:104-145 */
  for (int k = 0; k < A.rows;) {
    int kb = k + nb > A.rows ? A.rows - k : nb;
    oview A11 = sub(A, k, k, kb, kb);
    oview A21 = sub(A, k + kb, k, A.rows - k - kb, kb);
    oview A22 = sub(A, k + kb, k + kb, A.rows - k - kb, A.rows - k - kb);
    unblkLowerLDLnoPiv(A11);
    o_SolveTrm(A21, A11, 1.0, O_LOWER | O_UNIT | O_RIGHT | O_TRANSA);
    solveDiagRight(A21, A11);
    /* wrk = A21 D1 (MultDiag RIGHT, diag.go) */
    double *w = malloc(sizeof(double) * (size_t)(A21.rows > 0 ? A21.rows : 1) * kb);
    oview W = {w, A21.rows, kb, A21.rows > 0 ? A21.rows : 1};
    for (int j = 0; j < kb; j++)
      for (int i = 0; i < A21.rows; i++)
        w[i + (ptrdiff_t)j * W.ld] = A21.p[i + (ptrdiff_t)j * A21.ld] * A11.p[(ptrdiff_t)j * (A11.ld + 1)];
    o_UpdateTrm(A22, A21, W, -1.0, 1.0, O_LOWER | O_TRANSB);
    free(w);
    k += kb;
  }
}

static void blkUpperLDLnoPiv(oview A, int nb) {
  /* ldlnp.go:148-208: blocks from the bottom right */
  for (int e = A.rows; e > 0;) {
    int kb = e - nb < 0 ? e : nb;
    int k = e - kb;
    oview A00 = sub(A, 0, 0, k, k);
    oview A01 = sub(A, 0, k, k, kb);
    oview A11 = sub(A, k, k, kb, kb);
    unblkUpperLDLnoPiv(A11);
    o_SolveTrm(A01, A11, 1.0, O_UPPER | O_UNIT | O_RIGHT | O_TRANSA);
    solveDiagRight(A01, A11);
    double *w = malloc(sizeof(double) * (size_t)(k > 0 ? k : 1) * kb);
    oview W = {w, k, kb, k > 0 ? k : 1};
    for (int j = 0; j < kb; j++)
      for (int i = 0; i < k; i++)
        w[i + (ptrdiff_t)j * W.ld] = A01.p[i + (ptrdiff_t)j * A01.ld] * A11.p[(ptrdiff_t)j * (A11.ld + 1)];
    o_UpdateTrm(A00, A01, W, -1.0, 1.0, O_UPPER | O_TRANSB);
    free(w);
    e = k;
  }
}

int o_DecomposeLDLnoPiv(oview A, int flags, int nb) {
  /* ldlnp.go:211-241 (the workspace W of the blocked form is allocated
   * here; a missing W is the Go layer's error) */
  if (A.cols != A.rows) return 1; /* "A not a square matrix" */
  if (A.cols < nb || nb == 0) {
    if (flags & O_LOWER)
      unblkLowerLDLnoPiv(A);
    else
      unblkUpperLDLnoPiv(A);
  } else {
    if (flags & O_LOWER)
      blkLowerLDLnoPiv(A, nb);
    else
      blkUpperLDLnoPiv(A, nb);
  }
  return 0;
}

void o_SolveLDLnoPiv(oview B, oview A, int flags) {
  /* ldlnp.go:251-264; SolveDiag(B, A, LEFT) with the matrix A: row k times
   * 1/a_kk (diag.go:111-116) */
  int tri = flags & O_UPPER ? O_UPPER : (flags & O_LOWER ? O_LOWER : 0);
  if (!tri) return;
  o_SolveTrm(B, A, 1.0, tri | O_UNIT);
  for (int k = 0; k < B.rows; k++) Scale(sub(B, k, 0, 1, B.cols), 1.0 / A.p[(ptrdiff_t)k * (A.ld + 1)]);
  o_SolveTrm(B, A, 1.0, tri | O_UNIT | O_TRANSA);
}

/* ---- InverseTrm, trinv.go:16-245 ---- */
static void unblockedInverseLower(oview A) { /* trinv.go:16-48 */
  for (int k = 0; k < A.rows; k++) {
    oview a10 = sub(A, k, 0, 1, k), a11 = sub(A, k, k, 1, 1);
    oview A20 = sub(A, k + 1, 0, A.rows - k - 1, k), a21 = sub(A, k + 1, k, A.rows - k - 1, 1);
    double aval = a11.p[0];
    InvScale(a21, -aval);
    MVRankUpdate(A20, a21, a10, 1.0);
    InvScale(a10, aval);
    a11.p[0] = 1.0 / aval;
  }
}

static void unblockedInverseUpper(oview A) { /* trinv.go:51-83 */
  for (int k = 0; k < A.rows; k++) {
    oview a01 = sub(A, 0, k, k, 1), a11 = sub(A, k, k, 1, 1);
    oview a12 = sub(A, k, k + 1, 1, A.cols - k - 1), A02 = sub(A, 0, k + 1, k, A.cols - k - 1);
    double aval = a11.p[0];
    InvScale(a12, -aval);
    MVRankUpdate(A02, a01, a12, 1.0);
    InvScale(a01, aval);
    a11.p[0] = 1.0 / aval;
  }
}

static void unblockedInverseUnitLower(oview A) { /* trinv.go:86-113 */
  for (int k = 0; k < A.rows; k++) {
    oview a10 = sub(A, k, 0, 1, k);
    oview A20 = sub(A, k + 1, 0, A.rows - k - 1, k), a21 = sub(A, k + 1, k, A.rows - k - 1, 1);
    Scale(a21, -1.0);
    MVRankUpdate(A20, a21, a10, 1.0);
  }
}

static void unblockedInverseUnitUpper(oview A) { /* trinv.go:116-143 */
  for (int k = 0; k < A.rows; k++) {
    oview a01 = sub(A, 0, k, k, 1);
    oview a12 = sub(A, k, k + 1, 1, A.cols - k - 1), A02 = sub(A, 0, k + 1, k, A.cols - k - 1);
    Scale(a12, -1.0);
    MVRankUpdate(A02, a01, a12, 1.0);
  }
}

static void blockedInverseLower(oview A, int flags, int nb) { /* trinv.go:146-180 */
  for (int k = 0; k < A.rows;) {
    int kb = k + nb > A.rows ? A.rows - k : nb;
    oview A10 = sub(A, k, 0, kb, k), A11 = sub(A, k, k, kb, kb);
    oview A20 = sub(A, k + kb, 0, A.rows - k - kb, k), A21 = sub(A, k + kb, k, A.rows - k - kb, kb);
    o_SolveTrm(A21, A11, -1.0, flags | O_RIGHT);
    o_Mult(A20, A21, A10, 1.0, 1.0, 0);
    o_SolveTrm(A10, A11, 1.0, flags);
    if (flags & O_UNIT)
      unblockedInverseUnitLower(A11);
    else
      unblockedInverseLower(A11);
    k += kb;
  }
}

static void blockedInverseUpper(oview A, int flags, int nb) { /* trinv.go:183-216 */
  for (int k = 0; k < A.rows;) {
    int kb = k + nb > A.rows ? A.rows - k : nb;
    oview A00 = sub(A, 0, 0, k, k), A01 = sub(A, 0, k, k, kb), A11 = sub(A, k, k, kb, kb);
    o_MultTrm(A01, A00, 1.0, flags);
    o_SolveTrm(A01, A11, -1.0, flags | O_RIGHT);
    if (flags & O_UNIT)
      unblockedInverseUnitUpper(A11);
    else
      unblockedInverseUpper(A11);
    k += kb;
  }
}

int o_InverseTrm(oview A, int flags, int nb) { /* trinv.go:219-245 */
  if (nb == 0 || A.cols < nb) {
    if (flags & O_UNIT) {
      if (flags & O_LOWER)
        unblockedInverseUnitLower(A);
      else
        unblockedInverseUnitUpper(A);
    } else {
      if (flags & O_LOWER)
        unblockedInverseLower(A);
      else
        unblockedInverseUpper(A);
    }
  } else {
    if (flags & O_LOWER)
      blockedInverseLower(A, flags, nb);
    else
      blockedInverseUpper(A, flags, nb);
  }
  return 0;
}

/* ---- DecomposeLDL / SolveLDL, ldl.go (diagonal pivoting) ----
 * FloatMatrix helpers the file uses, restated: Diag (a 1 x n view stepping
 * by ld + 1), CopyTo on vectors (element k to element k, for the shorter
 * of the two), MVUpdateDiag (diag.go:141-166), SolveDiag / MultDiag with a
 * diagonal taken from a square block (diag.go:14-139). */
static oview diagv(oview A) {
  int n = A.rows < A.cols ? A.rows : A.cols;
  oview v = {A.p, 1, n, A.ld + 1};
  if (n == 1) v.cols = 1;
  return v;
}
static double *vref(oview X, int k) {
  return X.rows == 1 ? &X.p[(ptrdiff_t)k * X.ld] : &X.p[k];
}
static void copyToV(oview src, oview dst) {
  int n = numel(src) < numel(dst) ? numel(src) : numel(dst);
  for (int k = 0; k < n; k++) *vref(dst, k) = *vref(src, k);
}
static void copyTo(oview src, oview dst) { /* same shape */
  for (int j = 0; j < src.cols; j++)
    for (int i = 0; i < src.rows; i++) dst.p[i + (ptrdiff_t)j * dst.ld] = src.p[i + (ptrdiff_t)j * src.ld];
}
static int IAMax(oview X) { /* mvec.go:405-420 -> DIAMax (cmops.go:619-634) */
  if (X.rows != 1 && X.cols != 1) return -1;
  int N = numel(X);
  if (N <= 0) return -1;
  o_mvec_t v = mv_cols(X);
  return kern()->dvec_iamax(&v, N);
}
int o_iamax(const double *x, int inc, int n) {
  if (n <= 0) return -1;
  o_mvec_t v = {(double *)x, inc};
  return kern()->dvec_iamax(&v, n);
}
static void MVUpdateDiag(oview d, oview x, oview y, double alpha) {
  for (int k = 0; k < numel(d); k++) *vref(d, k) += velem(x, k) * velem(y, k) * alpha;
}
static void swapCols(oview A, int src, int dst) { /* pivot.go:37-45 */
  if (src == dst || A.rows == 0) return;
  Swap(sub(A, 0, src, A.rows, 1), sub(A, 0, dst, A.rows, 1));
}
static void scalePivots(int64_t *p, int n, int offset) { /* pivot.go:47-53 */
  for (int k = 0; k < n; k++)
    if (p[k] > 0) p[k] += offset;
}
/* applyRowPivots / applyColPivots, pivot.go:63-97 (dir 1 FORWARD, 2 BACKWARD) */
static void applyRowPivotsD(oview A, const int64_t *p, int n, int offset, int dir) {
  if (dir == 1) {
    for (int k = 0; k < n; k++)
      if (p[k] > 0) swapRows(A, (int)p[k] - 1 - offset, k);
  } else {
    for (int k = n - 1; k >= 0; k--)
      if (p[k] > 0) swapRows(A, (int)p[k] - 1 - offset, k);
  }
}
static void applyColPivotsD(oview A, const int64_t *p, int n, int offset, int dir) {
  if (dir == 1) {
    for (int k = 0; k < n; k++)
      if (p[k] > 0) swapCols(A, (int)p[k] - 1 - offset, k);
  } else {
    for (int k = n - 1; k >= 0; k--)
      if (p[k] > 0) swapCols(A, (int)p[k] - 1 - offset, k);
  }
}

/* applyPivotSym, pivot.go:145-228: diagonal interchange of AR[0, 0] with
 * AR[index, index] in a symmetric matrix stored in one triangle. LOWER: AL
 * is the block left of AR (its rows swap too); UPPER: AL is [ATL ATR] and AR
 * the row [a11 a12]. */
static void applyPivotSym(oview AL, oview AR, int index, int flags) {
  if (flags & O_LOWER) {
    Swap(sub(AL, 0, 0, 1, AL.cols), sub(AL, index, 0, 1, AL.cols));
    Swap(sub(AR, 1, 0, index - 1, 1), sub(AR, index, 1, 1, index - 1));
    Swap(sub(AR, index + 1, 0, AR.rows - index - 1, 1),
         sub(AR, index + 1, index, AR.rows - index - 1, 1));
    double p1 = AR.p[0], p3 = AR.p[index + (ptrdiff_t)index * AR.ld];
    AR.p[0] = p3;
    AR.p[index + (ptrdiff_t)index * AR.ld] = p1;
    return;
  }
  if (flags & O_UPPER) {
    int colno = AL.cols - AR.cols;
    Swap(sub(AR, 0, 1, 1, AR.cols - 1), sub(AL, index, colno + 1, 1, AR.cols - 1));
    Swap(sub(AL, index + 1, colno, AL.rows - index - 2, 1),
         sub(AL, index, index + 1, 1, colno - index - 1));
    Swap(sub(AL, 0, index, index, 1), sub(AL, 0, colno, index, 1));
    double p1 = AR.p[0], p3 = AL.p[index + (ptrdiff_t)index * AL.ld];
    AR.p[0] = p3;
    AL.p[index + (ptrdiff_t)index * AL.ld] = p1;
  }
}

static int unblkLowerLDL(oview A, int64_t *p, const char **msg) { /* ldl.go:25-88 */
  for (int k = 0; k < A.rows; k++) {
    oview ABL = sub(A, k, 0, A.rows - k, k), ABR = sub(A, k, k, A.rows - k, A.cols - k);
    int imax = IAMax(diagv(ABR));
    if (imax > 0) {
      applyPivotSym(ABL, ABR, imax, O_LOWER);
      p[k] = imax + k + 1;
    } else {
      p[k] = 0;
    }
    double a11 = ABR.p[0];
    if (a11 == 0.0) {
      *msg = "zero value on diagonal";
      return 1;
    }
    oview a21 = sub(A, k + 1, k, A.rows - k - 1, 1);
    MVUpdateTrm(sub(A, k + 1, k + 1, A.rows - k - 1, A.cols - k - 1), a21, a21, -1.0 / a11, O_LOWER);
    InvScale(a21, a11);
  }
  return 0;
}

/* findAndBuildPivot, ldl.go:90-133. The reference's column views of WR run
 * over all rows of W; the rows past AR's are computed but never read back,
 * so they are left out here (the reference's GEMV reads past the end of AL
 * for them). */
static int findAndBuildPivot(oview AL, oview AR, oview WL, oview WR, int k) {
  oview dg = sub(WR, 0, WR.cols - 1, AR.rows, 1);
  int dmax = IAMax(dg);
  oview wcol = sub(WR, 0, 0, AR.rows, 1);
  if (dmax == 0) {
    copyToV(sub(AR, 0, 0, AR.rows, 1), wcol);
  } else {
    copyToV(sub(AR, dmax, 0, 1, dmax + 1), wcol);
    if (dmax < AR.rows - 1)
      copyToV(sub(AR, dmax, dmax, AR.rows - dmax, 1), sub(WR, dmax, 0, wcol.rows - dmax, 1));
  }
  if (k > 0) MVMult(wcol, AL, sub(WL, dmax, 0, 1, WL.cols), -1.0, 1.0, 0);
  if (dmax > 0) {
    double t0 = WR.p[0];
    WR.p[0] = WR.p[dmax];
    WR.p[dmax] = t0;
    /* dg is WR's last column: the same column as wcol when WR has one */
    t0 = dg.p[0];
    dg.p[0] = dg.p[dmax];
    dg.p[dmax] = t0;
  }
  return dmax;
}

static int unblkBoundedLowerLDL(oview A, oview W, int64_t *p, int ncol) { /* ldl.go:135-216 */
  copyToV(diagv(A), sub(W, 0, W.cols - 1, A.rows, 1));
  int nc = 0;
  while (A.cols - nc > 0 && nc < ncol) {
    oview w10 = sub(W, nc, 0, W.rows - nc, nc), w11 = sub(W, nc, nc, W.rows - nc, W.cols - nc);
    oview ABL = sub(A, nc, 0, A.rows - nc, nc), ABR = sub(A, nc, nc, A.rows - nc, A.cols - nc);
    int dmax = findAndBuildPivot(ABL, ABR, w10, w11, nc);
    if (dmax > 0) {
      applyPivotSym(ABL, ABR, dmax, O_LOWER);
      swapRows(w10, 0, dmax);
      p[nc] = dmax + nc + 1;
    } else {
      p[nc] = 0;
    }
    oview a21 = sub(A, nc + 1, nc, A.rows - nc - 1, 1);
    oview wc = sub(w11, 1, 0, a21.rows, 1);
    copyToV(wc, a21);
    A.p[nc + (ptrdiff_t)nc * A.ld] = w11.p[0];
    InvScale(a21, A.p[nc + (ptrdiff_t)nc * A.ld]);
    if (ncol - nc > 1)
      MVUpdateDiag(sub(w11, 1, w11.cols - 1, a21.rows, 1), wc, wc,
                   -1.0 / A.p[nc + (ptrdiff_t)nc * A.ld]);
    nc++;
  }
  return nc;
}

static int blkLowerLDL(oview A, oview W, int64_t *p, int nb, const char **msg) { /* ldl.go:227-295 */
  int k = 0;
  while (A.cols - k > nb) {
    int nblk = unblkBoundedLowerLDL(sub(A, k, k, A.rows - k, A.cols - k), W, p + k, nb);
    oview A21 = sub(A, k + nblk, k, A.rows - k - nblk, nblk);
    oview A22 = sub(A, k + nblk, k + nblk, A.rows - k - nblk, A.cols - k - nblk);
    o_UpdateTrm(A22, A21, sub(W, nblk, 0, A21.rows, nblk), -1.0, 1.0, O_LOWER | O_TRANSB);
    applyRowPivotsD(sub(A, k, 0, A.rows - k, k), p + k, nblk, 0, 1);
    scalePivots(p + k, nblk, k);
    k += nblk;
  }
  if (A.cols - k > 0) {
    oview ABR = sub(A, k, k, A.rows - k, A.cols - k);
    if (unblkLowerLDL(ABR, p + k, msg)) return 1;
    applyRowPivotsD(sub(A, k, 0, A.rows - k, k), p + k, ABR.rows, 0, 1);
    scalePivots(p + k, ABR.rows, k);
  }
  return 0;
}

static int unblkUpperLDL(oview A, int64_t *p, const char **msg) { /* ldl.go:297-352 */
  for (int k = A.rows - 1; k >= 0; k--) {
    int imax = IAMax(diagv(sub(A, 0, 0, k + 1, k + 1)));
    if (imax < k) {
      applyPivotSym(sub(A, 0, 0, k + 1, A.cols), sub(A, k, k, 1, A.cols - k), imax, O_UPPER);
      p[k] = imax + 1;
    } else {
      p[k] = 0;
    }
    double a11 = A.p[k + (ptrdiff_t)k * A.ld];
    if (a11 == 0.0) {
      *msg = "zero on diagonal.";
      return 1;
    }
    oview a01 = sub(A, 0, k, k, 1);
    MVUpdateTrm(sub(A, 0, 0, k, k), a01, a01, -1.0 / a11, O_UPPER);
    InvScale(a01, a11);
  }
  return 0;
}

static int blkUpperLDL(oview A, oview W, int64_t *p, int nb, const char **msg) { /* ldl.go:356-418 */
  for (int e = A.rows; e > 0;) {
    int kb = e < nb ? e : nb, k = e - kb;
    oview A01 = sub(A, 0, k, k, kb), A11 = sub(A, k, k, kb, kb);
    oview A12 = sub(A, k, k + kb, kb, A.cols - k - kb);
    if (unblkUpperLDL(A11, p + k, msg)) return 1;
    applyColPivotsD(A01, p + k, kb, 0, 2);
    applyRowPivotsD(A12, p + k, kb, 0, 2);
    scalePivots(p + k, kb, e - kb);
    o_SolveTrm(A01, A11, 1.0, O_UPPER | O_UNIT | O_RIGHT | O_TRANSA);
    solveDiagRight(A01, A11);
    if (k > 0) {
      /* W = U01 D1 (MultDiag RIGHT) */
      oview wrk = sub(W, 0, 0, k, kb);
      copyTo(A01, wrk);
      for (int j = 0; j < kb; j++) Scale(sub(wrk, 0, j, k, 1), A11.p[(ptrdiff_t)j * (A11.ld + 1)]);
      o_UpdateTrm(sub(A, 0, 0, k, k), A01, wrk, -1.0, 1.0, O_UPPER | O_TRANSB);
    }
    e = k;
  }
  return 0;
}

int o_DecomposeLDL(oview A, oview W, int64_t *ipiv, int npiv, int flags, int nb,
                   const char **msg) { /* ldl.go:439-484 */
  *msg = 0;
  if (A.cols != A.rows) {
    *msg = "A not a square matrix";
    return 1;
  }
  for (int k = 0; k < npiv; k++) ipiv[k] = 0;
  if (A.cols < nb || nb == 0)
    return flags & O_LOWER ? unblkLowerLDL(A, ipiv, msg) : unblkUpperLDL(A, ipiv, msg);
  return flags & O_LOWER ? blkLowerLDL(A, W, ipiv, nb, msg) : blkUpperLDL(A, W, ipiv, nb, msg);
}

/* SolveDiag(B, A, LEFT) with A square (diag.go:85-139): rows times 1/d_k;
 * a 1 x 1 A is a column vector there, and the column is divided. */
static void solveDiagLeft(oview B, oview A) {
  if (A.cols == 1) {
    for (int k = 0; k < B.cols; k++) B.p[(ptrdiff_t)k * B.ld] /= A.p[0];
    return;
  }
  for (int k = 0; k < B.rows; k++) Scale(sub(B, k, 0, 1, B.cols), 1.0 / A.p[(ptrdiff_t)k * (A.ld + 1)]);
}

void o_SolveLDL(oview B, oview A, const int64_t *ipiv, int npiv, int flags) { /* ldl.go:486-511 */
  if (flags & O_UPPER) {
    applyRowPivotsD(B, ipiv, npiv, 0, 2);
    o_SolveTrm(B, A, 1.0, O_UPPER | O_UNIT);
    solveDiagLeft(B, A);
    o_SolveTrm(B, A, 1.0, O_UPPER | O_UNIT | O_TRANSA);
    applyRowPivotsD(B, ipiv, npiv, 0, 1);
  } else if (flags & O_LOWER) {
    applyRowPivotsD(B, ipiv, npiv, 0, 1);
    o_SolveTrm(B, A, 1.0, O_LOWER | O_UNIT);
    solveDiagLeft(B, A);
    o_SolveTrm(B, A, 1.0, O_LOWER | O_UNIT | O_TRANSA);
    applyRowPivotsD(B, ipiv, npiv, 0, 2);
  }
}

/* ---- DecomposeBK / SolveBK, ldlbk.go, ldlbks.go (Bunch-Kaufman) ----
 * Restated: the unblocked factorizations (ldlbk.go:212-456, the forms of
 * LAPACK's dsytf2) and the LOWER solve. DecomposeBK's blocked branch (the
 * dlasyf-like ldlbk.go:459-1111) is not restated: the oracle runs the
 * unblocked algorithm for it, which makes the same pivot decisions on the
 * same (exactly updated) values. SolveBK's UPPER branch (ldlbks.go:159-301)
 * is not restated: it steps its pivot counter the wrong way and indexes
 * past the pivot array at its second step. */
static const double bkALPHA = 0.6403882032022075; /* (1 + sqrt(17)) / 8 */

static void applyBKPivotSym(oview AR, int srcix, int dstix, int flags) { /* ldlbk.go:48-105 */
  if (flags & O_LOWER) {
    Swap(sub(AR, srcix + 1, srcix, dstix - srcix - 1, 1), sub(AR, dstix, srcix + 1, 1, dstix - srcix - 1));
    Swap(sub(AR, dstix + 1, srcix, AR.rows - dstix - 1, 1), sub(AR, dstix + 1, dstix, AR.rows - dstix - 1, 1));
  } else if (flags & O_UPPER) {
    Swap(sub(AR, dstix + 1, srcix, srcix - dstix - 1, 1), sub(AR, dstix, dstix + 1, 1, srcix - dstix - 1));
    Swap(sub(AR, 0, srcix, dstix, 1), sub(AR, 0, dstix, dstix, 1));
  } else {
    return;
  }
  double p1 = AR.p[srcix + (ptrdiff_t)srcix * AR.ld], p3 = AR.p[dstix + (ptrdiff_t)dstix * AR.ld];
  AR.p[srcix + (ptrdiff_t)srcix * AR.ld] = p3;
  AR.p[dstix + (ptrdiff_t)dstix * AR.ld] = p1;
}

#define AT(M, i, j) ((M).p[(i) + (ptrdiff_t)(j) * (M).ld])

static void findBKPivot(oview A, int flags, int *r_out, int *np_out) { /* ldlbk.go:107-210 */
  *r_out = 0;
  *np_out = 1;
  if (A.rows == 1) return;
  if (flags & O_LOWER) {
    double amax = fabs(AT(A, 0, 0));
    int r = IAMax(sub(A, 1, 0, A.rows - 1, 1)) + 1;
    double rmax = fabs(AT(A, r, 0));
    if (amax >= bkALPHA * rmax) return;
    int q = IAMax(sub(A, r, 0, 1, r));
    double qmax = fabs(AT(A, r, q));
    if (r < A.rows - 1) {
      oview qrow = sub(A, r + 1, r, A.rows - r - 1, 1);
      q = IAMax(qrow);
      double qmax2 = fabs(AT(qrow, q, 0));
      if (qmax2 > qmax) qmax = qmax2;
    }
    if (amax >= bkALPHA * rmax * (rmax / qmax)) return;
    *r_out = r;
    *np_out = fabs(AT(A, r, r)) >= bkALPHA * qmax ? 1 : 2;
    return;
  }
  if (flags & O_UPPER) {
    int lastcol = A.rows - 1;
    double amax = fabs(AT(A, lastcol, lastcol));
    int r = IAMax(sub(A, 0, lastcol, lastcol, 1));
    double rmax = fabs(AT(A, r, lastcol));
    *r_out = -1;
    if (amax >= bkALPHA * rmax) return;
    double qmax = 0.0;
    if (r > 0) {
      int q = IAMax(sub(A, 0, r, r, 1));
      qmax = fabs(AT(A, q, r));
    }
    oview qrow = sub(A, r, r + 1, 1, lastcol - r);
    int q = IAMax(qrow);
    double qmax2 = fabs(AT(qrow, 0, q));
    if (qmax2 > qmax) qmax = qmax2;
    if (amax >= bkALPHA * rmax * (rmax / qmax)) return;
    *r_out = r;
    *np_out = fabs(AT(A, r, r)) >= bkALPHA * qmax ? 1 : 2;
  }
}

/* the inverse of the 2 x 2 pivot block [a b; b d], as ldlbk.go:274-285:
 * inv = scale [d/b -1; -1 a/b], scale = 1 / ((a/b)(d/b) - 1) / b */
static double bk_inverse(oview a11inv, double a, double b, double d) {
  AT(a11inv, 0, 0) = d / b;
  AT(a11inv, 1, 1) = a / b;
  double scale = 1.0 / ((a / b) * (d / b) - 1.0);
  return scale / b;
}

static void unblkDecompBKLower(oview A, oview wrk, int64_t *p) { /* ldlbk.go:212-332 */
  oview a11inv = sub(wrk, 0, wrk.cols - 2, 2, 2);
  AT(a11inv, 1, 0) = -1.0;
  AT(a11inv, 0, 1) = -1.0;
  for (int k = 0; A.cols - k > 0;) {
    oview ABR = sub(A, k, k, A.rows - k, A.cols - k);
    int r, np;
    findBKPivot(ABR, O_LOWER, &r, &np);
    if (r != 0 && r != np - 1) {
      applyBKPivotSym(ABR, np - 1, r, O_LOWER);
      if (np == 2) {
        double t = AT(ABR, 1, 0);
        AT(ABR, 1, 0) = AT(ABR, r, 0);
        AT(ABR, r, 0) = t;
      }
    }
    oview a11 = sub(A, k, k, np, np), a21 = sub(A, k + np, k, A.rows - k - np, np);
    oview A22 = sub(A, k + np, k + np, A.rows - k - np, A.cols - k - np);
    if (np == 1) {
      MVUpdateTrm(A22, a21, a21, -1.0 / AT(a11, 0, 0), O_LOWER);
      InvScale(a21, AT(a11, 0, 0));
      p[k] = r + k + 1;
    } else {
      double scale = bk_inverse(a11inv, AT(a11, 0, 0), AT(a11, 1, 0), AT(a11, 1, 1));
      oview cwrk = sub(wrk, 2, 0, a21.rows, a21.cols);
      copyTo(a21, cwrk);
      o_Mult(a21, cwrk, a11inv, scale, 0.0, 0);
      o_UpdateTrm(A22, a21, cwrk, -1.0, 1.0, O_LOWER | O_TRANSB);
      p[k] = -(r + k + 1);
      p[k + 1] = p[k];
    }
    k += np;
  }
}

static void unblkDecompBKUpper(oview A, oview wrk, int64_t *p) { /* ldlbk.go:335-456 */
  oview a11inv = sub(wrk, 0, wrk.cols - 2, 2, 2);
  AT(a11inv, 1, 0) = -1.0;
  AT(a11inv, 0, 1) = -1.0;
  for (int e = A.rows; e > 0;) {
    oview ATL = sub(A, 0, 0, e, e);
    int nr = e - 1, r, np;
    findBKPivot(ATL, O_UPPER, &r, &np);
    if (r != -1) {
      applyBKPivotSym(ATL, e - np, r, O_UPPER);
      if (np == 2) {
        double t = AT(ATL, nr - 1, nr);
        AT(ATL, nr - 1, nr) = AT(ATL, r, nr);
        AT(ATL, r, nr) = t;
      }
    }
    int k = e - np;
    oview a11 = sub(A, k, k, np, np), a01 = sub(A, 0, k, k, np), A00 = sub(A, 0, 0, k, k);
    if (np == 1) {
      MVUpdateTrm(A00, a01, a01, -1.0 / AT(a11, 0, 0), O_UPPER);
      InvScale(a01, AT(a11, 0, 0));
      p[k] = r == -1 ? e : r + 1;
    } else {
      double scale = bk_inverse(a11inv, AT(a11, 0, 0), AT(a11, 0, 1), AT(a11, 1, 1));
      oview cwrk = sub(wrk, 2, 0, a01.rows, a01.cols);
      copyTo(a01, cwrk);
      o_Mult(a01, cwrk, a11inv, scale, 0.0, 0);
      o_UpdateTrm(A00, a01, cwrk, -1.0, 1.0, O_UPPER | O_TRANSB);
      p[k] = -(r + 1);
      p[k + 1] = p[k];
    }
    e -= np;
  }
}

int o_DecomposeBK(oview A, oview W, int64_t *ipiv, int npiv, int flags, int nb,
                  const char **msg) { /* ldlbk.go:1114-1140 */
  *msg = 0;
  for (int k = 0; k < npiv; k++) ipiv[k] = 0;
  const int blocked = !(A.cols < nb || nb == 0);
  if (blocked ? (W.cols < nb + 1 || W.rows < A.rows) : (W.cols < 2 || W.rows < A.rows)) {
    *msg = "Workspace too small";
    return 1;
  }
  if (flags & O_LOWER)
    unblkDecompBKLower(A, W, ipiv);
  else if (flags & O_UPPER)
    unblkDecompBKUpper(A, W, ipiv);
  return 0;
}

/* unblkSolveBKLower, ldlbks.go:15-157: phase 1 Z = D^-1 L^-1 P B with the
 * interchanges applied step by step, phase 2 X = P^T L^-T Z */
static void solveBKLower(oview B, oview A, const int64_t *p, int phase) {
  const int n = A.rows;
  if (phase == 1) {
    for (int k = 0; k < n;) { /* k = BT.Rows(), nc = k + 1 */
      int np = p[k] < 0 ? 2 : 1;
      oview BB = sub(B, k, 0, B.rows - k, B.cols);
      oview b1 = sub(B, k, 0, np, B.cols), B2 = sub(B, k + np, 0, B.rows - k - np, B.cols);
      oview a11 = sub(A, k, k, np, np), a21 = sub(A, k + np, k, n - k - np, np);
      if (np == 1) {
        if (p[k] != k + 1) swapRows(BB, 0, (int)p[k] - k - 1);
        MVRankUpdate(B2, a21, b1, -1.0);
        InvScale(b1, AT(a11, 0, 0));
      } else {
        if (p[k] != -(k + 1)) swapRows(BB, 1, (int)(-p[k]) - k - 1);
        double b = AT(a11, 1, 0), apb = AT(a11, 0, 0) / b, dpb = AT(a11, 1, 1) / b;
        double scale = (apb * dpb - 1.0) * b;
        o_Mult(B2, a21, b1, -1.0, 1.0, 0);
        for (int j = 0; j < b1.cols; j++) {
          double s0 = AT(b1, 0, j), s1 = AT(b1, 1, j);
          AT(b1, 0, j) = (dpb * s0 - s1) / scale;
          AT(b1, 1, j) = (apb * s1 - s0) / scale;
        }
      }
      k += np;
    }
    return;
  }
  for (int nc = n; nc > 0;) { /* nc: 1-based row of the block's last row */
    int np = p[nc - 1] < 0 ? 2 : 1;
    int k = nc - np; /* first row of the block; BT.Rows() = nc */
    oview b1 = sub(B, k, 0, np, B.cols), B2 = sub(B, nc, 0, B.rows - nc, B.cols);
    oview a21 = sub(A, nc, k, n - nc, np);
    oview Bx = sub(B, k, 0, B.rows - k, B.cols);
    if (np == 1) {
      MVMult(b1, B2, a21, -1.0, 1.0, O_TRANSA);
      if (p[k] != nc) swapRows(Bx, 0, (int)p[k] - nc);
    } else {
      o_Mult(b1, a21, B2, -1.0, 1.0, O_TRANSA);
      if (p[k] != -nc) swapRows(Bx, 1, (int)(-p[k]) - nc + 1);
    }
    nc -= np;
  }
}

void o_SolveBK(oview B, oview A, const int64_t *ipiv, int npiv, int flags) { /* ldlbks.go:303-316 */
  (void)npiv;
  if (flags & O_LOWER) {
    solveBKLower(B, A, ipiv, 1);
    solveBKLower(B, A, ipiv, 2);
  }
}
#undef AT

/* ---- reference Cholesky adjoint (no counterpart in matops) ----
 * Abar = 1/2 (S + S^T), S = L^-T Phi(L^T Lbar) L^-1 (the closed form of the
 * Cholesky derivative, e.g. Murray 2016, arXiv:1602.07527, eq. 10), computed
 * with the oracle's own GEMM/TRSM. */
void o_chol_adjoint(int n, const double *L, int ldl, const double *Lbar, int ldb, double *Abar,
                    int lda) {
  double *Lt = calloc((size_t)n * n, sizeof(double));
  double *Lb = calloc((size_t)n * n, sizeof(double));
  double *P = calloc((size_t)n * n, sizeof(double));
  for (int j = 0; j < n; j++)
    for (int i = j; i < n; i++) {
      Lt[i + (size_t)j * n] = L[i + (size_t)j * ldl];
      Lb[i + (size_t)j * n] = Lbar[i + (size_t)j * ldb];
    }
  oview vL = {Lt, n, n, n}, vB = {Lb, n, n, n}, vP = {P, n, n, n};
  o_Mult(vP, vL, vB, 1.0, 0.0, O_TRANSA); /* P = L^T Lbar */
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < j; i++) P[i + (size_t)j * n] = 0.0;
    P[j + (size_t)j * n] *= 0.5;
  }
  o_SolveTrm(vP, vL, 1.0, O_LEFT | O_LOWER | O_TRANSA); /* L^-T P */
  o_SolveTrm(vP, vL, 1.0, O_RIGHT | O_LOWER);           /* ... L^-1 */
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++)
      Abar[i + (size_t)j * lda] = 0.5 * (P[i + (size_t)j * n] + P[j + (size_t)i * n]);
  free(Lt);
  free(Lb);
  free(P);
}

/* ---- the kernel-level boundary (calgo/cmops.go:68-91, 171-240) ---- */
void o_kern_gemm(double *C, int ldc, const double *A, int lda, const double *B, int ldb,
                 double alpha, double beta, int flags, int P, int S, int L, int R, int E) {
  o_mdata_t Cm = {C, ldc}, Am = {(double *)A, lda}, Bm = {(double *)B, ldb};
  kern()->dmult_mm_blocked3(&Cm, &Am, &Bm, alpha, beta, flags, P, S, L, R, E, vpLen, nB, mB);
}

void o_kern_trsm(double *B, int ldb, const double *A, int lda, double alpha, int flags, int N,
                 int S, int E) {
  o_mdata_t Bm = {B, ldb}, Am = {(double *)A, lda};
  kern()->dmmat_solve_blk(&Bm, &Am, alpha, flags, N, S, E, nB);
}

void o_kern_trsm_unb(double *B, int ldb, const double *A, int lda, double alpha, int flags, int N,
                     int S, int E) {
  o_mdata_t Bm = {B, ldb}, Am = {(double *)A, lda};
  kern()->dmmat_solve_unb(&Bm, &Am, alpha, flags, N, S, E);
}

void o_kern_syrk(double *C, int ldc, const double *A, int lda, double alpha, double beta,
                 int flags, int P, int S, int E) {
  o_mdata_t Cm = {C, ldc}, Am = {(double *)A, lda};
  kern()->dmmat_rank_blk(&Cm, &Am, alpha, beta, flags, P, S, E, vpLen, nB);
}
