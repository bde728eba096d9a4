/*
 * Restatement of the calgo C kernels on the hot path -- TEST INFRASTRUCTURE
 * (the oracle's "port" kernels; the product never links this file).
 *
 * Each function keeps the interface and the semantics of the calgo routine
 * it restates (cited file:line) -- argument meaning, blocking of the outer
 * loops, beta handling, division-based substitution -- with plain scalar
 * loops in place of the SSE micro-kernels. Summation inside a dot product
 * uses two alternating partial sums, like the 2-lane SSE accumulators of
 * calgo/inner_ddot.h:13.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>

#include "oracle.h"

#define MAX_VP_ROWS 196 /* calgo/cmops.h:27 */
#define MAX_VP_COLS 68  /* calgo/cmops.h:28 */

/* dscale_tile, calgo/vec_scale.c:114: beta == 1 skips, beta == 0 writes
 * zeros (clearing NaN/Inf), otherwise multiplies. */
static void p_scale_tile(double *X, int ldX, double f0, int M, int N) {
  if (f0 == 1.0) return;
  for (int j = 0; j < N; j++) {
    double *c = X + (ptrdiff_t)j * ldX;
    for (int i = 0; i < M; i++) c[i] = f0 == 0.0 ? 0.0 : c[i] * f0;
  }
}

/* dscale_vec, calgo/vec_scale.c:94 */
static void p_scale_vec(double *X, int incX, double f0, int N) {
  if (f0 == 1.0) return;
  for (int i = 0; i < N; i++) X[(ptrdiff_t)i * incX] = f0 == 0.0 ? 0.0 : X[(ptrdiff_t)i * incX] * f0;
}

/* element (i,k) of op(A) and (k,j) of op(B) */
static inline double opA(const o_mdata_t *A, int flags, int i, int k) {
  return flags & O_TRANSA ? A->md[k + (ptrdiff_t)i * A->step] : A->md[i + (ptrdiff_t)k * A->step];
}
static inline double opB(const o_mdata_t *B, int flags, int k, int j) {
  return flags & O_TRANSB ? B->md[j + (ptrdiff_t)k * B->step] : B->md[k + (ptrdiff_t)j * B->step];
}

/* dmult_mm_blocked3, calgo/dmmat_gemm.c:364-416: for each NB column block
 * of C[R:E, S:L], scale by beta once (:398-400), then for each vlen-long
 * panel of the inner dimension add alpha * (panel dot product) to every
 * element (the _dblock_ddot_sse kernel, :14, adds alpha*(c0+c1)). */
static void p_dmult_mm_blocked3(o_mdata_t *C, const o_mdata_t *A, const o_mdata_t *B,
                                double alpha, double beta, int flags, int P, int S, int L, int R,
                                int E, int vlen, int NB, int MB) {
  (void)MB;
  if (L - S <= 0 || E - R <= 0) return;
  if (NB > MAX_VP_ROWS || NB <= 0) NB = MAX_VP_ROWS;
  if (vlen > MAX_VP_ROWS || vlen <= 0) vlen = MAX_VP_ROWS;
  for (int j = S; j < L; j += NB) {
    int nJ = L - j < NB ? L - j : NB;
    p_scale_tile(&C->md[(ptrdiff_t)j * C->step + R], C->step, beta, E - R, nJ);
    for (int k = 0; k < P; k += vlen) {
      int nP = P - k < vlen ? P - k : vlen;
      for (int jj = j; jj < j + nJ; jj++) {
        for (int i = R; i < E; i++) {
          double c0 = 0.0, c1 = 0.0;
          int kk = 0;
          for (; kk + 1 < nP; kk += 2) {
            c0 += opA(A, flags, i, k + kk) * opB(B, flags, k + kk, jj);
            c1 += opA(A, flags, i, k + kk + 1) * opB(B, flags, k + kk + 1, jj);
          }
          if (kk < nP) c0 += opA(A, flags, i, k + kk) * opB(B, flags, k + kk, jj);
          C->md[i + (ptrdiff_t)jj * C->step] += alpha * (c0 + c1);
        }
      }
    }
  }
}

/* Element (i,k) of op(A) for the triangular solves: TRANSA transposes. */
static inline double triA(const o_mdata_t *A, int flags, int i, int k) {
  return flags & O_TRANSA ? A->md[k + (ptrdiff_t)i * A->step] : A->md[i + (ptrdiff_t)k * A->step];
}

/* dmmat_solve_blk, calgo/dmmat_solve.c:864-909. Side: RIGHT bit, else
 * LEFT; triangle: UPPER bit, else lower (:879-906); TRANSA transposes; UNIT
 * never reads the diagonal. The blocked variants (:464-861) solve the
 * diagonal blocks by division-based substitution (dmmat_solve_unb, :400,
 * scalar variants :45-397) and update the rest with GEMM, then apply alpha
 * by a final scaling (e.g. :495-499). Restated here as the same column
 * substitution over the whole order N followed by the alpha scaling; the
 * blocking only regroups the same operations. */
static void p_dmmat_solve_blk(o_mdata_t *B, const o_mdata_t *A, double alpha, int flags, int N,
                              int S, int E, int NB) {
  (void)NB;
  if (E - S <= 0) return;
  const int upper = (flags & O_UPPER) != 0;
  const int trans = (flags & O_TRANSA) != 0;
  const int unit = (flags & O_UNIT) != 0;
  const int eff_lower = upper == trans; /* op(A) lower triangular */
  if (!(flags & O_RIGHT)) {
    /* op(A) X = B, columns S:E of B */
    for (int j = S; j < E; j++) {
      double *x = B->md + (ptrdiff_t)j * B->step;
      if (eff_lower) {
        for (int i = 0; i < N; i++) {
          double v = x[i];
          if (!unit) v = v / triA(A, flags, i, i);
          x[i] = v;
          for (int r = i + 1; r < N; r++) x[r] -= triA(A, flags, r, i) * v;
        }
      } else {
        for (int i = N - 1; i >= 0; i--) {
          double v = x[i];
          if (!unit) v = v / triA(A, flags, i, i);
          x[i] = v;
          for (int r = 0; r < i; r++) x[r] -= triA(A, flags, r, i) * v;
        }
      }
    }
    if (alpha != 1.0)
      p_scale_tile(B->md + (ptrdiff_t)S * B->step, B->step, alpha, N, E - S);
  } else {
    /* X op(A) = B, rows S:E of B: column c of X from columns already solved */
    for (int r = S; r < E; r++) {
      double *x = B->md + r;
      const int ld = B->step;
      if (!eff_lower) {
        /* op(A) upper: X[:,c] = (B[:,c] - sum_{k<c} X[:,k] op(A)[k][c]) / op(A)[c][c] */
        for (int c = 0; c < N; c++) {
          double v = x[(ptrdiff_t)c * ld];
          if (!unit) v = v / triA(A, flags, c, c);
          x[(ptrdiff_t)c * ld] = v;
          for (int q = c + 1; q < N; q++) x[(ptrdiff_t)q * ld] -= v * triA(A, flags, c, q);
        }
      } else {
        for (int c = N - 1; c >= 0; c--) {
          double v = x[(ptrdiff_t)c * ld];
          if (!unit) v = v / triA(A, flags, c, c);
          x[(ptrdiff_t)c * ld] = v;
          for (int q = 0; q < c; q++) x[(ptrdiff_t)q * ld] -= v * triA(A, flags, c, q);
        }
      }
    }
    if (alpha != 1.0) p_scale_tile(B->md + S, B->step, alpha, E - S, N);
  }
}

/* one row-panel product used by SYRK: C[i][j] += alpha * dot(row i, row j)
 * of A (or of A^T with TRANSA) over the P inner elements */
static double syrk_dot(const o_mdata_t *A, int trans, int i, int j, int P) {
  double c0 = 0.0, c1 = 0.0;
  int k = 0;
  for (; k + 1 < P; k += 2) {
    if (trans) {
      c0 += A->md[k + (ptrdiff_t)i * A->step] * A->md[k + (ptrdiff_t)j * A->step];
      c1 += A->md[k + 1 + (ptrdiff_t)i * A->step] * A->md[k + 1 + (ptrdiff_t)j * A->step];
    } else {
      c0 += A->md[i + (ptrdiff_t)k * A->step] * A->md[j + (ptrdiff_t)k * A->step];
      c1 += A->md[i + (ptrdiff_t)(k + 1) * A->step] * A->md[j + (ptrdiff_t)(k + 1) * A->step];
    }
  }
  if (k < P)
    c0 += trans ? A->md[k + (ptrdiff_t)i * A->step] * A->md[k + (ptrdiff_t)j * A->step]
                : A->md[i + (ptrdiff_t)k * A->step] * A->md[j + (ptrdiff_t)k * A->step];
  return c0 + c1;
}

/* dmmat_rank_blk, calgo/dmmat_rank.c:62-126: NB-wide diagonal blocks of
 * C[S:E,S:E] are scaled by beta row by row and updated (_dmmat_rank_diag,
 * :18-58); the rectangles left of (LOWER) or right of (UPPER) each diagonal
 * block are updated without beta (:109-124). For LOWER the rectangle runs
 * over columns 0..i and rows S.. of A, as the reference indexes it. */
static void p_dmmat_rank_blk(o_mdata_t *C, const o_mdata_t *A, double alpha, double beta,
                             int flags, int P, int S, int E, int vlen, int NB) {
  (void)vlen;
  if (E - S <= 0 || P <= 0) return;
  if (NB > MAX_VP_COLS || NB <= 0) NB = MAX_VP_COLS;
  if (NB > E - S) NB = E - S;
  const int trans = (flags & O_TRANSA) != 0;
  const int lower = (flags & O_LOWER) != 0;
  for (int i = S; i < E; i += NB) {
    int nI = E - i < NB ? E - i : NB;
    /* diagonal block, row by row (:26-57) */
    for (int r = 0; r < nI; r++) {
      const int row = i + r;
      if (!(flags & O_UPPER)) {
        p_scale_vec(&C->md[row + (ptrdiff_t)i * C->step], C->step, beta, r + 1);
        for (int c = i; c <= row; c++)
          C->md[row + (ptrdiff_t)c * C->step] += alpha * syrk_dot(A, trans, row, c, P);
      } else {
        p_scale_vec(&C->md[row + (ptrdiff_t)row * C->step], C->step, beta, nI - r);
        for (int c = row; c < i + nI; c++)
          C->md[row + (ptrdiff_t)c * C->step] += alpha * syrk_dot(A, trans, row, c, P);
      }
    }
    /* off-diagonal rectangle (:110-124) */
    if (lower) {
      for (int r = i; r < i + nI; r++)
        for (int c = 0; c < i; c++)
          C->md[r + (ptrdiff_t)c * C->step] += alpha * syrk_dot(A, trans, r, S + c, P);
    } else {
      for (int r = i; r < i + nI; r++)
        for (int c = i + nI; c < E; c++)
          C->md[r + (ptrdiff_t)c * C->step] += alpha * syrk_dot(A, trans, r, c, P);
    }
  }
}

/* dmvec_symv_rank, calgo/dmvec_rank.c:239-262: A += alpha*x*x^T on the
 * lower (default) or upper (UPPER bit) triangle of columns S:L, each column
 * j updated as A[i][j] += x[i] * (alpha*x[j]) (_dmvec_vpur_syr_lower :192,
 * cf = alpha*x1[0]). */
static void p_dmvec_symv_rank(o_mdata_t *A, const o_mvec_t *X, double alpha, int flags, int S,
                              int L, int NB) {
  (void)NB;
  if (L - S <= 0) return;
  for (int j = S; j < L; j++) {
    const double cf = alpha * X->md[(ptrdiff_t)j * X->inc];
    if (flags & O_UPPER) {
      for (int i = 0; i <= j; i++)
        A->md[i + (ptrdiff_t)j * A->step] += X->md[(ptrdiff_t)i * X->inc] * cf;
    } else {
      for (int i = j; i < L; i++)
        A->md[i + (ptrdiff_t)j * A->step] += X->md[(ptrdiff_t)i * X->inc] * cf;
    }
  }
}

/* dmult_gemv_blocked, calgo/dmvec_gemv.c:353-432: Y[R:E] = beta*Y +
 * alpha*op(A)[R:E, S:L]*X[S:L]; without TRANSA by AXPY over columns, with
 * TRANSA by dot products. */
static void p_dmult_gemv_blocked(o_mvec_t *Y, const o_mdata_t *A, const o_mvec_t *X,
                                 double alpha, double beta, int flags, int S, int L, int R, int E,
                                 int vlen, int MB) {
  (void)vlen;
  (void)MB;
  if (L - S <= 0 || E - R <= 0) return;
  if (beta != 1.0) p_scale_vec(&Y->md[(ptrdiff_t)R * Y->inc], Y->inc, beta, E - R);
  if (flags & O_TRANSA) {
    for (int i = R; i < E; i++) {
      double c = 0.0;
      for (int k = S; k < L; k++)
        c += A->md[k + (ptrdiff_t)i * A->step] * X->md[(ptrdiff_t)k * X->inc];
      Y->md[(ptrdiff_t)i * Y->inc] += alpha * c;
    }
  } else {
    for (int k = S; k < L; k++) {
      const double xk = alpha * X->md[(ptrdiff_t)k * X->inc];
      for (int i = R; i < E; i++)
        Y->md[(ptrdiff_t)i * Y->inc] += A->md[i + (ptrdiff_t)k * A->step] * xk;
    }
  }
}

/* dmvec_solve_blocked, calgo/dmvec_solve.c:165-219 (LOWER: NB-blocked
 * forward substitution, :186-199 with _dmvec_solve_forward :109; UPPER:
 * backward, :201-218). Division by the diagonal unless UNIT. */
static void p_dmvec_solve_blocked(o_mvec_t *X, const o_mdata_t *A, int flags, int N, int NB) {
  (void)NB;
  const int unit = (flags & O_UNIT) != 0;
  double *x = X->md;
  const ptrdiff_t inc = X->inc;
  if (flags & O_LOWER) {
    for (int i = 0; i < N; i++) {
      double v = x[i * inc];
      if (!unit) v = v / A->md[i + (ptrdiff_t)i * A->step];
      x[i * inc] = v;
      for (int r = i + 1; r < N; r++) x[r * inc] -= A->md[r + (ptrdiff_t)i * A->step] * v;
    }
  } else {
    for (int i = N - 1; i >= 0; i--) {
      double v = x[i * inc];
      if (!unit) v = v / A->md[i + (ptrdiff_t)i * A->step];
      x[i * inc] = v;
      for (int r = 0; r < i; r++) x[r * inc] -= A->md[r + (ptrdiff_t)i * A->step] * v;
    }
  }
}

/* dvec_dot, calgo/dvec_util.c:46-65: two alternating partial sums */
static double p_dvec_dot(const o_mvec_t *X, const o_mvec_t *Y, double alpha, int N) {
  double c0 = 0.0, c1 = 0.0;
  int i = 0;
  for (; i < N - 1; i += 2) {
    c0 += X->md[(ptrdiff_t)i * X->inc] * Y->md[(ptrdiff_t)i * Y->inc];
    c1 += X->md[(ptrdiff_t)(i + 1) * X->inc] * Y->md[(ptrdiff_t)(i + 1) * Y->inc];
  }
  if (i < N) c0 += X->md[(ptrdiff_t)i * X->inc] * Y->md[(ptrdiff_t)i * Y->inc];
  return alpha * (c0 + c1);
}

/* dvec_swap, calgo/dvec_util.c:200-226 */
static void p_dvec_swap(o_mvec_t *X, o_mvec_t *Y, int N) {
  for (int i = 0; i < N; i++) {
    double t = X->md[(ptrdiff_t)i * X->inc];
    X->md[(ptrdiff_t)i * X->inc] = Y->md[(ptrdiff_t)i * Y->inc];
    Y->md[(ptrdiff_t)i * Y->inc] = t;
  }
}

/* dvec_invscal, calgo/dvec_util.c:230-245: X = X / alpha */
static void p_dvec_invscal(o_mvec_t *X, double alpha, int N) {
  for (int i = 0; i < N; i++) X->md[(ptrdiff_t)i * X->inc] /= alpha;
}

/* dmmat_solve_unb, calgo/dmmat_solve.c:400-440: alpha is applied first
 * (dscale_tile of the block, :405-408 and :419-422), then the scalar
 * substitution; the lower triangle is used iff LOWER is set (:409, :423),
 * where the blocked form tests UPPER first. */
static void p_dmmat_solve_unb(o_mdata_t *B, const o_mdata_t *A, double alpha, int flags, int N,
                              int S, int E) {
  if (E - S <= 0) return;
  if (flags & O_RIGHT) {
    if (alpha != 1.0) p_scale_tile(B->md + S, B->step, alpha, E - S, N);
  } else {
    if (alpha != 1.0) p_scale_tile(B->md + (ptrdiff_t)S * B->step, B->step, alpha, N, E - S);
  }
  const int f = (flags & O_LOWER) ? (flags & ~O_UPPER) : ((flags & ~O_LOWER) | O_UPPER);
  p_dmmat_solve_blk(B, A, 1.0, f, N, S, E, 0);
}

/* dmult_symm_blocked2, calgo/dmmat_symm.c:253-390: C[R:E, S:L] = beta C +
 * alpha S B (LEFT) or alpha B S (otherwise), S the symmetric matrix whose
 * UPPER (else lower) triangle A holds (the off-diagonal blocks pick their
 * offsets by UPPER, :316 and :366); beta applied once per MB x NB block of
 * C (dscale_tile, :313 and :363), then the three block products added. */
static inline double symA(const o_mdata_t *A, int upper, int i, int k) {
  const int stored = upper ? i <= k : i >= k;
  return stored ? A->md[i + (ptrdiff_t)k * A->step] : A->md[k + (ptrdiff_t)i * A->step];
}
static void p_dmult_symm_blocked2(o_mdata_t *C, const o_mdata_t *A, const o_mdata_t *B,
                                  double alpha, double beta, int flags, int P, int S, int L, int R,
                                  int E, int vlen, int NB, int MB) {
  (void)vlen;
  (void)NB;
  (void)MB;
  if (L - S <= 0 || E - R <= 0) return;
  const int upper = (flags & O_UPPER) != 0;
  p_scale_tile(&C->md[(ptrdiff_t)S * C->step + R], C->step, beta, E - R, L - S);
  for (int j = S; j < L; j++)
    for (int i = R; i < E; i++) {
      double c0 = 0.0, c1 = 0.0;
      int k = 0;
      for (; k + 1 < P; k += 2) {
        if (flags & O_LEFT) {
          c0 += symA(A, upper, i, k) * B->md[k + (ptrdiff_t)j * B->step];
          c1 += symA(A, upper, i, k + 1) * B->md[k + 1 + (ptrdiff_t)j * B->step];
        } else {
          c0 += B->md[i + (ptrdiff_t)k * B->step] * symA(A, upper, k, j);
          c1 += B->md[i + (ptrdiff_t)(k + 1) * B->step] * symA(A, upper, k + 1, j);
        }
      }
      if (k < P)
        c0 += (flags & O_LEFT) ? symA(A, upper, i, k) * B->md[k + (ptrdiff_t)j * B->step]
                               : B->md[i + (ptrdiff_t)k * B->step] * symA(A, upper, k, j);
      C->md[i + (ptrdiff_t)j * C->step] += alpha * (c0 + c1);
    }
}

/* dmmat_trmm_blk, calgo/dmmat_trid.c:780-827: B = alpha op(A) B on columns
 * S:E (LEFT) or B = alpha B op(A) on rows S:E (RIGHT); triangle UPPER, else
 * lower (:799, :815); UNIT never reads the diagonal. Restated as the
 * product with the triangle, each output from the inputs before the
 * update. */
static void p_dmmat_trmm_blk(o_mdata_t *B, const o_mdata_t *A, double alpha, int flags, int N,
                             int S, int E, int NB) {
  (void)NB;
  if (E - S <= 0 || N <= 0) return;
  const int upper = (flags & O_UPPER) != 0;
  const int unit = (flags & O_UNIT) != 0;
  double *x = malloc(sizeof(double) * (size_t)N);
  for (int v = S; v < E; v++) {
    /* one column (LEFT) or row (RIGHT) of B at a time */
    const ptrdiff_t base = flags & O_RIGHT ? v : (ptrdiff_t)v * B->step;
    const ptrdiff_t inc = flags & O_RIGHT ? B->step : 1;
    for (int i = 0; i < N; i++) x[i] = B->md[base + i * inc];
    for (int i = 0; i < N; i++) {
      double c = 0.0;
      for (int k = 0; k < N; k++) {
        /* LEFT: out[i] = sum_k op(A)[i][k] x[k]; RIGHT: out[i] = sum_k x[k] op(A)[k][i] */
        const int r = flags & O_RIGHT ? k : i, q = flags & O_RIGHT ? i : k;
        const int ar = flags & O_TRANSA ? q : r, ac = flags & O_TRANSA ? r : q;
        if (upper ? ar > ac : ar < ac) continue;
        const double a = (ar == ac && unit) ? 1.0 : A->md[ar + (ptrdiff_t)ac * A->step];
        c += a * x[k];
      }
      B->md[base + i * inc] = alpha * c;
    }
  }
  free(x);
}

/* dmmat_rank2_blk, calgo/dmmat_rank.c:132-239: for each NB block i of
 * [S, E): the diagonal triangle gets beta and the two products
 * (_dmmat_rank_diag, :18-58; without TRANSA beta is applied in both halves,
 * :213 and :226, with TRANSA only in the first, :184), the rectangle left
 * of it (LOWER: columns 0..i, rows S.. of the second operand) or right of
 * it (UPPER) the two products only. B is stepped with A's leading
 * dimension (Bd.step = A->step, :160), as the reference does. */
static double r2dot(const o_mdata_t *X, const o_mdata_t *Y, int trans, int i, int j, int P) {
  double c0 = 0.0, c1 = 0.0;
  int k = 0;
  for (; k + 1 < P; k += 2) {
    if (trans) {
      c0 += X->md[k + (ptrdiff_t)i * X->step] * Y->md[k + (ptrdiff_t)j * Y->step];
      c1 += X->md[k + 1 + (ptrdiff_t)i * X->step] * Y->md[k + 1 + (ptrdiff_t)j * Y->step];
    } else {
      c0 += X->md[i + (ptrdiff_t)k * X->step] * Y->md[j + (ptrdiff_t)k * Y->step];
      c1 += X->md[i + (ptrdiff_t)(k + 1) * X->step] * Y->md[j + (ptrdiff_t)(k + 1) * Y->step];
    }
  }
  if (k < P)
    c0 += trans ? X->md[k + (ptrdiff_t)i * X->step] * Y->md[k + (ptrdiff_t)j * Y->step]
                : X->md[i + (ptrdiff_t)k * X->step] * Y->md[j + (ptrdiff_t)k * Y->step];
  return c0 + c1;
}
static void p_dmmat_rank2_blk(o_mdata_t *C, const o_mdata_t *A, const o_mdata_t *B, double alpha,
                              double beta, int flags, int P, int S, int E, int vlen, int NB) {
  (void)vlen;
  if (E - S <= 0 || P <= 0) return;
  if (NB > MAX_VP_COLS || NB <= 0) NB = MAX_VP_COLS;
  if (NB > E - S) NB = E - S;
  const int trans = (flags & O_TRANSA) != 0;
  const int lower = (flags & O_LOWER) != 0;
  o_mdata_t Bs = {B->md, A->step};
  const double beta2 = trans ? 1.0 : beta;
  for (int half = 0; half < 2; half++) {
    const o_mdata_t *X = half ? &Bs : A, *Y = half ? A : &Bs;
    for (int i = S; i < E; i += NB) {
      const int nI = E - i < NB ? E - i : NB;
      for (int r = 0; r < nI; r++) {
        const int row = i + r;
        const int c0 = !(flags & O_UPPER) ? i : row, c1 = !(flags & O_UPPER) ? row : i + nI - 1;
        for (int c = c0; c <= c1; c++) {
          double *cp = &C->md[row + (ptrdiff_t)c * C->step];
          const double b = half ? beta2 : beta;
          *cp = b == 0.0 ? 0.0 : (b == 1.0 ? *cp : b * *cp);
          *cp += alpha * r2dot(X, Y, trans, row, c, P);
        }
      }
      if (lower) {
        for (int c = 0; c < i; c++)
          for (int row = i; row < i + nI; row++)
            C->md[row + (ptrdiff_t)c * C->step] += alpha * r2dot(X, Y, trans, row, S + c, P);
      } else {
        for (int c = i + nI; c < E; c++)
          for (int row = i; row < i + nI; row++)
            C->md[row + (ptrdiff_t)c * C->step] += alpha * r2dot(X, Y, trans, row, c, P);
      }
    }
  }
}

/* dmmat_trmupd_blk, calgo/dmmat_rank.c:292-360: per NB block i of [S, E),
 * the diagonal triangle is scaled by beta and updated row by row
 * (_dmmat_trmupd_diag, :247-290); for UPPER the rows above it (rows 0..i
 * of C with the rows of op(A) from S, :320-324; the reference's offset
 * &A[S] equals that only without TRANSA or with S = 0), for LOWER the rows
 * below it (:349-354) get alpha op(A) op(B) without beta. */
static void p_dmmat_trmupd_blk(o_mdata_t *C, const o_mdata_t *A, const o_mdata_t *B,
                               double alpha, double beta, int flags, int P, int S, int E, int vlen,
                               int NB) {
  (void)vlen;
  if (E - S <= 0 || P <= 0) return;
  if (NB > MAX_VP_COLS || NB <= 0) NB = MAX_VP_COLS;
  if (NB > E - S) NB = E - S;
  const int upper = (flags & O_UPPER) != 0;
  for (int i = S; i < E; i += NB) {
    const int nI = E - i < NB ? E - i : NB;
    /* diagonal triangle */
    for (int r = 0; r < nI; r++) {
      const int row = i + r;
      const int c0 = upper ? row : i, c1 = upper ? i + nI - 1 : row;
      for (int c = c0; c <= c1; c++) {
        double *cp = &C->md[row + (ptrdiff_t)c * C->step];
        *cp = beta == 0.0 ? 0.0 : (beta == 1.0 ? *cp : beta * *cp);
        double c0s = 0.0, c1s = 0.0;
        int k = 0;
        for (; k + 1 < P; k += 2) {
          c0s += opA(A, flags, row, k) * opB(B, flags, k, c);
          c1s += opA(A, flags, row, k + 1) * opB(B, flags, k + 1, c);
        }
        if (k < P) c0s += opA(A, flags, row, k) * opB(B, flags, k, c);
        *cp += alpha * (c0s + c1s);
      }
    }
    /* off-diagonal rectangle: rows [r0, r1) of C, columns [i, i + nI) */
    const int r0 = upper ? 0 : i + nI, r1 = upper ? i : E;
    for (int c = i; c < i + nI; c++)
      for (int row = r0; row < r1; row++) {
        const int arow = upper ? S + row : row;
        double c0s = 0.0;
        for (int k = 0; k < P; k++) c0s += opA(A, flags, arow, k) * opB(B, flags, k, c);
        C->md[row + (ptrdiff_t)c * C->step] += alpha * c0s;
      }
  }
}

/* dvec_iamax, calgo/dvec_util.c:164-195: elements are taken in pairs; the
 * larger of a pair (the first on a tie) becomes the current maximum if it
 * is strictly above it (starting from 0.0); an odd last element likewise.
 * Index 0 when nothing is above 0.0 (all zeros, NaN). A NaN first element
 * of a pair hides its partner: the pair's candidate is then NaN, which
 * never compares above the maximum. */
static int p_dvec_iamax(const o_mvec_t *X, int N) {
  if (N <= 1) return 0;
  const double *Xc = X->md;
  double max = 0.0;
  int ix = 0, i;
  for (i = 0; i < N - 1; i += 2) {
    double c0 = fabs(Xc[0]);
    const double c1 = fabs(Xc[X->inc]);
    int n = 0;
    if (c1 > c0) {
      n = 1;
      c0 = c1;
    }
    if (c0 > max) {
      ix = i + n;
      max = c0;
    }
    Xc += 2 * (ptrdiff_t)X->inc;
  }
  if (i < N) {
    const double c0 = fabs(Xc[0]);
    ix = c0 > max ? N - 1 : ix;
  }
  return ix;
}

const okern_t o_port_kernels = {
    p_dmult_mm_blocked3, p_dmmat_solve_blk,    p_dmmat_rank_blk, p_dmvec_symv_rank,
    p_dmvec_solve_blocked, p_dmult_gemv_blocked, p_dvec_dot,       p_dvec_swap,
    p_dvec_invscal,      p_dmmat_solve_unb,    p_dmult_symm_blocked2, p_dmmat_trmm_blk,
    p_dmmat_rank2_blk,   p_dmmat_trmupd_blk,   p_dvec_iamax,
};
