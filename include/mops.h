/*
 * mops.h - C ABI of the gfx950 (MI355X) dense FP64 path of matops.
 *
 * Everything is column-major FP64 with a leading dimension ("step") equal to
 * the parent matrix's row count, exactly as calgo's views (calgo/cmops.h:47-53).
 * Flag bits are calgo's (calgo/cmops.h:13-24, calgo/cmops.go:19-31).
 *
 * Two families of entry points:
 *  - host-memory drop-ins: same parameter lists as the calgo C kernels the
 *    Go package binds through cgo (calgo/cmops.go). They copy the touched
 *    blocks to the GPU, compute, copy results back and return when done.
 *  - device-resident twins (mops_d_*): device pointers and a HIP stream
 *    (hipStream_t passed as void*, NULL = the legacy default stream). They
 *    are asynchronous unless stated otherwise.
 *
 * Return value: 0 on success; a negative value -k means argument k (1-based)
 * is invalid; other negative values are runtime failures whose text is in
 * mops_last_error(). Positive values are LAPACK-style info codes documented
 * per routine. Every call is re-entrant: state (stream, workspace, error
 * text) is kept per host thread.
 *
 * Bounded waits. Kernels that wait for other workgroups of the same launch
 * (the k split of GEMMs with few output tiles, the chained triangular solve)
 * give up after a spin limit instead of hanging, and then raise a bit in a
 * sticky status word of the calling thread. The host-memory entry points
 * read it before they return and fail with a runtime error (text in
 * mops_last_error()) when a bit is set; the device-resident twins leave it
 * for mops_d_status(). The factorizations report their own waits through
 * their info / status arguments as well.
 */
#ifndef MOPS_H
#define MOPS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Column-major view, pointer at element (0,0); same layout as calgo's
 * mdata_t (calgo/cmops.h:50-53) so a cgo caller can pass its struct as is. */
typedef struct mops_mdata {
  double *md;
  int step;
} mops_mdata_t;

enum {
  MOPS_NOTRANS = 0,
  MOPS_TRANSA = 0x1,
  MOPS_TRANSB = 0x2,
  MOPS_LOWER = 0x4,
  MOPS_UPPER = 0x8,
  MOPS_LEFT = 0x10,
  MOPS_RIGHT = 0x20,
  MOPS_UNIT = 0x40,
  MOPS_TRANS = 0x80,
  MOPS_NOSCALE = 0x800
};

/* ---- runtime ------------------------------------------------------------ */

/* Text of the calling thread's last failure ("" if none). */
const char *mops_last_error(void);
/* Number of visible GPUs (0 when no GPU / no driver). */
int mops_device_count(void);
/* Frees the calling thread's device workspace and stream. */
void mops_thread_release(void);
/* Synchronizes `stream`, then reads and clears the calling thread's status
 * word: 0 when no internal wait timed out since the last check, else a
 * runtime failure (negative) naming the wait in mops_last_error(); the
 * results of the device-resident calls since the last check are then not
 * valid. No counterpart in the reference (its CPU kernels never wait). */
int mops_d_status(void *stream);

/* Launch timing for measurement (no counterpart in the reference). When
 * enabled, the trailing updates of the factorizations are bracketed by HIP
 * events on their stream; collect sums the elapsed time (ms) and the
 * algorithmic flops of one class: 0 = SYRK trailing update of the Cholesky,
 * 1 = GEMM trailing update of the LU, 2 = LU panel. Enabling again clears. */
void mops_trace_enable(int on);
int mops_trace_collect(int cls, int *count, double *ms, double *flops);

/* ---- GEMM: C = beta*C + alpha*op(A)*op(B) -------------------------------- */

/* Replaces dmult_mm_blocked3 (calgo/cmops.h:133-138, calgo/dmmat_gemm.c:364),
 * bound by calgo.DMult (calgo/cmops.go:68-91) under matops.Mult
 * (mmult.go:198-256). Updates the block C[R:E, S:L] with panel length P;
 * rows R:E of op(A) and columns S:L of op(B) are used. beta == 0 writes
 * zeros first (NaN in C is cleared), beta == 1 leaves C as is. vlen/NB/MB
 * are the CPU blocking hints; they are accepted and ignored. */
int mops_dmult_mm_blocked3(mops_mdata_t *C, const mops_mdata_t *A, const mops_mdata_t *B,
                           double alpha, double beta, int flags, int P, int S, int L, int R,
                           int E, int vlen, int NB, int MB);

/* Device-resident GEMM. C is M x N, op(A) is M x P, op(B) is P x N; flags
 * TRANSA / TRANSB select op(). */
int mops_d_dgemm(int flags, int M, int N, int P, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc, void *stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP stream capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments and results as mops_d_dgemm, but the
 * scratch of the k split (partial tiles and flags) and the status word live
 * in one static device array of 34 MiB per GPU (enough for every shape the
 * split plan makes; a product needing more is rejected with ST_BADARG and a
 * message), so the call allocates nothing. The array is shared by every
 * call on that GPU and keeps the last call's contents: calls that overlap
 * on one GPU race and must be serialized by the caller. A timed-out wait is
 * reported by mops_d_status(). */
int mops_d_dgemm_noalloc(int flags, int M, int N, int P, double alpha, const double *A, int lda,
                         const double *B, int ldb, double beta, double *C, int ldc,
                         void *stream);

/* ---- TRSM: B = alpha*op(A)^-1*B (LEFT) or alpha*B*op(A)^-1 (RIGHT) ------- */

/* Replaces dmmat_solve_blk (calgo/cmops.h:207-209, calgo/dmmat_solve.c:864),
 * bound by calgo.DSolveBlk (calgo/cmops.go:195-214) under matops.SolveTrm
 * (mmult.go:388-422). A is N x N; LEFT (no RIGHT bit) solves the N x (E-S)
 * block of columns S:E of B, RIGHT the (E-S) x N block of rows S:E. UPPER
 * selects the upper triangle (otherwise lower), TRANSA transposes A, UNIT
 * never reads the diagonal. NB is a CPU hint, ignored. */
int mops_dmmat_solve_blk(mops_mdata_t *B, const mops_mdata_t *A, double alpha, int flags, int N,
                         int S, int E, int NB);

/* Replaces dmmat_solve_unb (calgo/cmops.h:202-204, calgo/dmmat_solve.c:400),
 * the unblocked form of the same solve: identical arguments and results
 * except for how the triangle is chosen, which follows dmmat_solve_unb:
 * LOWER selects the lower triangle, otherwise the upper one (the blocked
 * form tests UPPER first). Host memory, synchronous. */
int mops_dmmat_solve_unb(mops_mdata_t *B, const mops_mdata_t *A, double alpha, int flags, int N,
                         int S, int E);

/* Device-resident TRSM: B is N x nrhs (LEFT) or nrhs x N (RIGHT). */
int mops_d_dtrsm(int flags, int N, int nrhs, double alpha, const double *A, int lda, double *B,
                 int ldb, void *stream);

/* ---- SYRK: C = beta*C + alpha*A*A^T (or A^T*A with TRANSA), one triangle - */

/* Replaces dmmat_rank_blk (calgo/cmops.h:212-214, calgo/dmmat_rank.c:62),
 * bound by calgo.DSymmRankBlk (calgo/cmops.go:219-240) under
 * matops.RankUpdateSym (mmult.go:426-445). Updates the LOWER (else UPPER)
 * triangle of the block C[S:E, S:E] with rows S:E of A (columns with
 * TRANSA), inner length P. beta scales the whole triangle (BLAS semantics;
 * the reference scales only its diagonal blocks, a difference that only
 * shows for beta != 1, which no caller on the factorization path uses).
 * With S > 0 the LOWER case also differs: calgo updates C[i, 0:i] from
 * rows S..S+i of A (calgo/dmmat_rank.c:116-123, two index bases that agree
 * only for S = 0, the only value RankUpdateSym passes, mmult.go:436);
 * here the block is the lower triangle of C[S:E, S:E], as for UPPER. */
int mops_dmmat_rank_blk(mops_mdata_t *C, const mops_mdata_t *A, double alpha, double beta,
                        int flags, int P, int S, int E, int vlen, int NB);

/* Device-resident SYRK: C is N x N, A is N x P (P x N with TRANSA). */
int mops_d_dsyrk(int flags, int N, int P, double alpha, const double *A, int lda, double beta,
                 double *C, int ldc, void *stream);

/* ---- the other BLAS-3 kernels -------------------------------------------- */

/* Replaces dmult_symm_blocked2 (calgo/cmops.h:148-152, calgo/dmmat_symm.c:253),
 * bound by calgo.DMultSymm (calgo/cmops.go:95-118) under matops.MultSym
 * (mmult.go:266-326). A holds one triangle (UPPER, else lower) of a P x P
 * symmetric matrix S. LEFT: C[R:E, S:L] = beta C + alpha S[R:E, :] B[:, S:L];
 * otherwise (RIGHT): C[R:E, S:L] = beta C + alpha B[R:E, :] S[:, S:L].
 * Differences from calgo, which only arise off its tested forms (LEFT or
 * RIGHT with one triangle bit, no TRANSB; calgo/tmult_test.go:347-455): the
 * product is the symmetric one whatever other bits are set (calgo's RIGHT
 * branch reads B at inconsistent offsets under TRANSB, and its diagonal
 * blocks take the other triangle when no triangle bit is given). beta == 0
 * writes zeros. Host memory, synchronous. */
int mops_dmult_symm_blocked2(mops_mdata_t *C, const mops_mdata_t *A, const mops_mdata_t *B,
                             double alpha, double beta, int flags, int P, int S, int L, int R,
                             int E, int vlen, int NB, int MB);
/* Device-resident SYMM: C (M x N); A M x M (LEFT) or N x N (RIGHT). */
int mops_d_dsymm(int flags, int M, int N, double alpha, const double *A, int lda, const double *B,
                 int ldb, double beta, double *C, int ldc, void *stream);

/* Replaces dmmat_trmm_blk (calgo/cmops.h:199-200, calgo/dmmat_trid.c:780),
 * bound by calgo.DTrmmBlk (calgo/cmops.go:146-166) under matops.MultTrm
 * (mmult.go:338-372): B = alpha op(A) B on columns S:E of B (LEFT, no RIGHT
 * bit) or B = alpha B op(A) on rows S:E (RIGHT); A N x N triangular (UPPER,
 * else lower), TRANSA transposes, UNIT never reads the diagonal. NB is a
 * CPU hint, ignored. Host memory, synchronous. */
int mops_dmmat_trmm_blk(mops_mdata_t *B, const mops_mdata_t *A, double alpha, int flags, int N,
                        int S, int E, int NB);
/* Device-resident TRMM: B N x nrhs (LEFT) or nrhs x N (RIGHT). */
int mops_d_dtrmm(int flags, int N, int nrhs, double alpha, const double *A, int lda, double *B,
                 int ldb, void *stream);

/* Replaces dmmat_rank2_blk (calgo/cmops.h:216-219, calgo/dmmat_rank.c:132),
 * bound by calgo.DSymmRank2Blk (calgo/cmops.go:244-268) under
 * matops.RankUpdate2Sym (mmult.go:453-476): the LOWER (else UPPER) triangle
 * of C[S:E, S:E] = beta C + alpha (A B^T + B A^T) with rows S:E of A and B,
 * or alpha (A^T B + B^T A) with columns S:E under TRANSA; inner length P.
 * BLAS semantics; calgo differs for beta != 1 (it scales only the diagonal
 * blocks, and twice without TRANSA), for ldB != ldA (it steps B with A's
 * leading dimension) and, for LOWER, with S > 0 (as dmmat_rank_blk).
 * Host memory, synchronous. */
int mops_dmmat_rank2_blk(mops_mdata_t *C, const mops_mdata_t *A, const mops_mdata_t *B,
                         double alpha, double beta, int flags, int P, int S, int E, int vlen,
                         int NB);
/* Device-resident SYR2K: C N x N, A and B N x P (P x N with TRANSA). */
int mops_d_dsyr2k(int flags, int N, int P, double alpha, const double *A, int lda,
                  const double *B, int ldb, double beta, double *C, int ldc, void *stream);

/* Replaces dmmat_trmupd_blk (calgo/cmops.h:222-226, calgo/dmmat_rank.c:292),
 * bound by calgo.DTrmUpdBlk (calgo/cmops.go:273-297) under matops.UpdateTrm
 * (mmult.go:501-520): the UPPER (else lower) triangle of C[S:E, S:E] =
 * beta C + alpha op(A)[S:E, :] op(B)[:, S:E], op per TRANSA / TRANSB, inner
 * length P. BLAS semantics; calgo scales only its diagonal blocks by beta
 * and, for UPPER with S > 0, indexes the rows above them from row 0.
 * Host memory, synchronous. */
int mops_dmmat_trmupd_blk(mops_mdata_t *C, const mops_mdata_t *A, const mops_mdata_t *B,
                          double alpha, double beta, int flags, int P, int S, int E, int vlen,
                          int NB);
/* Device-resident triangular update: C N x N, op(A) N x P, op(B) P x N. */
int mops_d_dtrmupd(int flags, int N, int P, double alpha, const double *A, int lda,
                   const double *B, int ldb, double beta, double *C, int ldc, void *stream);

/* ---- Cholesky ------------------------------------------------------------ */

/* Replaces DecomposeCHOL / blockedCHOL / unblockedCHOL (chol.go:17-127):
 * A = L*L^T (LOWER) or U^T*U (otherwise), in place, outer block nb (0 or
 * nb > N: one block). The other triangle is not modified. Returns 0, or
 * 1 + the 0-based global index of the first diagonal whose square root is
 * NaN, with mops_last_error() = "illegal value at <index>: <value>" (the
 * panic text of chol.go:35). Host memory, synchronous. */
int mops_dpotrf(mops_mdata_t *A, int flags, int N, int nb);

/* Replaces SolveCHOL (chol.go:147-158): B (N x nrhs) = A^-1 B using the
 * factor in A. Host memory, synchronous. */
int mops_dpotrs(mops_mdata_t *B, const mops_mdata_t *A, int flags, int N, int nrhs);

/* Device-resident twins. info: device int, set to -1 when the factorization
 * succeeded, -2 when an internal wait timed out, else the 0-based index of
 * the first NaN square root (may be NULL). */
int mops_d_dpotrf(int flags, int N, int nb, double *A, int lda, int *info, void *stream);
int mops_d_dpotrs(int flags, int N, int nrhs, const double *A, int lda, double *B, int ldb,
                  void *stream);

/* Reverse-mode derivative of DecomposeCHOL(A, LOWER, nb) (no counterpart in
 * the reference; for callers that differentiate through the factorization).
 * L: the lower Cholesky factor (N x N, lower triangle read). Lbar: on entry
 * d loss / dL in its lower triangle (the upper triangle is ignored); on
 * return the symmetric d loss / dA (both triangles written), so that for a
 * symmetric perturbation E, d loss = sum_ij Abar_ij E_ij. Closed form
 * (Murray 2016, eq. 10): Abar = (X + X^T) / 2, X = L^-T Phi(L^T Lbar) L^-1,
 * evaluated in place in blocks of order nb (0 -> 512; rounded up to a
 * multiple of 64, at most 1024): a blocked in-place TRMM and two blocked
 * in-place TRSMs (5 N^3 / 3 flops). Uses L, the Lbar buffer and O(nb N)
 * workspace. Host memory, synchronous. */
int mops_dpotrf_adjoint(const mops_mdata_t *L, mops_mdata_t *Lbar, int N, int nb);
/* Device-resident twin: Abar holds Lbar on entry. */
int mops_d_dpotrf_adjoint(int N, int nb, const double *L, int ldl, double *Abar, int lda,
                          void *stream);

/* ---- LU with partial pivoting -------------------------------------------- */

/* Replaces DecomposeLU / blockedLUpiv / unblockedLUpiv (lu.go:103-273) with
 * pivotIndex / applyPivots (pivot.go:55-61, 98-107): A (M x N) = P^T L U in
 * place, unit L below the diagonal, U on and above it. pivots: npiv int64
 * entries (Go []int); all are cleared, then the first min(M,N) receive the
 * 0-based global pivot rows (p[k] >= k, p[k] == k: no interchange), equal to
 * LAPACK ipiv - 1. Ties between |candidates| go to the lowest row. A zero
 * pivot is not an error (Inf/NaN propagate, as in the reference). Returns
 * -3 with "pivot array < min(A.Rows(),A.Cols())" when npiv < min(M,N).
 * M is limited to 32768 rows (the pivoting leaf keeps one row per thread
 * of at most 64 workgroups of 512): taller matrices are rejected with a
 * message; the reference has no such limit. Host memory, synchronous. */
int mops_dgetrf(mops_mdata_t *A, int64_t *pivots, int npiv, int M, int N, int nb);

/* Replaces SolveLU (lu.go:320-334): B (N x nrhs) = A^-1 B from the factors
 * and pivots of mops_dgetrf; TRANSA solves A^T X = B (correctly; the
 * reference applies the pivots first on that branch, which does not).
 * Host memory, synchronous. */
int mops_dgetrs(mops_mdata_t *B, const mops_mdata_t *A, const int64_t *pivots, int npiv,
                int flags, int N, int nrhs);

/* Device-resident twins (same row limit). ipiv: device int64[min(M,N)]
 * (may be NULL); perm: device int[M] (may be NULL) receiving the row
 * permutation (perm[i] = original row now at i); status: device int (may
 * be NULL), 0 on success, 1 if a panel's grid barrier timed out.
 * mops_d_dgetrs uses perm when given, else rebuilds it from the npiv
 * pivots. */
int mops_d_dgetrf(int M, int N, int nb, double *A, int lda, int64_t *ipiv, int *perm,
                  int *status, void *stream);
int mops_d_dgetrs(int flags, int N, int nrhs, const double *A, int lda, const int64_t *ipiv,
                  int npiv, const int *perm, double *B, int ldb, void *stream);

/* ---- factorizations without pivoting, the triangular inverse ------------- */

/* Replaces DecomposeLUnoPiv / blockedLUnoPiv / unblockedLUnoPiv (lu.go:38-96,
 * 291-299): A (M x N) = L U in place, unit L below the diagonal, no row
 * interchanges; a zero pivot gives Inf/NaN, not an error, as in the
 * reference. nb is accepted and ignored (every blocking gives the same
 * factors to rounding). Host memory, synchronous. */
int mops_dgetrf_nopiv(mops_mdata_t *A, int M, int N, int nb);
int mops_d_dgetrf_nopiv(int M, int N, int nb, double *A, int lda, void *stream);

/* Replaces DecomposeLDLnoPiv and its blocked / unblocked forms
 * (ldlnp.go:23-241): A (N x N) = L D L^T (LOWER) or U D U^T (otherwise) in
 * place, unit L / U in the selected strict triangle, D on the diagonal, the
 * other triangle not written. No workspace argument: the GPU keeps its own
 * (the Go layer's "No workspace for blocking invocation" check stays in the
 * caller). Host memory, synchronous. */
int mops_dsytrf_nopiv(mops_mdata_t *A, int flags, int N, int nb);
int mops_d_dsytrf_nopiv(int flags, int N, int nb, double *A, int lda, void *stream);

/* Replaces SolveLDLnoPiv (ldlnp.go:251-264): B (N x nrhs) = A^-1 B from the
 * factors of mops_dsytrf_nopiv; UPPER selects U D U^T, else LOWER L D L^T;
 * neither bit: B is left as is. Host memory, synchronous. */
int mops_dsytrs_nopiv(mops_mdata_t *B, const mops_mdata_t *A, int flags, int N, int nrhs);
int mops_d_dsytrs_nopiv(int flags, int N, int nrhs, const double *A, int lda, double *B, int ldb,
                        void *stream);

/* ---- symmetric indefinite factorizations with pivoting ------------------- */

/* Replaces DecomposeLDL (ldl.go:439-460) and its blocked / unblocked forms
 * (ldl.go:25-418): A (N x N) = P^T L D L^T P (LOWER) or P^T U D U^T P
 * (otherwise) in place, 1x1 pivots (the largest |diagonal| of the part not
 * yet factored). ipiv: npiv int64 entries (Go []int), all cleared, then for
 * each k: k+1-based row exchanged with row k, 0 = none (ldl.go:427). The
 * blocked UPPER form (nb > 0, N >= nb) pivots inside each diagonal block of
 * order nb, as the reference does. The blocked LOWER form gives the
 * unblocked result (the reference's does with a workspace of nb + 1
 * columns; with nb columns its diagonal store collides with the panel).
 * A zero pivot stops the factorization: returns 1 + its 0-based index with
 * mops_last_error() "zero value on diagonal" (LOWER, ldl.go:66) or "zero on
 * diagonal." (UPPER, ldl.go:335); in the blocked LOWER form only the
 * columns the reference factors unblocked are checked (ldl.go:283). No
 * workspace argument: the GPU keeps its own. Returns -3 when npiv < N.
 * Host memory, synchronous. */
int mops_dsytrf_ldl(mops_mdata_t *A, int64_t *ipiv, int npiv, int flags, int N, int nb);
/* Device-resident twin: ipiv device int64[npiv]; info: device int (may be
 * NULL), -1 on success, else the 0-based index of the zero pivot. The
 * blocked UPPER form synchronizes the stream once per block. */
int mops_d_dsytrf_ldl(int flags, int N, int nb, double *A, int lda, int64_t *ipiv, int npiv,
                      int *info, void *stream);

/* Replaces SolveLDL (ldl.go:486-511): B (N x nrhs) = A^-1 B from the factors
 * and pivots of mops_dsytrf_ldl; UPPER selects U D U^T (checked first), else
 * LOWER L D L^T; neither bit: B is left as is. Pivot entries outside
 * [0, N] are ignored. Host memory, synchronous. */
int mops_dsytrs_ldl(mops_mdata_t *B, const mops_mdata_t *A, const int64_t *ipiv, int npiv,
                    int flags, int N, int nrhs);
int mops_d_dsytrs_ldl(int flags, int N, int nrhs, const double *A, int lda, const int64_t *ipiv,
                      int npiv, double *B, int ldb, void *stream);

/* Replaces DecomposeBK (ldlbk.go:1114-1140) with its unblocked and blocked
 * forms (ldlbk.go:107-1111): A (N x N) = L D L^T (LOWER) or U D U^T (UPPER,
 * else nothing but clearing ipiv), D with 1x1 and 2x2 blocks, Bunch-Kaufman
 * pivoting (alpha = (1 + sqrt(17)) / 8). ipiv in LAPACK's format, 1-based:
 * ipiv[k] > 0 a 1x1 block with rows k and ipiv[k]-1 exchanged (itself when
 * none), ipiv[k] = ipiv[k+1] < 0 a 2x2 block. Every nb gives the unblocked
 * decisions (the reference's blocked form makes the same). Singular blocks
 * give Inf/NaN, not an error, as in the reference. The reference's
 * "Workspace too small" check on its W argument stays in the caller.
 * Returns -3 when npiv < N. Host memory, synchronous. */
int mops_dsytrf_bk(mops_mdata_t *A, int64_t *ipiv, int npiv, int flags, int N, int nb);
int mops_d_dsytrf_bk(int flags, int N, int nb, double *A, int lda, int64_t *ipiv, int npiv,
                     void *stream);

/* Replaces SolveBK (ldlbks.go:303-316): B (N x nrhs) = A^-1 B from the
 * factors and pivots of mops_dsytrf_bk; LOWER (checked first) or UPPER,
 * neither: B is left as is. The UPPER solve follows LAPACK's dsytrs (the
 * reference's UPPER branch, ldlbks.go:159-301, steps its block counter the
 * wrong way). Pivot entries outside [-N, N] are ignored. One workgroup per
 * right-hand side. Host memory, synchronous. */
int mops_dsytrs_bk(mops_mdata_t *B, const mops_mdata_t *A, const int64_t *ipiv, int npiv,
                   int flags, int N, int nrhs);
int mops_d_dsytrs_bk(int flags, int N, int nrhs, const double *A, int lda, const int64_t *ipiv,
                     int npiv, double *B, int ldb, void *stream);

/* Replaces InverseTrm (trinv.go:16-245): the LOWER (else upper) triangle of
 * A (N x N) replaced by that of its inverse; UNIT takes the diagonal as 1
 * and leaves it as stored; the other triangle is not written. nb is
 * accepted and ignored. Host memory, synchronous. */
int mops_dinverse_trm(mops_mdata_t *A, int flags, int N, int nb);
int mops_d_dinverse_trm(int flags, int N, int nb, double *A, int lda, void *stream);

/* ---- Mixed precision LU with iterative refinement ------------------------ */

/* No counterpart in the reference: for callers that want the LU solve of a
 * well-conditioned system faster than the FP64 factorization allows. */
enum { MOPS_PREC_FP64 = 0, MOPS_PREC_FP32 = 1, MOPS_PREC_BF16 = 2 };

typedef struct {
  int iterations;        /* solves with the factors (1 = no refinement step) */
  int converged;         /* 1 if backward_error <= tol */
  double backward_error; /* max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) */
  double first_error;    /* the same after the first solve */
  double factor_ms;      /* device time of the factorization */
  double refine_ms;      /* device time of the solves and residuals */
} mops_ir_report;

/* X (N x nrhs) = A^-1 B: LU with partial pivoting in precision prec
 * (MOPS_PREC_FP32: FP32 trailing updates on the FP32 matrix cores, FP64
 * panels; MOPS_PREC_BF16: the same with the trailing products on BF16
 * operands, U12 scaled by a power of two per panel, FP32 accumulation and
 * storage; MOPS_PREC_FP64: the DecomposeLU path), then FP64 refinement
 * X += solve(B - A X) until backward_error <= tol or maxit solves. A and B
 * are not modified. Returns 0 even when not converged (see rep->converged).
 * Host memory, synchronous. */
int mops_dgesv_ir(const mops_mdata_t *A, const mops_mdata_t *B, mops_mdata_t *X, int N, int nrhs,
                  int prec, double tol, int maxit, mops_ir_report *rep);
/* Device-resident twin (rep is a host pointer; the call synchronizes the
 * stream once per refinement step). */
int mops_d_dgesv_ir(int N, int nrhs, const double *A, int lda, const double *B, int ldb, double *X,
                    int ldx, int prec, double tol, int maxit, mops_ir_report *rep, void *stream);

/* The BF16 product of the mixed LU on its own (no counterpart in the
 * reference; for callers that check it): C (M x N, FP32) += alpha * A * B
 * with A (M x K) and B (K x N) FP32 in memory, rounded to BF16 (A after a
 * power-of-two scaling that brings max |A| into [1/2, 1), undone in the
 * result), products on the BF16 matrix cores with FP32 accumulation.
 * Device pointers, asynchronous on stream. */
int mops_d_gemm_bf16(int M, int N, int K, float alpha, const float *A, int lda, const float *B,
                     int ldb, float *C, int ldc, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MOPS_H */
