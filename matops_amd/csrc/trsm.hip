// Triangular solves with many right-hand sides (TRSM) and the triangular
// inverses they are built from.
//
// Replaces dmmat_solve_blk (calgo/dmmat_solve.c:864-909) and its eight
// blocked variants (:464-861), bound by calgo.DSolveBlk
// (calgo/cmops.go:195-214) under matops.SolveTrm (mmult.go:388-422).
//
// The reference solves NB x NB diagonal blocks by scalar substitution and
// updates the rest with GEMM. Here the diagonal blocks are inverted up front
// (all 64x64 leaves in one launch, then combined level by level with batched
// GEMMs: inv([A11 0; A21 A22]) = [X11 0; -X22*A21*X11  X22]), so that every
// step of the blocked solve is a GEMM on the matrix cores:
//   X_b = op(inv(A_bb)) * B_b,   B_rest -= op(A)_{rest,b} * X_b.
// Solutions are written to a workspace copy and moved back once at the end.
#include "leaf.hpp"
#include "mops_solvers.hpp"

namespace mops {
namespace {

constexpr int LEAF = 64;
constexpr int LEAF_LPC = 4;

// Inverts the 64x64 (or smaller, at the end) diagonal leaves of nblk
// blocks of order bb. Block b of A starts at A + b*sA, its inverse at
// X + b*sX. The leaf also clears its columns of the other triangle inside
// the block, so that X ends up dense with explicit zeros. The triangle is
// staged in LDS (an upper one transposed, inv(U) = inv(U^T)^T) and each
// column of the inverse is solved by LEAF_LPC lanes (leaf.hpp).
__global__ __launch_bounds__(LEAF* LEAF_LPC) void k_trtri_leaf(int bb, int lower, int unit,
                                                               const double* A, long lda, long sA,
                                                               double* X, long ldx, long sX) {
  constexpr int LDL = LEAF + 1;
  constexpr int RPL = LEAF / LEAF_LPC;
  __shared__ double Ls[LEAF * LDL];
  const int o = blockIdx.x * LEAF;
  const int lb = min(LEAF, bb - o);
  const double* Ab = A + blockIdx.y * sA + o + (long)o * lda;
  double* Xblk = X + blockIdx.y * sX;
  double* Xd = Xblk + o + (long)o * ldx;
  const int t = threadIdx.x;
  const int c = t / LEAF_LPC, sub = t % LEAF_LPC;
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int r = sub + LEAF_LPC * q;
    double v = 0.0;
    if (r < lb && c < lb && r >= c && !(unit && r == c))
      v = lower ? Ab[r + (long)c * lda] : Ab[c + (long)r * lda];
    Ls[r * LDL + c] = v;
  }
  __syncthreads();
  double x[RPL];
  leaf::inv_lower_column<LEAF, LEAF_LPC, LDL>(Ls, lb, c, sub, unit != 0, x);
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int r = sub + LEAF_LPC * q;
    if (r < lb && c < lb) {
      if (lower)
        Xd[r + (long)c * ldx] = x[q];
      else
        Xd[c + (long)r * ldx] = x[q];
    }
  }
  // explicit zeros in this leaf's columns on the other side of the diagonal
  if (lower) {
    for (int cc = 0; cc < lb; ++cc)
      for (int i = t; i < o; i += LEAF * LEAF_LPC) Xblk[i + (long)(o + cc) * ldx] = 0.0;
  } else {
    for (int cc = 0; cc < lb; ++cc)
      for (int i = o + lb + t; i < bb; i += LEAF * LEAF_LPC) Xblk[i + (long)(o + cc) * ldx] = 0.0;
  }
}

// All nblk blocks (order bb each, same strides) are inverted together.
int trtri_core(hipStream_t s, bool lower, bool unit, int bb, int nblk, const double* A, long lda,
               long sA, double* X, long ldx, long sX, double* T) {
  if (bb <= 0 || nblk <= 0) return ST_OK;
  if (nblk > 65535) {
    set_error("trtri: too many blocks");
    return ST_BADARG;
  }
  hipLaunchKernelGGL(k_trtri_leaf, dim3(ceil_div(bb, LEAF), nblk), dim3(LEAF * LEAF_LPC), 0, s, bb,
                     lower ? 1 : 0, unit ? 1 : 0, A, lda, sA, X, ldx, sX);
  MOPS_CHECK_HIP(hipGetLastError());
  // Level h joins the inverses of [o, o+h) and [o+h, min(o+2h, bb)).
  for (int h = LEAF; h < bb; h *= 2) {
    const int full = bb / (2 * h);  // pairs with both halves complete
    const int rem_o = full * 2 * h;
    const int rem_r = bb - rem_o - h;  // right half of a trailing partial pair
    const long tstride = (long)h * h;  // per pair in T
    for (int pass = 0; pass < 2; ++pass) {
      int npair, r;
      long o;
      if (pass == 0) {
        npair = full;
        r = h;
        o = 0;
      } else {
        if (rem_r <= 0) break;
        npair = 1;
        r = rem_r;
        o = rem_o;
      }
      if (npair == 0) continue;
      const long dA = 2L * h * (lda + 1), dX = 2L * h * (ldx + 1);
      const long tb = (long)nblk * npair;
      (void)tb;
      if (lower) {
        // T = A21 * X11 ; X21 = -X22 * T
        const double* A21 = A + (o + h) + o * lda;
        const double* X11 = X + o + o * ldx;
        const double* X22 = X + (o + h) + (o + h) * ldx;
        double* X21 = X + (o + h) + o * ldx;
        MOPS_TRY(gemm_batched2_d(s, false, false, r, h, h, 1.0, A21, lda, sA, dA, X11, ldx, sX,
                                 dX, 0.0, T, r, tstride * npair, tstride, nblk, npair,
                                 TRI_B_LOWER));
        MOPS_TRY(gemm_batched2_d(s, false, false, r, h, r, -1.0, X22, ldx, sX, dX, T, r,
                                 tstride * npair, tstride, 0.0, X21, ldx, sX, dX, nblk, npair,
                                 TRI_A_LOWER));
      } else {
        // T = A12 * X22 ; X12 = -X11 * T
        const double* A12 = A + o + (o + h) * lda;
        const double* X22 = X + (o + h) + (o + h) * ldx;
        const double* X11 = X + o + o * ldx;
        double* X12 = X + o + (o + h) * ldx;
        MOPS_TRY(gemm_batched2_d(s, false, false, h, r, r, 1.0, A12, lda, sA, dA, X22, ldx, sX,
                                 dX, 0.0, T, h, tstride * npair, tstride, nblk, npair,
                                 TRI_B_UPPER));
        MOPS_TRY(gemm_batched2_d(s, false, false, h, r, h, -1.0, X11, ldx, sX, dX, T, h,
                                 tstride * npair, tstride, 0.0, X12, ldx, sX, dX, nblk, npair,
                                 TRI_A_UPPER));
      }
    }
  }
  return ST_OK;
}

size_t trtri_core_ws(int bb, int nblk) {
  // at level h: pairs*h*h <= bb*h/2 <= bb*bb/2 per block
  return (size_t)nblk * ((size_t)bb * bb / 2 + LEAF * LEAF) * sizeof(double);
}

inline int trsm_block(int N) {
  if (N <= 256) return N;
  return 512;
}

}  // namespace

size_t trtri_ws_bytes(int n) { return trtri_core_ws(n, 1) + 256; }

int trtri_d(hipStream_t s, bool lower, bool unit, int n, const double* A, long lda, double* X,
            long ldx) {
  if (n <= 0) return ST_OK;
  double* T = (double*)ws_carve(trtri_core_ws(n, 1));
  if (!T) {
    set_error("trtri: workspace too small");
    return ST_NOMEM;
  }
  return trtri_core(s, lower, unit, n, 1, A, lda, 0, X, ldx, 0, T);
}

int trtri_blocks_d(hipStream_t s, bool lower, bool unit, int n, int bs, const double* A,
                   long lda, double* X) {
  if (n <= 0) return ST_OK;
  const int nfull = n / bs, rem = n - nfull * bs;
  double* T = (double*)ws_carve(trtri_core_ws(bs, nfull > 0 ? nfull : 1));
  if (!T) {
    set_error("trtri_blocks: workspace too small");
    return ST_NOMEM;
  }
  if (nfull > 0)
    MOPS_TRY(trtri_core(s, lower, unit, bs, nfull, A, lda, (long)bs * (lda + 1), X, bs,
                        (long)bs * bs, T));
  if (rem > 0)
    MOPS_TRY(trtri_core(s, lower, unit, rem, 1, A + (long)nfull * bs * (lda + 1), lda, 0,
                        X + (long)nfull * bs * bs, bs, 0, T));
  return ST_OK;
}

size_t trsm_ws_bytes(int flags, int N, int nrhs) {
  (void)flags;
  if (N <= 0 || nrhs <= 0) return 0;
  const int bs = trsm_block(N);
  const int nblk = ceil_div(N, bs);
  size_t inv = (size_t)nblk * bs * bs * sizeof(double);
  size_t tri = trtri_core_ws(bs, nblk);
  size_t w = (size_t)N * nrhs * sizeof(double);
  return inv + tri + w + 4 * 256;
}

int trsm_d(hipStream_t s, int flags, int N, int nrhs, double alpha, const double* A, long lda,
           double* B, long ldb) {
  if (N <= 0 || nrhs <= 0) return ST_OK;
  const bool right = flags & F_RIGHT;
  const bool upper = flags & F_UPPER;  // dmmat_solve_blk tests UPPER first (:879-906)
  const bool trans = flags & F_TRANSA;
  const bool unit = flags & F_UNIT;
  const bool eff_lower = upper == trans;  // op(A) is lower triangular
  if (alpha != 1.0) MOPS_TRY(scale_tile_d(s, right ? nrhs : N, right ? N : nrhs, alpha, B, ldb));
  if (alpha == 0.0) return ST_OK;

  const int bs = trsm_block(N);
  const int nblk = ceil_div(N, bs);
  double* Dinv = (double*)ws_carve((size_t)nblk * bs * bs * sizeof(double));
  const int wrows = right ? nrhs : N;
  double* W = (double*)ws_carve((size_t)N * nrhs * sizeof(double));
  if (!Dinv || !W) {
    set_error("trsm: workspace too small");
    return ST_NOMEM;
  }
  MOPS_TRY(trtri_blocks_d(s, !upper, unit, N, bs, A, lda, Dinv));

  // op(A)[r0:r0+nr, c0:c0+nc] as a GEMM operand
  auto opA = [&](int r0, int c0, bool& t) -> const double* {
    t = trans;
    return trans ? A + c0 + (long)r0 * lda : A + r0 + (long)c0 * lda;
  };
  const int tri_left = eff_lower ? TRI_A_LOWER : TRI_A_UPPER;
  const int tri_right = eff_lower ? TRI_B_LOWER : TRI_B_UPPER;
  const bool forward = right ? !eff_lower : eff_lower;

  for (int step = 0; step < nblk; ++step) {
    const int b = forward ? step : nblk - 1 - step;
    const int r0 = b * bs;
    const int rb = min(bs, N - r0);
    const double* D = Dinv + (long)b * bs * bs;
    bool t;
    if (!right) {
      double* Bb = B + r0;
      double* Wb = W + r0;
      MOPS_TRY(gemm_d(s, trans, false, rb, nrhs, rb, 1.0, D, bs, Bb, ldb, 0.0, Wb, wrows, GM_FULL,
                      tri_left));
      if (forward && r0 + rb < N) {
        const double* Ap = opA(r0 + rb, r0, t);
        MOPS_TRY(gemm_d(s, t, false, N - r0 - rb, nrhs, rb, -1.0, Ap, lda, Wb, wrows, 1.0,
                        B + r0 + rb, ldb));
      } else if (!forward && r0 > 0) {
        const double* Ap = opA(0, r0, t);
        MOPS_TRY(gemm_d(s, t, false, r0, nrhs, rb, -1.0, Ap, lda, Wb, wrows, 1.0, B, ldb));
      }
    } else {
      double* Bb = B + (long)r0 * ldb;
      double* Wb = W + (long)r0 * wrows;
      MOPS_TRY(gemm_d(s, false, trans, nrhs, rb, rb, 1.0, Bb, ldb, D, bs, 0.0, Wb, wrows,
                      GM_FULL, tri_right));
      if (forward && r0 + rb < N) {
        const double* Ap = opA(r0, r0 + rb, t);
        MOPS_TRY(gemm_d(s, false, t, nrhs, N - r0 - rb, rb, -1.0, Wb, wrows, Ap, lda, 1.0,
                        B + (long)(r0 + rb) * ldb, ldb));
      } else if (!forward && r0 > 0) {
        const double* Ap = opA(r0, 0, t);
        MOPS_TRY(gemm_d(s, false, t, nrhs, r0, rb, -1.0, Wb, wrows, Ap, lda, 1.0, B, ldb));
      }
    }
  }
  return copy_tile_d(s, wrows, right ? N : nrhs, W, wrows, B, ldb);
}

}  // namespace mops
