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

constexpr int LEAF = leaf::NL;

// Inverts the 64x64 (or smaller, at the end) diagonal leaves of nblk
// blocks of order bb. Block b of A starts at A + b*sA, its inverse at
// X + b*sX. The leaf also clears its columns of the other triangle inside
// the block, so that X ends up dense with explicit zeros. The triangle is
// staged in LDS (an upper one transposed, inv(U) = inv(U^T)^T) and each
// column of the inverse is solved by 4 lanes (leaf.hpp).
__global__ __launch_bounds__(256) void k_trtri_leaf(int bb, int lower, int unit,
                                                    const double* A, long lda, long sA,
                                                    double* X, long ldx, long sX) {
  using namespace leaf;
  __shared__ double Ls[NL * LDL];
  __shared__ double dinv[NL];
  const int o = blockIdx.x * LEAF;
  const int lb = min(LEAF, bb - o);
  const double* Ab = A + blockIdx.y * sA + o + (long)o * lda;
  double* Xblk = X + blockIdx.y * sX;
  double* Xd = Xblk + o + (long)o * ldx;
  const int t = threadIdx.x;
  const int c = t / LPC, sub = t % LPC;
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int r = sub + LPC * q;
    double v = 0.0;
    if (r < lb && c < lb && r >= c && !(unit && r == c))
      v = lower ? Ab[r + (long)c * lda] : Ab[c + (long)r * lda];
    Ls[r * LDL + c] = v;
  }
  __syncthreads();
  if (t < NL) dinv[t] = (t < lb && !unit) ? 1.0 / Ls[t * LDL + t] : 1.0;
  __syncthreads();
  double x[RPL];
  inv_lower_column(Ls, dinv, lb, c, sub, unit != 0, x);
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int r = sub + LPC * q;
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
      for (int i = t; i < o; i += 256) Xblk[i + (long)(o + cc) * ldx] = 0.0;
  } else {
    for (int cc = 0; cc < lb; ++cc)
      for (int i = o + lb + t; i < bb; i += 256) Xblk[i + (long)(o + cc) * ldx] = 0.0;
  }
}

// ---- LEFT solve for few right-hand sides: one kernel, blocks chained by flags
//
// Workgroup (b, g), 1024 threads, solves block row b (64 rows) of
// op(A) X = B for the column group g (64 columns); each of its 16 waves
// owns one 16x16 tile of the block. With p the block solved just before b
// (b-1 forward, b+1 backward) and D_b = op(inv(A_bb)):
//   off the chain:  acc = B_b - sum_{c solved before p} op(A)_bc X_c
//                   (each X_c as soon as its flag is up), y = D_b acc,
//                   M_b = D_b op(A)_bp
//   on the chain:   X_b = y - M_b X_p   (one 64x64x64 product)
// X_b is written over B_b with agent-scope atomic stores, then flag (b, g)
// goes up; readers load X with agent-scope atomic loads, so no cache
// maintenance is needed. Workgroups are numbered in solve order, so every
// block a workgroup waits for belongs to a workgroup dispatched before it.
// The products run on the FP64 matrix cores (v_mfma_f64_16x16x4).
constexpr int SK_BS = 64, SK_CG = 64;
constexpr int SK_LD = SK_BS + 1;
constexpr unsigned SK_SPIN_LIMIT = 1u << 24;

__device__ __forceinline__ void sk_wait(const int* flag, int* err) {
  unsigned spins = 0;
  while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
    if (++spins > SK_SPIN_LIMIT) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

__global__ __launch_bounds__(1024) void k_trsm_skinny(int N, int nrhs, int eff_lower, int trans,
                                                      const double* __restrict__ A, long lda,
                                                      const double* __restrict__ Dinv,
                                                      double* B, long ldb, int* flags, int* err) {
  __shared__ double As[SK_BS * SK_LD];  // op(A) block [row][k]
  __shared__ double Xs[SK_BS * SK_LD];  // X block [k][col], or acc
  __shared__ double Ds[SK_BS * SK_LD];  // D_b [row][k]
  __shared__ double Ms[SK_BS * SK_LD];  // M_b [row][k]
  const int nblk = (N + SK_BS - 1) / SK_BS;
  const int ngrp = (nrhs + SK_CG - 1) / SK_CG;
  const int step = blockIdx.x / ngrp, g = blockIdx.x % ngrp;
  const int b = eff_lower ? step : nblk - 1 - step;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int li = lane & 15, lq = lane >> 4;
  const int ti = w & 3, tj = w >> 2;  // this wave's tile: rows 16ti.., columns 16tj..
  const int r0 = b * SK_BS, c0 = g * SK_CG;
  const int nb_rows = min(SK_BS, N - r0);
  const int ncols = min(SK_CG, nrhs - c0);

  // op(A) block (b, c) into S[i][k] (rows of block b, columns of block c)
  auto load_opA = [&](double* S, int c) {
    const int nk = min(SK_BS, N - c * SK_BS);
    for (int idx = t; idx < SK_BS * SK_BS; idx += 1024) {
      const int x = idx % SK_BS, y = idx / SK_BS;  // x fastest: contiguous in memory
      if (!trans) {  // op(A)[i = x][k = y] = A[r0 + x][c*64 + y]
        S[x * SK_LD + y] = (x < nb_rows && y < nk)
                               ? A[(long)r0 + x + ((long)c * SK_BS + y) * lda] : 0.0;
      } else {       // op(A)[i = y][k = x] = A[c*64 + x][r0 + y]
        S[y * SK_LD + x] = (y < nb_rows && x < nk)
                               ? A[(long)c * SK_BS + x + ((long)r0 + y) * lda] : 0.0;
      }
    }
  };
  // X block c (rows of block c, this group's columns) into Xs[k][col]
  // (all loads in flight before the first LDS write: one memory latency)
  auto load_x = [&](int c) {
    const int nk = min(SK_BS, N - c * SK_BS);
    constexpr int XPER = SK_BS * SK_CG / 1024;
    double v[XPER];
#pragma unroll
    for (int q = 0; q < XPER; ++q) {
      const int idx = t + q * 1024;
      const int k = idx % SK_BS, col = idx / SK_BS;
      v[q] = (k < nk && col < ncols)
                 ? __hip_atomic_load(&B[(long)c * SK_BS + k + (long)(c0 + col) * ldb],
                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                 : 0.0;
    }
#pragma unroll
    for (int q = 0; q < XPER; ++q) {
      const int idx = t + q * 1024;
      const int k = idx % SK_BS, col = idx / SK_BS;
      Xs[k * SK_LD + col] = v[q];
    }
  };
  // tile (ti, tj) of S (64 x 64, [row][k]) times Xs ([k][col])
  auto mm = [&](const double* S) {
    v4d d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s4 = 0; s4 < SK_BS; s4 += 4)
      d = __builtin_amdgcn_mfma_f64_16x16x4f64(S[(16 * ti + li) * SK_LD + s4 + lq],
                                               Xs[(s4 + lq) * SK_LD + 16 * tj + li], d, 0, 0, 0);
    return d;
  };

  // acc = B_b tile in the MFMA result layout: reg r <-> row 16ti + lq + 4r, col 16tj + li
  v4d acc;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = 16 * ti + lq + 4 * r, col = 16 * tj + li;
    acc[r] = (row < nb_rows && col < ncols) ? B[r0 + row + (long)(c0 + col) * ldb] : 0.0;
  }
  const int pred = eff_lower ? b - 1 : b + 1;  // the block solved just before b
  const bool has_pred = step > 0;
  // D_b (the inverse of A_bb, transposed for op = transpose) and, through
  // Xs, M_b = D_b op(A)_{b,pred}: neither depends on any solution, so both
  // are ready before the first wait
  {
    const double* D = Dinv + (long)b * SK_BS * SK_BS;
    for (int idx = t; idx < SK_BS * SK_BS; idx += 1024) {
      const int x = idx % SK_BS, y = idx / SK_BS;
      const double v = (x < nb_rows && y < nb_rows) ? D[x + (long)y * SK_BS] : 0.0;
      if (!trans)
        Ds[x * SK_LD + y] = v;
      else
        Ds[y * SK_LD + x] = v;
    }
  }
  if (has_pred) {
    load_opA(Xs, pred);  // Xs[k][col] = op(A)_bp[k][col] (rows of b, columns of pred)
    __syncthreads();
    const v4d m = mm(Ds);
#pragma unroll
    for (int r = 0; r < 4; ++r) Ms[(16 * ti + lq + 4 * r) * SK_LD + 16 * tj + li] = m[r];
  }
  // off the chain: every block solved before pred; the next op(A) block is
  // fetched into registers while the current one is used
  constexpr int PER = SK_BS * SK_BS / 1024;
  double pre[PER];
  auto fetch_opA = [&](int c) {
    const int nk = min(SK_BS, N - c * SK_BS);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int idx = t + q * 1024;
      const int x = idx % SK_BS, y = idx / SK_BS;
      pre[q] = !trans ? ((x < nb_rows && y < nk) ? A[(long)r0 + x + ((long)c * SK_BS + y) * lda] : 0.0)
                      : ((y < nb_rows && x < nk) ? A[(long)c * SK_BS + x + ((long)r0 + y) * lda] : 0.0);
    }
  };
  auto store_opA = [&]() {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int idx = t + q * 1024;
      const int x = idx % SK_BS, y = idx / SK_BS;
      if (!trans)
        As[x * SK_LD + y] = pre[q];
      else
        As[y * SK_LD + x] = pre[q];
    }
  };
  const int noff = step - 1;  // blocks solved before pred
  if (noff > 0) fetch_opA(eff_lower ? 0 : nblk - 1);
  for (int it = 0; it < noff; ++it) {
    const int c = eff_lower ? it : nblk - 1 - it;
    __syncthreads();  // As and Xs are free
    store_opA();
    if (it + 1 < noff) fetch_opA(eff_lower ? it + 1 : nblk - 2 - it);
    if (t == 0) sk_wait(&flags[c * ngrp + g], err);
    __syncthreads();
    load_x(c);
    __syncthreads();
    const v4d d = mm(As);
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] -= d[r];
  }
  // y = D_b acc
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) Xs[(16 * ti + lq + 4 * r) * SK_LD + 16 * tj + li] = acc[r];
  __syncthreads();
  v4d y = mm(Ds);
  if (has_pred) {
    // on the chain: X_b = y - M_b X_pred
    __syncthreads();
    if (t == 0) sk_wait(&flags[pred * ngrp + g], err);
    __syncthreads();
    load_x(pred);
    __syncthreads();
    const v4d z = mm(Ms);
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] -= z[r];
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = 16 * ti + lq + 4 * r, col = 16 * tj + li;
    if (row < nb_rows && col < ncols)
      __hip_atomic_store(&B[r0 + row + (long)(c0 + col) * ldb], y[r], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (t == 0) __hip_atomic_store(&flags[b * ngrp + g], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_zero_i32(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// All nblk blocks (order bb each, same strides) are inverted together.
int trtri_core(hipStream_t s, bool lower, bool unit, int bb, int nblk, const double* A, long lda,
               long sA, double* X, long ldx, long sX, double* T) {
  if (bb <= 0 || nblk <= 0) return ST_OK;
  if (nblk > 65535) {
    set_error("trtri: too many blocks");
    return ST_BADARG;
  }
  hipLaunchKernelGGL(k_trtri_leaf, dim3(ceil_div(bb, LEAF), nblk), dim3(256), 0, s, bb,
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

constexpr int SKINNY_MAX_RHS = 128;

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
  const int nb64 = ceil_div(N, 64);
  size_t sk = (size_t)nb64 * 64 * 64 * sizeof(double) + (size_t)nb64 * 64 * sizeof(int) +
              trtri_core_ws(64, nb64);
  return (inv + tri + w > sk ? inv + tri + w : sk) + 6 * 256;
}

namespace {
int trsm_impl(hipStream_t s, int flags, int N, int nrhs, double alpha, const double* A, long lda,
              double* B, long ldb);
}

int trsm_d(hipStream_t s, int flags, int N, int nrhs, double alpha, const double* A, long lda,
           double* B, long ldb) {
  const size_t mark = ws_mark();
  const int st = trsm_impl(s, flags, N, nrhs, alpha, A, lda, B, ldb);
  ws_release(mark);
  return st;
}

namespace {
int trsm_impl(hipStream_t s, int flags, int N, int nrhs, double alpha, const double* A, long lda,
              double* B, long ldb) {
  if (N <= 0 || nrhs <= 0) return ST_OK;
  const bool right = flags & F_RIGHT;
  const bool upper = flags & F_UPPER;  // dmmat_solve_blk tests UPPER first (:879-906)
  const bool trans = flags & F_TRANSA;
  const bool unit = flags & F_UNIT;
  const bool eff_lower = upper == trans;  // op(A) is lower triangular
  if (alpha != 1.0) MOPS_TRY(scale_tile_d(s, right ? nrhs : N, right ? N : nrhs, alpha, B, ldb));
  if (alpha == 0.0) return ST_OK;

  if (!right && nrhs <= SKINNY_MAX_RHS && N > SK_BS) {
    const int nb64 = ceil_div(N, SK_BS), ngrp = ceil_div(nrhs, SK_CG);
    double* D = (double*)ws_carve((size_t)nb64 * SK_BS * SK_BS * sizeof(double));
    int* fl = (int*)ws_carve((size_t)(nb64 * ngrp + 1) * sizeof(int));
    if (!D || !fl) {
      set_error("trsm: workspace too small");
      return ST_NOMEM;
    }
    MOPS_TRY(trtri_blocks_d(s, !upper, unit, N, SK_BS, A, lda, D));
    hipLaunchKernelGGL(k_zero_i32, dim3(ceil_div(nb64 * ngrp + 1, 256)), dim3(256), 0, s, fl,
                       nb64 * ngrp + 1);
    hipLaunchKernelGGL(k_trsm_skinny, dim3(nb64 * ngrp), dim3(1024), 0, s, N, nrhs,
                       eff_lower ? 1 : 0, trans ? 1 : 0, A, lda, D, B, ldb, fl,
                       fl + nb64 * ngrp);
    MOPS_CHECK_HIP(hipGetLastError());
    return ST_OK;
  }
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
}  // namespace

}  // namespace mops
