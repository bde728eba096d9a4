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

// ---- LEFT solve for few right-hand sides: one kernel, blocks chained
//
// Workgroup (b, g), 256 threads, solves block row b (64 rows) of
// op(A) X = B for the column group g (16 columns); wave w owns rows
// 16w..16w+15 of the block. With p the block solved just before b (b-1
// forward, b+1 backward), q the one before p, and D_b = inv(op(A)_bb):
//   off the chain:  acc = B_b - sum_{c solved before p} op(A)_bc X_c,
//                   y = D_b acc,   M_b = D_b op(A)_bp
//   on the chain:   X_b = y - M_b X_p   (one 64x16x64 product)
// X_b goes to B (the result) and to the workspace W, which was filled with
// a signalling-NaN sentinel that no arithmetic result can equal. The
// workgroups of the next two blocks (p = b and q = b) poll their X values
// in W until no sentinel is left: the data is the signal, so a hand-off
// costs one one-way store and one load round trip. Blocks solved earlier are
// waited for on a flag per (block, group), raised once every W store of the
// block has completed (one polled word instead of the whole block, for the
// many workgroups that wait at the front of the chain). Every exchanged value
// is an agent-scope atomic load or store, so no cache maintenance is needed.
// Workgroups are numbered in solve order, so every block a workgroup waits
// for belongs to a workgroup dispatched before it. The products run on the
// FP64 matrix cores (v_mfma_f64_16x16x4): A operand lane l = (row l%16,
// k l/16), B operand lane l = (k l/16, column l%16), result lane l, reg r =
// (row l/16 + 4r, column l%16). D_b and M_b stay in registers in the A
// operand layout (M_b computed as M_b^T = op(A)_bp^T D_b^T, whose result
// layout is that A operand layout).
constexpr int TC_BS = 64, TC_CG = 16, TC_T = 256, TC_KS = TC_BS / 4;
constexpr int TC_LDA = TC_BS + 1;  // padded row of a 64x64 LDS block
constexpr int TC_LDX = TC_CG + 1;  // padded row of a 64x16 LDS block
constexpr unsigned long long TC_SENT = 0x7FF5A5A5A5A5A5A5ull;  // quiet bit clear
constexpr unsigned long long TC_QNAN = 0x7FF8000000000000ull;
constexpr unsigned TC_SPIN_LIMIT = 1u << 24;

__device__ __forceinline__ bool tc_is_sent(double v) {
  return (unsigned long long)__double_as_longlong(v) == TC_SENT;
}

// Waits until *flag == ready for at most `limit` polls; then raises the
// status bit and gives up (limit 0: gives up at once, without polling).
__device__ __forceinline__ void tc_wait_flag(const int* flag, int ready, int* err,
                                             unsigned limit) {
  for (unsigned spins = 0;; ++spins) {
    if (spins >= limit) {
      raise_status(err, STATUS_TRSM_WAIT);
      break;
    }
    if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == ready) break;
    __builtin_amdgcn_s_sleep(1);
  }
}

template <bool TRANS>
__global__ __launch_bounds__(TC_T) void k_trsm_chain(int N, int nrhs, int fwd,
                                                     const double* __restrict__ A, long lda,
                                                     const double* __restrict__ Dinv, double* B,
                                                     long ldb, double* W, long ldw, int* flags,
                                                     int ready, int* err, unsigned spin_limit) {
  __shared__ double As[TC_BS * TC_LDA];  // op(A) block [row][k]
  __shared__ double Xs[TC_BS * TC_LDX];  // X block [k][col]
  __shared__ double Ys[TC_BS * TC_LDX];  // acc [k][col]
  const int nblk = (N + TC_BS - 1) / TC_BS;
  const int ngrp = (nrhs + TC_CG - 1) / TC_CG;
  const int step = blockIdx.x / ngrp, g = blockIdx.x % ngrp;
  auto blk = [&](int it) { return fwd ? it : nblk - 1 - it; };  // block solved at step it
  const int b = blk(step);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int li = lane & 15, lq = lane >> 4;
  const int r0 = b * TC_BS, c0 = g * TC_CG;
  const int nb_rows = min(TC_BS, N - r0);
  const int ncols = min(TC_CG, nrhs - c0);
  const int row_a = 16 * w + li;  // this lane's row in the A operand layout

  // op(A) rows of block b, columns of block c, staged as As[i][k] (loads
  // run along the contiguous dimension of A)
  constexpr int SPER = TC_BS * TC_BS / TC_T;
  auto stage_fetch = [&](int c, double* v) {
    const int nk = min(TC_BS, N - c * TC_BS);
#pragma unroll
    for (int q = 0; q < SPER; ++q) {
      const int idx = t + q * TC_T;
      const int x = idx % TC_BS, y = idx / TC_BS;  // x runs along memory
      if (!TRANS)  // op(A)[i = x][k = y] = A[r0 + x][c*64 + y]
        v[q] = (x < nb_rows && y < nk) ? A[(long)r0 + x + ((long)c * TC_BS + y) * lda] : 0.0;
      else         // op(A)[i = y][k = x] = A[c*64 + x][r0 + y]
        v[q] = (y < nb_rows && x < nk) ? A[(long)c * TC_BS + x + ((long)r0 + y) * lda] : 0.0;
    }
  };
  auto stage_store = [&](const double* v) {
#pragma unroll
    for (int q = 0; q < SPER; ++q) {
      const int idx = t + q * TC_T;
      const int x = idx % TC_BS, y = idx / TC_BS;
      if (!TRANS)
        As[x * TC_LDA + y] = v[q];
      else
        As[y * TC_LDA + x] = v[q];
    }
  };
  // A operand of op(A)_bc for this wave straight from memory (no transpose
  // needed: lanes run along the rows of A)
  auto direct_fetch = [&](int c, double* a) {
    const int nk = min(TC_BS, N - c * TC_BS);
#pragma unroll
    for (int s = 0; s < TC_KS; ++s) {
      const int k = 4 * s + lq;
      a[s] = (row_a < nb_rows && k < nk) ? A[(long)r0 + row_a + ((long)c * TC_BS + k) * lda] : 0.0;
    }
  };
  // X block c (this group's columns) into Xs[k][col]; with poll, each value
  // is loaded again until it is no longer the sentinel
  auto load_x = [&](int c, bool poll) {
    const int nk = min(TC_BS, N - c * TC_BS);
    constexpr int XPER = TC_BS * TC_CG / TC_T;
    double v[XPER];
    bool need[XPER];
#pragma unroll
    for (int q = 0; q < XPER; ++q) {
      const int idx = t + q * TC_T;
      const int k = idx % TC_BS, col = idx / TC_BS;
      need[q] = k < nk && col < ncols;
      v[q] = need[q] ? __hip_atomic_load(&W[(long)c * TC_BS + k + (long)(c0 + col) * ldw],
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                     : 0.0;
    }
    if (poll) {
      for (unsigned spins = 0;; ++spins) {
        if (spins >= spin_limit) {
          raise_status(err, STATUS_TRSM_WAIT);
          break;
        }
        bool pending = false;
#pragma unroll
        for (int q = 0; q < XPER; ++q) pending |= need[q] && tc_is_sent(v[q]);
        if (!pending) break;
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int q = 0; q < XPER; ++q) {
          const int idx = t + q * TC_T;
          const int k = idx % TC_BS, col = idx / TC_BS;
          if (need[q] && tc_is_sent(v[q]))
            v[q] = __hip_atomic_load(&W[(long)c * TC_BS + k + (long)(c0 + col) * ldw],
                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < XPER; ++q) {
      const int idx = t + q * TC_T;
      Xs[(idx % TC_BS) * TC_LDX + idx / TC_BS] = v[q];
    }
  };
  // this wave's 16x16 tile of (64 x 64 operand a, A layout) times S[k][col]
  auto mm = [&](const double* a, const double* S) {
    v4d d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < TC_KS; ++s)
      d = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], S[(4 * s + lq) * TC_LDX + li], d, 0, 0, 0);
    return d;
  };

  // D_b in the A operand layout: dst[s] = op(D_b)[row_a][4s + lq]
  double dst[TC_KS];
  {
    const double* D = Dinv + (long)b * TC_BS * TC_BS;
#pragma unroll
    for (int s = 0; s < TC_KS; ++s) {
      const int k = 4 * s + lq;
      dst[s] = (row_a < nb_rows && k < nb_rows)
                   ? (TRANS ? D[k + (long)row_a * TC_BS] : D[row_a + (long)k * TC_BS]) : 0.0;
    }
  }
  // acc = B_b tile in the result layout: reg r <-> row 16w + lq + 4r, column li
  v4d acc;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = 16 * w + lq + 4 * r;
    acc[r] = (row < nb_rows && li < ncols) ? B[r0 + row + (long)(c0 + li) * ldb] : 0.0;
  }
  // M_b = D_b op(A)_bp, as M_b^T tiles t: result reg r of tile t is
  // M_b[row_a][16t + lq + 4r] = the A operand value of k-step 4t + r
  double mst[TC_KS];
  if (step > 0) {
    double v[SPER];
    stage_fetch(blk(step - 1), v);
    stage_store(v);
    __syncthreads();
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
      v4d d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < TC_KS; ++s)
        d = __builtin_amdgcn_mfma_f64_16x16x4f64(As[(4 * s + lq) * TC_LDA + 16 * tt + li], dst[s],
                                                 d, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) mst[4 * tt + r] = d[r];
    }
  }
  // off the chain: every block solved before p, in solve order; the last of
  // them (q) by polling its data, the others after their flag. The next
  // block's op(A) values are fetched while the current one is used.
  const int noff = step - 1;
  double pre[TRANS ? SPER : TC_KS];
  if (noff > 0) {
    if (TRANS)
      stage_fetch(blk(0), pre);
    else
      direct_fetch(blk(0), pre);
  }
  for (int it = 0; it < noff; ++it) {
    const int c = blk(it);
    const bool last = it + 1 == noff;
    double a[TC_KS];
    __syncthreads();  // As and Xs are free
    if (TRANS) {
      stage_store(pre);
    } else {
#pragma unroll
      for (int s = 0; s < TC_KS; ++s) a[s] = pre[s];
    }
    if (!last) {
      if (TRANS)
        stage_fetch(blk(it + 1), pre);
      else
        direct_fetch(blk(it + 1), pre);
    }
    if (!last) {
      if (t == 0) tc_wait_flag(&flags[c * ngrp + g], ready, err, spin_limit);
      __syncthreads();
    }
    load_x(c, last);
    __syncthreads();
    if (TRANS) {
#pragma unroll
      for (int s = 0; s < TC_KS; ++s) a[s] = As[row_a * TC_LDA + 4 * s + lq];
    }
    const v4d d = mm(a, Xs);
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] -= d[r];
  }
  // y = D_b acc
#pragma unroll
  for (int r = 0; r < 4; ++r) Ys[(16 * w + lq + 4 * r) * TC_LDX + li] = acc[r];
  __syncthreads();  // also: every wave is done with Xs
  v4d y = mm(dst, Ys);
  if (step > 0) {
    // on the chain: X_b = y - M_b X_p
    load_x(blk(step - 1), true);
    __syncthreads();
    const v4d z = mm(mst, Xs);
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] -= z[r];
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = 16 * w + lq + 4 * r;
    if (row < nb_rows && li < ncols) {
      double v = y[r];
      if (tc_is_sent(v)) v = __longlong_as_double((long long)TC_QNAN);
      __hip_atomic_store(&W[r0 + row + (long)(c0 + li) * ldw], v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      B[r0 + row + (long)(c0 + li) * ldb] = v;
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (t == 0)
    __hip_atomic_store(&flags[b * ngrp + g], ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_fill_u64(unsigned long long* p, long n, unsigned long long v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = v;
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

size_t trtri_blocks_ws_bytes(int n, int bs) {
  if (n <= 0 || bs <= 0) return 0;
  const int nfull = n / bs;
  return trtri_core_ws(bs, nfull > 0 ? nfull : 1) + 256;
}

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
  const int ngrp = ceil_div(nrhs, 16);
  size_t sk = (size_t)nb64 * 64 * 64 * sizeof(double) + (size_t)nb64 * 64 * sizeof(int) +
              (size_t)nb64 * 64 * ngrp * 16 * sizeof(double) + trtri_core_ws(64, nb64);
  return (inv + tri + w > sk ? inv + tri + w : sk) + 6 * 256;
}

namespace {
// Workspace of the chained LEFT solve: the inverses of the 64x64 diagonal
// blocks and the sentinel-filled copy of the solution (the flags come from
// the thread's flag pool, per solve).
struct ChainWs {
  double* D;
  double* W;
  long ldw;
  int nb64, ngrp;
};

int chain_ws(int N, int nrhs, ChainWs& w) {
  w.nb64 = ceil_div(N, TC_BS);
  w.ngrp = ceil_div(nrhs, TC_CG);
  w.ldw = (long)w.nb64 * TC_BS;
  w.D = (double*)ws_carve((size_t)w.nb64 * TC_BS * TC_BS * sizeof(double));
  w.W = (double*)ws_carve((size_t)w.ldw * w.ngrp * TC_CG * sizeof(double));
  if (!w.D || !w.W) {
    set_error("trsm: workspace too small");
    return ST_NOMEM;
  }
  return ST_OK;
}

// op(A) X = B for a LEFT solve, with w.D holding the inverses of A's
// diagonal blocks (of the triangle the flags select).
int chain_solve(hipStream_t s, int flags, int N, int nrhs, const double* A, long lda, double* B,
                long ldb, const ChainWs& w) {
  const bool trans = flags & F_TRANSA;
  const bool eff_lower = ((flags & F_UPPER) != 0) == trans;  // op(A) is lower triangular
  int* fl = nullptr;
  int ready = 0;
  if (!signal_flags((long)w.nb64 * w.ngrp, &fl, &ready)) {
    set_error("trsm: signal flags could not be allocated");
    return ST_NOMEM;
  }
  const long nw = w.ldw * w.ngrp * TC_CG;
  hipLaunchKernelGGL(k_fill_u64, dim3((int)std::min<long>((nw + 255) / 256, 4096)), dim3(256), 0,
                     s, (unsigned long long*)w.W, nw, TC_SENT);
  const dim3 grid(w.nb64 * w.ngrp), block(TC_T);
  // a wait that gives up raises STATUS_TRSM_WAIT in the caller's status word
  int* err = active_status();
  if (!err) {
    set_error("trsm: status word could not be allocated");
    return ST_NOMEM;
  }
  if (trans)
    hipLaunchKernelGGL(k_trsm_chain<true>, grid, block, 0, s, N, nrhs, eff_lower ? 1 : 0, A, lda,
                       w.D, B, ldb, w.W, w.ldw, fl, ready, err, wait_spin_limit(TC_SPIN_LIMIT));
  else
    hipLaunchKernelGGL(k_trsm_chain<false>, grid, block, 0, s, N, nrhs, eff_lower ? 1 : 0, A, lda,
                       w.D, B, ldb, w.W, w.ldw, fl, ready, err, wait_spin_limit(TC_SPIN_LIMIT));
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

int trsm_impl(hipStream_t s, int flags, int N, int nrhs, double alpha, const double* A, long lda,
              double* B, long ldb);
}  // namespace

int trsm_pair_d(hipStream_t s, int flags1, int flags2, int N, int nrhs, const double* A,
                long lda, double* B, long ldb) {
  const bool same_triangle = ((flags1 ^ flags2) & ~F_TRANSA) == 0;
  if (!same_triangle || (flags1 & F_RIGHT) || nrhs > SKINNY_MAX_RHS || N <= TC_BS || nrhs <= 0) {
    MOPS_TRY(trsm_d(s, flags1, N, nrhs, 1.0, A, lda, B, ldb));
    return trsm_d(s, flags2, N, nrhs, 1.0, A, lda, B, ldb);
  }
  const size_t mark = ws_mark();
  ChainWs w;
  int st = chain_ws(N, nrhs, w);
  if (st == ST_OK)
    st = trtri_blocks_d(s, !(flags1 & F_UPPER), (flags1 & F_UNIT) != 0, N, TC_BS, A, lda, w.D);
  if (st == ST_OK) st = chain_solve(s, flags1, N, nrhs, A, lda, B, ldb, w);
  if (st == ST_OK) st = chain_solve(s, flags2, N, nrhs, A, lda, B, ldb, w);
  ws_release(mark);
  return st;
}

static bool trsm_chained(int flags, int N, int nrhs) {
  return !(flags & F_RIGHT) && nrhs > 0 && nrhs <= SKINNY_MAX_RHS && N > TC_BS;
}

size_t trsm_prep_w_bytes(int N, int nrhs) {
  return (size_t)ceil_div(N, TC_BS) * TC_BS * ceil_div(nrhs, TC_CG) * TC_CG * sizeof(double) + 256;
}

size_t trsm_prep_ws_bytes(int N, int nrhs) {
  const int nb64 = ceil_div(N, TC_BS);
  return (size_t)nb64 * TC_BS * TC_BS * sizeof(double) + trsm_prep_w_bytes(N, nrhs) +
         trtri_core_ws(TC_BS, nb64) + 3 * 256;
}

int trsm_prep_d(hipStream_t s, int flags, int N, int nrhs, const double* A, long lda, double* W,
                TrsmPrep& p) {
  p = TrsmPrep();
  if (!trsm_chained(flags, N, nrhs)) return ST_OK;
  p.nb64 = ceil_div(N, TC_BS);
  p.ngrp = ceil_div(nrhs, TC_CG);
  p.ldw = (long)p.nb64 * TC_BS;
  p.D = (double*)ws_carve((size_t)p.nb64 * TC_BS * TC_BS * sizeof(double));
  p.W = W ? W : (double*)ws_carve((size_t)p.ldw * p.ngrp * TC_CG * sizeof(double));
  if (!p.D || !p.W) {
    set_error("trsm: workspace too small");
    return ST_NOMEM;
  }
  const size_t mark = ws_mark();  // the inversion's scratch only
  const int st = trtri_blocks_d(s, !(flags & F_UPPER), (flags & F_UNIT) != 0, N, TC_BS, A, lda,
                                p.D);
  ws_release(mark);
  return st;
}

int trsm_prep_solve_d(hipStream_t s, int flags, int N, int nrhs, const double* A, long lda,
                      double* B, long ldb, const TrsmPrep& p) {
  if (!p.D) return trsm_d(s, flags, N, nrhs, 1.0, A, lda, B, ldb);
  ChainWs w;
  w.D = p.D;
  w.W = p.W;
  w.ldw = p.ldw;
  w.nb64 = p.nb64;
  w.ngrp = p.ngrp;
  return chain_solve(s, flags, N, nrhs, A, lda, B, ldb, w);
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

  if (!right && nrhs <= SKINNY_MAX_RHS && N > TC_BS) {
    ChainWs w;
    MOPS_TRY(chain_ws(N, nrhs, w));
    MOPS_TRY(trtri_blocks_d(s, !upper, unit, N, TC_BS, A, lda, w.D));
    return chain_solve(s, flags, N, nrhs, A, lda, B, ldb, w);
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
