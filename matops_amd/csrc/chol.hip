// Blocked Cholesky factorization and solve on the GPU.
//
// Replaces the panel loop of blockedCHOL / unblockedCHOL (chol.go:17-94),
// DecomposeCHOL (chol.go:116-127) and SolveCHOL (chol.go:147-158). The
// reference's step for a diagonal block of width nb is
//   A11 = chol(A11)  (per column: sqrt, a21 /= a11, SYR; chol.go:29-46)
//   A21 = A21 * tril(A11)^-T          (SolveTrm RIGHT|LOWER|TRANSA, :79)
//   A22 = A22 - A21 * A21^T           (RankUpdateSym LOWER, :81)
// and the UPPER variant with A12 / U11^-T (:82-87). The same step runs here
// with the panel (diagonal block and the rows below it, up to 512 columns)
// factored in one launch of k_chol_panel: 64-column diagonal blocks in
// registers (the reference's column algorithm), their rows below solved and
// the rest of the panel updated on the matrix cores, the workgroups chained
// by flags. Wider panels are split recursively with trapezoidal GEMM
// updates between the parts. The trailing update of step k runs on the
// caller's stream while the next panel is factored on a second stream.
#include <climits>
#include <cmath>

#include "mops_solvers.hpp"

namespace mops {
namespace {

constexpr int LEAF = 64;

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

constexpr unsigned LEAF_SPIN_LIMIT = 1u << 26;

__device__ __forceinline__ double ld_agent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr int CL_T = 256;  // threads per workgroup of the leaf kernel
constexpr int CL_LD = LEAF + 1;

// 1/sqrt(d) for a finite d > 0: the hardware estimate (v_rsq_f64) refined
// by two Newton-Raphson steps y <- y + y (1 - d y^2) / 2, each of which
// roughly doubles the number of correct bits.
__device__ __forceinline__ double rsqrt_newton(double d) {
  double y = __builtin_amdgcn_rsq(d);
  double e = fma(-d * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-d * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

// Element (r, c), r >= c, of the panel's factor L in memory: the panel's
// corner is P; LOWER stores L itself, UPPER its transpose U = L^T.
template <bool UPPER>
__device__ __forceinline__ double* lp(double* P, long lda, int r, int c) {
  return UPPER ? P + c + (long)r * lda : P + r + (long)c * lda;
}

// Thread 0 waits for a flag set by another workgroup of the same launch
// (always one with a lower index, so it has been dispatched); a wait that
// exceeds the spin limit raises err and gives up instead of hanging.
__device__ __forceinline__ void cp_wait(const int* f, int* err) {
  unsigned spins = 0;
  while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
    if (++spins > LEAF_SPIN_LIMIT) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

constexpr int CP_MAXB = 8;  // diagonal blocks of 64 in one panel launch (512 columns)
constexpr int CP_FLAGS = CP_MAXB + CP_MAXB * CP_MAXB;  // ints per panel launch

// One Cholesky panel of up to 512 columns (the diagonal blocks of 64 and
// every row below them) in one launch of 256-thread workgroups.
//
// Rows are split into 64-row chunks: workgroup b < nsteps owns the rows of
// diagonal block b, the others 64 rows each below the panel's diagonal.
// Block column j is the reference's step for 64 columns (chol.go:29-46 on
// the diagonal block, the TRSM of chol.go:79 below it, the SYR updates of
// the columns right of it); here step j is:
//  * workgroup j factors its diagonal block, kept since the start of the
//    launch as four 16 x 16 MFMA accumulator tiles per wave (wave w owns
//    columns 16w..16w+15). It moves its 16 columns to row layout (through
//    LDS) and factors them one column at a time: L[j][j] = sqrt(d), the
//    column below scaled by 1/L[j][j] (InvScale, chol.go:42; built with
//    -ffast-math, the reference's divisions are a multiply by the
//    reciprocal, here 1/sqrt(d) directly), the rank-1 update of the wave's
//    remaining columns (the SYR of chol.go:44). The next diagonal element is
//    formed from two readlanes ahead of the full update, which keeps the
//    serial chain per column short. After one barrier the waves right of it
//    apply the rank-16 update of their tiles on the matrix cores. A NaN
//    diagonal (chol.go:33-36) is reported once, for the first column. Wave
//    0 then inverts the four 16 x 16 diagonal blocks (lane = block and
//    column, forward substitution), and a flag publishes L and the
//    inverses;
//  * every workgroup below it waits for that flag and solves its rows,
//    X L^T = B, 16 rows per wave, 16 columns at a time:
//      X_b^T = inv(L_bb) (B_b^T - sum_{a<b} L_ba X_a^T)
//    (the MFMA result layout of a solved block X_a^T is the B-operand
//    layout the later blocks need, so solved blocks stay in registers);
//  * then it subtracts X X_c^T from its block in each column block c > j
//    of the panel, X_c being the rows just solved by the diagonal
//    workgroup c (published with agent-scope stores and a flag). The block
//    of column j + 1 stays in LDS as the next step's right-hand side; a
//    diagonal workgroup's own diagonal block is updated in its registers.
// So the critical path of a step is one diagonal factorization, one
// 64-row solve and one 64 x 64 update, all on chip; every wait is on a
// workgroup with a lower index. The UPPER case runs the same on U^T.
template <bool UPPER>
__global__ __launch_bounds__(CL_T) void k_chol_panel(int m, int pw, double* P, long lda,
                                                     double* Minv, int goff, int* info,
                                                     double* badval, int* flags, int* err) {
  constexpr bool upper = UPPER;
  __shared__ double s1[LEAF * CL_LD];
  __shared__ double s2[LEAF * CL_LD];
  __shared__ double dv[LEAF];  // diagonal before the square root
  const int t = threadIdx.x, l = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int li = l % 16, lq = l / 16;
  const int R0 = 16 * w;
  const int nsteps = (pw + LEAF - 1) / LEAF;
  const int b = blockIdx.x;
  const bool diag = b < nsteps;
  const int row0 = diag ? LEAF * b : pw + LEAF * (b - nsteps);
  const int nr = diag ? min(LEAF, pw - row0) : min(LEAF, m - row0);  // rows held
  const int jend = diag ? b : nsteps;                                 // solves
  int* dflag = flags;      // dflag[j]: diagonal block j factored, inverses published
  int* xflag = flags + CP_MAXB;  // xflag[CP_MAXB j + c]: rows of block c solved in step j
  double* Bs = s1;         // Bs[r * CL_LD + k]: right-hand side, then X, of step j
  double* Ls = s2;         // L below the 16 x 16 diagonal blocks, or a published X

  // diagonal workgroups: the whole symmetric block (rows and columns beyond
  // nr: identity); tile q of wave w, register e of lane l = A[16q + lq + 4e][16w + li]
  v4d dacc[4];
  if (diag) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = 16 * q + lq + 4 * e, c = R0 + li;
        double v = r == c ? 1.0 : 0.0;
        if (r < nr && c < nr)
          v = c <= r ? *lp<UPPER>(P, lda, row0 + r, row0 + c) : *lp<UPPER>(P, lda, row0 + c, row0 + r);
        dacc[q][e] = v;
      }
  }
  if (jend > 0) {
    double v[LEAF * LEAF / CL_T];
#pragma unroll
    for (int q = 0; q < LEAF * LEAF / CL_T; ++q) {
      const int idx = t + q * CL_T;
      const int a0 = idx % LEAF, a1 = idx / LEAF;  // contiguous index first
      const int r = upper ? a1 : a0, c = upper ? a0 : a1;
      v[q] = 0.0;
      if (r < nr && c < min(LEAF, pw)) v[q] = *lp<UPPER>(P, lda, row0 + r, c);
    }
#pragma unroll
    for (int q = 0; q < LEAF * LEAF / CL_T; ++q) {
      const int idx = t + q * CL_T;
      const int a0 = idx % LEAF, a1 = idx / LEAF;
      const int r = upper ? a1 : a0, c = upper ? a0 : a1;
      Bs[r * CL_LD + c] = v[q];
    }
  }

#pragma unroll 1
  for (int j = 0; j < jend; ++j) {
    const int c0 = LEAF * j, wj = min(LEAF, pw - c0);
    if (t == 0) {
      cp_wait(&dflag[j], err);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    // the acquire by thread 0 invalidated this CU's caches; after the
    // barrier every thread's loads see what the diagonal workgroup released
    __syncthreads();
    double mv[4][4];  // inverse blocks in A-operand layout: inv(L_bb)[li][4s + lq]
#pragma unroll
    for (int bb = 0; bb < 4; ++bb)
#pragma unroll
      for (int s = 0; s < 4; ++s) mv[bb][s] = Minv[j * 1024 + bb * 256 + li * 16 + 4 * s + lq];
    {
      double v[LEAF * LEAF / CL_T];
#pragma unroll
      for (int q = 0; q < LEAF * LEAF / CL_T; ++q) {
        const int idx = t + q * CL_T;
        const int a0 = idx % LEAF, a1 = idx / LEAF;
        const int r = upper ? a1 : a0, c = upper ? a0 : a1;  // L[r][c] of block j
        v[q] = 0.0;
        if (r < wj && c < wj && (c >> 4) < (r >> 4)) v[q] = *lp<UPPER>(P, lda, c0 + r, c0 + c);
      }
#pragma unroll
      for (int q = 0; q < LEAF * LEAF / CL_T; ++q) {
        const int idx = t + q * CL_T;
        const int a0 = idx % LEAF, a1 = idx / LEAF;
        const int r = upper ? a1 : a0, c = upper ? a0 : a1;
        Ls[r * CL_LD + c] = v[q];
      }
    }
    __syncthreads();
    // solve; lane l, register e of xt[bb] holds X[R0 + l%16][16bb + l/16 + 4e]
    v4d xt[4];
#pragma unroll
    for (int bb = 0; bb < 4; ++bb) {
      v4d y;
#pragma unroll
      for (int e = 0; e < 4; ++e) y[e] = Bs[(R0 + li) * CL_LD + 16 * bb + lq + 4 * e];
#pragma unroll
      for (int a = 0; a < bb; ++a)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          // Y^T[c][r] -= sum_k L[16bb + c][16a + k] X[r][16a + k]
          const double lv = -Ls[(16 * bb + li) * CL_LD + 16 * a + 4 * s + lq];
          y = __builtin_amdgcn_mfma_f64_16x16x4f64(lv, xt[a][s], y, 0, 0, 0);
        }
      v4d x = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4; ++s)
        x = __builtin_amdgcn_mfma_f64_16x16x4f64(mv[bb][s], y[s], x, 0, 0, 0);
      xt[bb] = x;
    }
    // X to LDS (own rows; zero beyond the block's width)
#pragma unroll
    for (int bb = 0; bb < 4; ++bb)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int c = 16 * bb + lq + 4 * e;
        Bs[(R0 + li) * CL_LD + c] = c < wj ? xt[bb][e] : 0.0;
      }
    __syncthreads();
    // X is final: to memory; a diagonal workgroup's rows are also a column
    // block of the panel, published for the other workgroups' updates
#pragma unroll
    for (int q = 0; q < LEAF * LEAF / CL_T; ++q) {
      const int idx = t + q * CL_T;
      const int a0 = idx % LEAF, a1 = idx / LEAF;
      const int r = upper ? a1 : a0, c = upper ? a0 : a1;
      if (r < nr && c < wj) {
        const double v = Bs[r * CL_LD + c];
        double* p = lp<UPPER>(P, lda, row0 + r, c0 + c);
        if (diag)
          __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
          *p = v;
      }
    }
    if (diag) {
      // own diagonal block: dacc -= X X^T (with lanes along the rows the
      // result lands in the layout of dacc: the block is symmetric); its
      // MFMAs run while the stores of X drain, ahead of the flag
#pragma unroll
      for (int s4 = 0; s4 < LEAF; s4 += 4) {
        const double xa = -Bs[(R0 + li) * CL_LD + s4 + lq];
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
          dacc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(Bs[(16 * tj + li) * CL_LD + s4 + lq], xa,
                                                          dacc[tj], 0, 0, 0);
      }
      // the flag of the last step is published after the first column
      // block of the factorization, by when the stores of X have drained
      // without holding up the factorization
      if (j + 1 < jend) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (t == 0) __hip_atomic_store(&xflag[CP_MAXB * j + b], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    // updates of the other column blocks right of j, right to left, so that
    // the block of column j + 1 (the next right-hand side) comes last
    const int chi = diag ? b - 1 : nsteps - 1;
#pragma unroll 1
    for (int cb = chi; cb > j; --cb) {
      const int cc0 = LEAF * cb, wcb = min(LEAF, pw - cc0);
      if (t == 0) cp_wait(&xflag[CP_MAXB * j + cb], err);
      __syncthreads();
      {
        double v[LEAF * LEAF / CL_T];  // all loads in flight before the first LDS write
#pragma unroll
        for (int q = 0; q < LEAF * LEAF / CL_T; ++q) {
          const int idx = t + q * CL_T;
          const int a0 = idx % LEAF, a1 = idx / LEAF;
          const int r = upper ? a1 : a0, c = upper ? a0 : a1;  // X_cb[r][c]
          v[q] = 0.0;
          if (r < wcb && c < wj) v[q] = ld_agent(lp<UPPER>(P, lda, cc0 + r, c0 + c));
        }
#pragma unroll
        for (int q = 0; q < LEAF * LEAF / CL_T; ++q) {
          const int idx = t + q * CL_T;
          const int a0 = idx % LEAF, a1 = idx / LEAF;
          const int r = upper ? a1 : a0, c = upper ? a0 : a1;
          Ls[r * CL_LD + c] = v[q];
        }
      }
      __syncthreads();
      // wave w: rows R0.. of this chunk, the 64 columns of block cb, with
      // lanes along the contiguous index in memory: LOWER acc[tj][e] =
      // (row R0 + li, column 16tj + lq + 4e), UPPER (row R0 + lq + 4e,
      // column 16tj + li)
      v4d acc[4];
#pragma unroll
      for (int tj = 0; tj < 4; ++tj)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = upper ? R0 + lq + 4 * e : R0 + li;
          const int c = upper ? 16 * tj + li : 16 * tj + lq + 4 * e;
          acc[tj][e] = r < nr && c < wcb ? *lp<UPPER>(P, lda, row0 + r, cc0 + c) : 0.0;
        }
#pragma unroll
      for (int s4 = 0; s4 < LEAF; s4 += 4) {
        const double xa = -Bs[(R0 + li) * CL_LD + s4 + lq];
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) {
          const double lb = Ls[(16 * tj + li) * CL_LD + s4 + lq];
          if constexpr (UPPER)
            acc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa, lb, acc[tj], 0, 0, 0);
          else
            acc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(lb, xa, acc[tj], 0, 0, 0);
        }
      }
      if (cb == j + 1) {
        __syncthreads();  // every wave is done with X in Bs
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int r = upper ? R0 + lq + 4 * e : R0 + li;
            const int c = upper ? 16 * tj + li : 16 * tj + lq + 4 * e;
            Bs[r * CL_LD + c] = acc[tj][e];
          }
      } else {
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int r = upper ? R0 + lq + 4 * e : R0 + li;
            const int c = upper ? 16 * tj + li : 16 * tj + lq + 4 * e;
            if (r < nr && c < wcb) *lp<UPPER>(P, lda, row0 + r, cc0 + c) = acc[tj][e];
          }
      }
      __syncthreads();  // Ls (and Bs) are reused
    }
  }
  if (!diag) return;

  // ---- factor the diagonal block b (n = nr columns) ----
  __syncthreads();
  const int n = nr;
  double* A = lp<UPPER>(P, lda, row0, row0);
  double* Mb = Minv + b * 1024;
  double* Lt = s1;         // Lt[c * 64 + r] = L[r][c], filled block by block
  double* Cs = s2;         // row-layout staging of the owner's tiles, pitch 17
  double* rv = s2 + 1152;  // rv[j] = 1 / L[j][j]
#pragma unroll 1
  for (int jb = 0; jb < 4; ++jb) {
    if (w == jb) {
      // tiles -> row layout a[i] = (row l, column 16jb + i)
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) Cs[(16 * q + lq + 4 * e) * 17 + li] = dacc[q][e];
      double a[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) a[i] = Cs[l * 17 + i];
      double dn = readlane_d(a[0], 16 * jb);
#pragma unroll
      for (int ji = 0; ji < 16; ++ji) {
        const int j = 16 * jb + ji;
        const double d = dn;
        double ri, r;
        if (d > 0.0 && d < INFINITY) {
          ri = rsqrt_newton(d);
          r = d * ri;
        } else {
          r = sqrt(d);
          ri = 1.0 / r;
        }
        const double x = l > j ? a[ji] * ri : 0.0;
        double xn = 0.0;  // x of row j + 1 (= Lt[j * 64 + j + 1])
        if (ji < 15) {
          xn = readlane_d(x, j + 1);
          dn = fma(-xn, xn, readlane_d(a[ji + 1], j + 1));
        }
        a[ji] = l == j ? r : x;
        Lt[j * 64 + l] = a[ji];
        dv[j] = d;
        rv[j] = ri;
        // the next column's update takes x of row j + 1 from the readlane
        // rather than back from LDS (the read after the store above sat on
        // the serial path); the later columns read it from LDS
#pragma unroll
        for (int i = ji + 1; i < 16; ++i)
          a[i] = fma(-x, i == ji + 1 ? xn : Lt[j * 64 + 16 * jb + i], a[i]);
        // keep this column's updates here: sinking them toward their use
        // would put chains of dependent FMAs on the serial path
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    const bool pub = jb == 0 && jend > 0;  // the last step's rows of X (see above)
    if (pub) __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (pub && t == 0)
      __hip_atomic_store(&xflag[CP_MAXB * (jend - 1) + b], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (w == jb) {
      // the finished columns back to A (off the serial path)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int c = 16 * jb + i;
        const double v = Lt[c * 64 + l];
        if (l < n && c < n && c <= l) *lp<UPPER>(A, lda, l, c) = v;
      }
    }
    if (w > jb) {
      // rank-16 update: tile q (rows 16q..) -= L[rows][block jb] L[16w..][block jb]^T
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q >= w) {
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const double* col = Lt + (16 * jb + 4 * s + lq) * 64;
            dacc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-col[16 * q + li], col[R0 + li],
                                                           dacc[q], 0, 0, 0);
          }
        }
      }
    }
  }
  if (w == 0) {
    // first NaN on the diagonal
    const unsigned long long bad = __ballot(l < n && isnan(Lt[l * 64 + l]));
    if (bad && l == 0 && info) {
      const int j = __ffsll((long long)bad) - 1;
      const int old = atomicMin(info, goff + row0 + j);
      if ((old < 0 || old > goff + row0 + j) && badval) *badval = dv[j];
    }
    // inverses of the diagonal blocks: lane l -> block bb = l/16, column
    // c = l%16; z = inv(L_bb) e_c by forward substitution
    const int bb = lq, c = li;
    double z[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) z[q] = q == c ? 1.0 : 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      z[q] *= rv[16 * bb + q];
#pragma unroll
      for (int q2 = q + 1; q2 < 16; ++q2)
        z[q2] = fma(-Lt[(16 * bb + q) * 64 + 16 * bb + q2], z[q], z[q2]);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) Mb[bb * 256 + q * 16 + c] = z[q];
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (t == 0) __hip_atomic_store(&dflag[b], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_zero_flags(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void k_init_info(int* info, double* badval) {
  if (info) *info = 0x7fffffff;
  if (badval) *badval = 0.0;
}

// info: -1 when the factorization succeeded, -2 when a leaf's wait for its
// diagonal block timed out, else the first NaN square root's index.
__global__ void k_fix_info(int* info, const int* err) {
  if (!info) return;
  if (*err) *info = -2;
  else if (*info == 0x7fffffff) *info = -1;
}

// Factor the panel of columns [j0, j0+w) (rows j0..N-1) in place. Up to
// CL_FUSE columns: one launch of k_chol_panel. Wider panels: the recursive
// (left part, trapezoidal update of the right part, right part) scheme. The
// UPPER case is the transpose of each step.
constexpr int CL_FUSE = CP_MAXB * LEAF;

struct LeafFlags {
  int* flags;    // CP_FLAGS zeroed flags per panel launch
  int next;
  int cap;
  int* err;
  double* minv;  // inverses of the current panel's 16 x 16 diagonal blocks (CP_MAXB x 1024)
};

int chol_rec(hipStream_t s, bool upper, int N, int j0, int w, double* A, long lda, LeafFlags& fl,
             int* info, double* badval) {
  if (w <= CL_FUSE) {
    const int m = N - j0;  // rows of the panel
    const int nsteps = ceil_div(w, LEAF);
    if (fl.next + CP_FLAGS > fl.cap) {
      set_error("potrf: panel flags exhausted");
      return ST_NOMEM;
    }
    int* flag = fl.flags + fl.next;
    fl.next += CP_FLAGS;
    hipLaunchKernelGGL(upper ? k_chol_panel<true> : k_chol_panel<false>,
                       dim3(nsteps + ceil_div(m - w, LEAF)), dim3(CL_T), 0, s, m, w,
                       A + j0 + (long)j0 * lda, lda, fl.minv, j0, info, badval, flag, fl.err);
    MOPS_CHECK_HIP(hipGetLastError());
    return ST_OK;
  }
  int h = ((w / 2 + CL_FUSE - 1) / CL_FUSE) * CL_FUSE;
  if (h >= w) h = w - LEAF;
  MOPS_TRY(chol_rec(s, upper, N, j0, h, A, lda, fl, info, badval));
  const int rows = N - j0 - h;
  double* C = A + (j0 + h) + (long)(j0 + h) * lda;
  if (!upper) {
    const double* P = A + (j0 + h) + (long)j0 * lda;
    MOPS_TRY(gemm_d(s, false, true, rows, w - h, h, -1.0, P, lda, P, lda, 1.0, C, lda,
                    GM_LOWER_T));
  } else {
    const double* P = A + j0 + (long)(j0 + h) * lda;
    MOPS_TRY(gemm_d(s, true, false, w - h, rows, h, -1.0, P, lda, P, lda, 1.0, C, lda,
                    GM_UPPER_T));
  }
  return chol_rec(s, upper, N, j0 + h, w - h, A, lda, fl, info, badval);
}

}  // namespace

// The outer block the GPU uses for a requested nb. DecomposeCHOL runs
// unblocked when N < nb or nb == 0 (chol.go:121); the unblocked and blocked
// algorithms compute the same factor, so both run here as the blocked one
// with the block clamped to N. Blocks of 256 and more take 512 columns: a
// panel of 512 is one launch of k_chol_panel, and the trailing updates then
// have k = 512 (the triangle-only GEMM runs 38 TFLOP/s at k = 256 and 49 at
// k = 512 for n = 7680; the Cholesky n = 8192 7.14 -> 6.61 ms). The factor
// differs from the nb = 256 one by rounding only.
static int potrf_block(int N, int nb) {
  if (nb <= 0 || nb > N) nb = N;
  if (nb >= 4 * LEAF) nb = min(N, CL_FUSE);
  return nb;
}

// Panel launches of one factorization: each block column of nb contributes
// at most ceil(kb / 256) launches (ceil(kb / CL_FUSE) when it is split), and
// its recursive split at most one more.
static int potrf_panel_count(int N, int nb) {
  nb = potrf_block(N, nb);
  return ceil_div(N, 4 * LEAF) + 2 * ceil_div(N, nb > 0 ? nb : 1) + 2;
}

size_t potrf_ws_bytes(int N, int nb) {
  return (size_t)(CP_FLAGS * potrf_panel_count(N, nb) + 8) * sizeof(int) +
         (size_t)CP_MAXB * 1024 * sizeof(double) + 2048;
}

namespace {

// Panel k (diagonal block of order kb and everything below / right of it).
int chol_panel(hipStream_t s, bool upper, int N, int k, int kb, double* A, long lda,
               LeafFlags& fl, int* info, double* badval) {
  return chol_rec(s, upper, N, k, kb, A, lda, fl, info, badval);
}

// Update of the trailing block that starts c0 columns after the panel's
// end (rows and columns j0 = k+kb+c0 ... j0+nc of the trailing matrix),
// with the panel's rows/columns r0 = k+kb+c0 ... : trap = only the next
// block column (a trapezoid of nc columns), else the square to the end.
int chol_update(hipStream_t s, bool upper, int N, int k, int kb, int c0, int nc, bool trap,
                double* A, long lda) {
  const int j0 = k + kb + c0;
  const int rows = N - j0;
  if (rows <= 0 || nc <= 0) return ST_OK;
  double* C = A + j0 + (long)j0 * lda;
  if (!upper) {
    const double* P = A + j0 + (long)k * lda;  // A21 rows j0..N
    if (trap)
      return gemm_d(s, false, true, rows, nc, kb, -1.0, P, lda, P, lda, 1.0, C, lda, GM_LOWER_T);
    return gemm_d(s, false, true, rows, rows, kb, -1.0, P, lda, P, lda, 1.0, C, lda, GM_LOWER);
  }
  const double* P = A + k + (long)j0 * lda;  // A12 columns j0..N
  if (trap)
    return gemm_d(s, true, false, nc, rows, kb, -1.0, P, lda, P, lda, 1.0, C, lda, GM_UPPER_T);
  return gemm_d(s, true, false, rows, rows, kb, -1.0, P, lda, P, lda, 1.0, C, lda, GM_UPPER);
}

}  // namespace

int potrf_d(hipStream_t s, int flags, int N, int nb, double* A, long lda, int* info,
            double* badval) {
  if (N <= 0) return ST_OK;
  // blockedCHOL takes the LOWER branch when the bit is set, otherwise UPPER
  // (chol.go:40, :77).
  const bool upper = !(flags & F_LOWER);
  nb = potrf_block(N, nb);
  // the flags of every panel launch, plus an error word
  const int nleaf = CP_FLAGS * potrf_panel_count(N, nb);
  LeafFlags fl;
  fl.flags = (int*)ws_carve((size_t)(nleaf + 1) * sizeof(int));
  fl.minv = (double*)ws_carve((size_t)CP_MAXB * 1024 * sizeof(double));
  fl.next = 0;
  fl.cap = nleaf;
  if (!fl.flags || !fl.minv) {
    set_error("potrf: workspace too small");
    return ST_NOMEM;
  }
  fl.err = fl.flags + nleaf;
  hipLaunchKernelGGL(k_zero_flags, dim3(ceil_div(nleaf + 1, 256)), dim3(256), 0, s, fl.flags,
                     nleaf + 1);
  hipLaunchKernelGGL(k_init_info, dim3(1), dim3(1), 0, s, info, badval);
  MOPS_CHECK_HIP(hipGetLastError());
  const int nsteps = ceil_div(N, nb);
  if (nsteps <= 2) {
    for (int k = 0; k < N; k += nb) {
      const int kb = min(nb, N - k);
      MOPS_TRY(chol_panel(s, upper, N, k, kb, A, lda, fl, info, badval));
      const int m = N - k - kb;
      if (m <= 0) break;
      trace_mark_begin(s, TR_SYRK_UPDATE);
      MOPS_TRY(chol_update(s, upper, N, k, kb, 0, m, false, A, lda));
      trace_mark_end(s, TR_SYRK_UPDATE, (double)m * (m + 1) * kb);
    }
  } else {
    // Look-ahead: the panel of step k+1 (and the update of its block column
    // by step k) runs on a second stream while the caller's stream applies
    // step k to the rest of the trailing matrix. The rest update starts only
    // when the block column update is done, at the moment the next panel
    // becomes ready too: both kernels then enter the dispatcher together,
    // the panel's workgroups are not left waiting for CU slots behind
    // thousands of queued tiles, and the block column update, on the panel
    // chain, has the GPU to itself (Cholesky n = 8192: 29.45 -> 29.7
    // TFLOP/s against starting block column k+2 beside it). The rest update
    // does block column k+2 first (a trapezoid) and then the triangle
    // beyond it: the next block column update needs only the first part, so
    // the panel chain does not wait for the whole triangle.
    hipStream_t s2 = aux_stream();
    hipEvent_t e_fork = sync_event(0), e_col = sync_event(1), e_next = sync_event(2),
               e_join = sync_event(3);
    if (!s2 || !e_fork || !e_col || !e_next || !e_join) {
      set_error("potrf: could not create the look-ahead stream");
      return ST_HIP;
    }
    MOPS_CHECK_HIP(hipEventRecord(e_fork, s));
    MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_fork, 0));
    MOPS_TRY(chol_panel(s2, upper, N, 0, min(nb, N), A, lda, fl, info, badval));
    bool have_next = false;
    for (int k = 0; k < N; k += nb) {
      const int kb = min(nb, N - k);
      const int m = N - k - kb;
      if (m <= 0) break;
      const int kb1 = min(nb, m);  // width of the next block column
      // s2 may touch block column k+1 only after step k-1 has updated it
      if (have_next) MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_next, 0));
      MOPS_TRY(chol_update(s2, upper, N, k, kb, 0, kb1, true, A, lda));
      MOPS_CHECK_HIP(hipEventRecord(e_col, s2));  // after panel k and this update
      have_next = false;
      if (m > kb1) {
        const int mr = m - kb1;
        const int kb2 = min(nb, mr);  // block column k+2
        // block column k+2 (which needs panel k only) and the triangle wait
        // for the block column update, to start together with the next panel
        MOPS_CHECK_HIP(hipStreamWaitEvent(s, e_col, 0));
        MOPS_TRY(chol_update(s, upper, N, k, kb, kb1, kb2, true, A, lda));
        MOPS_CHECK_HIP(hipEventRecord(e_next, s));
        have_next = true;
        if (mr > kb2) {
          // timed when tracing: every launch of the triangle-only GEMM
          const int mq = mr - kb2;
          trace_mark_begin(s, TR_SYRK_UPDATE);
          MOPS_TRY(chol_update(s, upper, N, k, kb, kb1 + kb2, mq, false, A, lda));
          trace_mark_end(s, TR_SYRK_UPDATE, (double)mq * (mq + 1) * kb);
        }
      }
      MOPS_TRY(chol_panel(s2, upper, N, k + kb, kb1, A, lda, fl, info, badval));
    }
    MOPS_CHECK_HIP(hipEventRecord(e_join, s2));
    MOPS_CHECK_HIP(hipStreamWaitEvent(s, e_join, 0));
  }
  hipLaunchKernelGGL(k_fix_info, dim3(1), dim3(1), 0, s, info, fl.err);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

size_t potrs_ws_bytes(int flags, int N, int nrhs) {
  return 2 * trsm_ws_bytes(flags | F_LEFT, N, nrhs);
}

int potrs_d(hipStream_t s, int flags, int N, int nrhs, const double* A, long lda, double* B,
            long ldb) {
  // SolveCHOL (chol.go:147-158): UPPER solves U^T then U; LOWER solves L then
  // L^T; neither flag leaves B unchanged.
  // (both solves share the inverses of the diagonal blocks)
  if (flags & F_UPPER)
    return trsm_pair_d(s, F_UPPER | F_TRANSA, F_UPPER, N, nrhs, A, lda, B, ldb);
  if (flags & F_LOWER)
    return trsm_pair_d(s, F_LOWER, F_LOWER | F_TRANSA, N, nrhs, A, lda, B, ldb);
  return ST_OK;
}

}  // namespace mops
