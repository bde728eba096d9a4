// LU factorization with partial pivoting and the LU solve on the GPU.
//
// Replaces DecomposeLU / blockedLUpiv / unblockedLUpiv (lu.go:103-273),
// pivotIndex / applyPivots (pivot.go:55-61, 98-107) and SolveLU
// (lu.go:320-334).
//
// Pivot choice is the reference's: per column, the first row holding the
// strictly largest |a| from the diagonal down (ties -> lowest row; a NaN is
// never chosen unless it sits on the diagonal, where nothing replaces it),
// pivots 0-based and global, p[k] >= k. The arithmetic is right-looking
// (LAPACK dgetrf order) instead of the reference's left-looking variant 3;
// the factors agree to rounding and the pivot vectors are identical unless
// two candidates tie to within rounding.
//
// MI355X layout: the factorization runs on T = A^T (A in row-major order),
// so a row interchange moves two contiguous runs of memory, and every row of
// the panel is a contiguous run. A copy-transpose brings A in and out.
//
// Panel (columns k..k+kb, rows k..m-1), right-looking in 32-column leaves.
// A leaf is one kernel of G co-resident workgroups of 512 threads, one row
// per thread held in registers. Per column every workgroup finds its best
// candidate (wave shuffles) and publishes its key into exchange records
// that start out as a sentinel, then the candidate row; every workgroup
// polls the candidates' keys until all are there, agrees on the pivot,
// loads the pivot row and the old row k, exchanges the two rows in
// registers and updates the next column (no grid barrier per column); the
// rest of the rank-1 update runs in the next column while its keys travel
// (n = 16384: 118.7 -> 116.8 ms against the whole update before the
// publish). The interchanges of
// the leaf's rows in the other panel columns are applied once, at the end
// of the leaf, as a gather of the rows that moved. After each leaf the
// U block right of it is solved with the leaf's unit triangle (one small
// kernel) and the rest of the panel gets one GEMM; after the panel the
// trailing matrix gets the row interchanges, the TRSM of U12 (by the
// inverse of the panel's unit triangle) and the GEMM A22 -= L21*U12 on the
// matrix cores.
#include <climits>
#include <cmath>

#include "mops_lu.hpp"
#include "mops_solvers.hpp"

namespace mops {
namespace {

constexpr int LEAF_W = 32;    // leaf width (columns)
constexpr int LEAF_R = 512;   // rows per workgroup in a leaf (= threads)
constexpr int LEAF_GMAX = 64;
constexpr int LU_NBMAX = 512;  // widest panel (columns)

// Panel width of the GPU schedule for blocking nb and min(m, n) = mn.
// Every blocking gives the same factorization (same pivots, to rounding):
// a blocking below 256 is used as given; 0 (the reference's unblocked
// case) and 256 or more take wmax. The FP64 factorization uses LU_NBMAX
// (n = 16384: 131 ms against 138.5 ms with 256-column panels); the mixed
// one 256, as its FP32 trailing updates are cheap and wider panels only
// lengthen the panel chain (FP32 factorization 118 -> 125 ms at 512).
constexpr int LU_NB_MIXED = 256;
inline int lu_panel_width(int nb, int mn, int wmax) {
  const int w = (nb <= 0 || nb >= 256) ? wmax : nb;
  return w < mn ? w : mn;
}
constexpr int LEAF_SW = LEAF_W + 4;  // exchange slot: key, row, row data, perm, state
constexpr unsigned SPIN_LIMIT = 1u << 26;
// Panel state (row movement, inverse of the unit triangle) rotates over
// three slots: the rest update of step k may still read slot k mod 3 while
// the panel chain factors panel k+2; panel k+3 reuses the slot only after
// the bulk stream has finished step k.
constexpr int LU_NPAR = 3;

// ---- copy-transpose of a column-major tile (64x64 through LDS) ----
template <typename S, typename D>
__global__ __launch_bounds__(256) void k_transpose(int rows, int cols, const S* __restrict__ A,
                                                   long lda, D* __restrict__ B, long ldb) {
  __shared__ D t[64][65];
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int c = ty; c < 64; c += 4) {
    const int gr = r0 + tx, gc = c0 + c;
    if (gr < rows && gc < cols) t[c][tx] = (D)A[gr + (long)gc * lda];
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int gc = c0 + tx, gr = r0 + r;
    if (gr < rows && gc < cols) B[gc + (long)gr * ldb] = t[tx][r];
  }
}

// pivot candidates: larger key wins, ties -> lower row
__device__ __forceinline__ bool better(double v1, int r1, double v2, int r2) {
  return v1 > v2 || (v1 == v2 && r1 < r2);
}

__device__ __forceinline__ void st_agent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_agent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Exchange records start out as this signalling-NaN pattern (quiet bit
// clear), which no arithmetic result can equal: a reader that finds any
// other value in a word has the published one.
constexpr unsigned long long XSENT = 0x7FF5A5A5A5A5A5A5ull;
__device__ __forceinline__ bool is_xsent(double v) {
  return (unsigned long long)__double_as_longlong(v) == XSENT;
}
// One column's exchange records: LEAF_GMAX candidate slots [key, row, row
// data, perm, state], then the old row k [data, perm, state]. Columns use a
// ring of LEAF_RING such areas.
constexpr int LEAF_XAREA = LEAF_GMAX * LEAF_SW + LEAF_W + 2;
constexpr int LEAF_RING = 3;

struct LeafArgs {
  double* T;       // A^T, column-major: T[c + r*ldT] = A[r][c]
  long ldT;
  int m;           // rows of A
  int c0;          // first column of the leaf == first row it pivots
  int w;           // leaf width (<= LEAF_W)
  int pc0, pc1;    // columns of the enclosing panel (rows swapped in them too)
  int* piv;        // global pivots, piv[c0 + j]
  int* perm;       // perm[i] = original row now at position i
  double* xa;      // exchange records: a ring of 3 areas of LEAF_XAREA doubles
  int ring0;       // ring position of the leaf's first column
  double* stage;   // moved values of the other panel columns: [2 * LEAF_W * LU_NBMAX]
  unsigned* ctr;   // grid barrier counter (monotonic over the factorization)
  unsigned base;   // its value when this leaf starts
  int* err;        // set to 1 if a barrier wait timed out
};

// Grid barrier among the workgroups of one leaf launch (the leaf's final
// gather). Everything the workgroups exchange inside the leaf is
// written and read with agent-scope atomics (coherent without cache
// maintenance), so the barrier only has to
// order those accesses: every thread's stores are complete (waitcnt) before
// the workgroup barrier, the arrival is a relaxed atomic add, and the reads
// that follow the spin depend on it. No L2 write-back or invalidate is
// needed, which matters while other kernels fill the caches.
__device__ __forceinline__ void grid_sync(unsigned* ctr, unsigned target, int* err) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (++spins > SPIN_LIMIT) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
}

// True, every 1024 polls, once some wait of the factorization has timed out
// (the error word is set): later waits then give up at once instead of
// each running to the spin limit.
__device__ __forceinline__ bool leaf_failed(const int* err, unsigned spins) {
  return (spins & 1023u) == 0 &&
         __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// One DPP lane move of a double (both halves); lanes whose source is out of
// range keep their own value.
template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, ROW_MASK,
                                             BANK_MASK, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, ROW_MASK,
                                             BANK_MASK, false);
  return __hiloint2double(hi, lo);
}

// Wave maximum of v (never NaN) and the lowest lane that holds it. The
// maximum is a DPP scan (row_shr 1, 2, 3, 4, 8, then row_bcast 15 and 31),
// which ends in lane 63; the lanes that hold it are found with one ballot.
// Rows increase with the lane (and with the wave and the workgroup), so the
// first maximal lane is the reference's tie rule: the lowest row wins.
__device__ __forceinline__ int wave_first_max(double v, double& vmax) {
  double m = v;
  m = fmax(m, dpp_d<0x111, 0xf, 0xf>(m));  // row_shr:1
  m = fmax(m, dpp_d<0x112, 0xf, 0xf>(m));  // row_shr:2
  m = fmax(m, dpp_d<0x113, 0xf, 0xf>(m));  // row_shr:3
  m = fmax(m, dpp_d<0x114, 0xf, 0xe>(m));  // row_shr:4
  m = fmax(m, dpp_d<0x118, 0xf, 0xc>(m));  // row_shr:8
  m = fmax(m, dpp_d<0x142, 0xa, 0xf>(m));  // row_bcast:15
  m = fmax(m, dpp_d<0x143, 0xc, 0xf>(m));  // row_bcast:31
  const int lo = __builtin_amdgcn_readlane(__double2loint(m), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(m), 63);
  m = __hiloint2double(hi, lo);
  vmax = m;
  return __ffsll((long long)__ballot(v == m)) - 1;
}

// Pivoted LU of the leaf columns [c0, c0+w) over rows c0..m-1 (the column
// loop of unblockedLUpiv, lu.go:103-170, with pivotIndex, pivot.go:98-107):
// per column, the first strictly largest |a| from the diagonal down (a NaN
// below the diagonal never wins; a NaN on the diagonal is never replaced),
// the interchange of rows k and p, the column of L scaled by the reciprocal
// of the pivot and the rank-1 update of the leaf columns right of it.
// Thread t of workgroup g holds row c0 + 512 g + t of the leaf in
// registers, rotated: at step j, x[c] is column j + c (mod 32), so the
// current column is always x[0] and every register index is static while
// the column loop stays rolled. Rows travel between workgroups in this
// rotated form, which every thread shares. At the end, the rows that moved
// are gathered in the other columns of the panel [pc0, pc1) (rows outside
// the panel move later, all at once, from the permutation).
__global__ __launch_bounds__(LEAF_R) void k_getf2_leaf(LeafArgs a) {
  constexpr int W = LEAF_W, SW = LEAF_SW, NWV = LEAF_R / 64;
  __shared__ double xs[2][2 * (W + 1)];  // by column parity: the pivot row, then the old row k
  __shared__ int pinfo;               // the pivot of the current column
  __shared__ double wv[NWV];
  __shared__ int wr[NWV];
  __shared__ double wrow[NWV][W + 1];  // each wave's best row (data, perm)
  __shared__ double okl[W + 1];        // old row k, when this workgroup holds it
  __shared__ int pv[W];
  __shared__ int pair_dst[2 * W], pair_src[2 * W];

  const int G = gridDim.x, g = blockIdx.x, t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int c0 = a.c0, w = a.w;
  const int rg0 = c0 + g * LEAF_R;
  const int row = rg0 + t;
  const bool mine = row < a.m;
  double* T = a.T;
  const long ldT = a.ldT;

  double x[W];
  int pm = 0;
#pragma unroll
  for (int c = 0; c < W; ++c) x[c] = (mine && c < w) ? T[c0 + c + (long)row * ldT] : 0.0;
  if (mine) pm = a.perm[row];
  unsigned target = a.base;

  // The rank-1 update of column j reaches only column j + 1 in step 7; the
  // columns after it get it in step 3 of column j + 1, after the key of
  // column j + 1 is out, so it runs while the keys travel. dpend: this row
  // takes that pending update, with multiplier dl and the pivot row of the
  // previous column (xs[(j - 1) & 1]).
  bool dpend = false;
  double dl = 0.0;

#pragma unroll 1
  for (int j = 0; j < w; ++j) {
    const int k = c0 + j;
    // This column's records are in ring area (ring0 + j) mod 3. An area is
    // used again 3 columns later: each workgroup puts the sentinel back into
    // its own slot (workgroup 0 also into the old row) two columns ahead,
    // once every workgroup is done reading it.
    double* xaj = a.xa + (long)((a.ring0 + j) % LEAF_RING) * LEAF_XAREA;
    double* xareset = a.xa + (long)((a.ring0 + j + 2) % LEAF_RING) * LEAF_XAREA;
    // 1. candidate of each wave (x[0], column j, is final). Keys: |a|; a
    // NaN is -1 (never chosen) except on the diagonal, where it is +inf
    // (nothing replaces it); rows that are not candidates -2.
    double v = -2.0;
    if (mine && row >= k) {
      const double ax = fabs(x[0]);
      v = isnan(ax) ? (row == k ? INFINITY : -1.0) : ax;
    }
    double vmax;
    const int wl = wave_first_max(v, vmax);
    if (lane == 0) {
      wv[wave] = vmax;
      wr[wave] = vmax > -2.0 ? rg0 + 64 * wave + wl : INT_MAX;
    }
    __syncthreads();
    // 2. wave 0 publishes this workgroup's key and candidate row index (the
    // row data follows in step 5). The wait first orders every earlier
    // store of this wave (the previous column's data, the sentinels put
    // back into a ring area) before anything of this column.
    int bq = 0;  // the workgroup's winning wave (wave 0 only)
    if (wave == 0) {
      __builtin_amdgcn_s_waitcnt(0);
      double bv;
      bq = wave_first_max(lane < NWV ? wv[lane] : -3.0, bv);
      const int br = wr[bq];
      double* slot = xaj + (long)g * SW;
      if (lane == 0) {
        st_agent(&slot[1], (double)br);
        st_agent(&slot[0], bv);
      }
    }
    // 3. the pending update of column j - 1 on the columns after j
    if (dpend) {
      const double* pp = xs[(j - 1) & 1];  // pp[c + 1] belongs to x[c]
      const int nd = __builtin_amdgcn_readfirstlane(w - 1 - j);
#pragma unroll
      for (int c = 1; c < W; ++c) {
        // computed for every c, kept only for the pending columns (no branch)
        const double nv = fma(-dl, pp[c + 1], x[c]);
        x[c] = c <= nd ? nv : x[c];
      }
    }
    // 4. the final rows: each wave's candidate row and the old row k
    if (lane == wl) {
#pragma unroll
      for (int c = 0; c < W; ++c) wrow[wave][c] = x[c];
      wrow[wave][W] = (double)pm;
    }
    if (row == k) {
#pragma unroll
      for (int c = 0; c < W; ++c) okl[c] = x[c];
      okl[W] = (double)pm;
    }
    __syncthreads();
    // 5. the row data: the candidate row (wave 0) and the old row k (wave 1
    // of the workgroup that holds it). Data that happens to equal the
    // sentinel is published "escaped": a state word, stored after the data
    // stores completed, tells readers the data is complete.
    if (wave == 0) {
      double* slot = xaj + (long)g * SW;
      const double dv = lane <= W ? wrow[bq][lane] : 0.0;
      const bool esc = __ballot(lane <= W && is_xsent(dv)) != 0;
      if (lane <= W) st_agent(&slot[2 + lane], dv);
      if (esc) {
        __builtin_amdgcn_s_waitcnt(0);
        if (lane == 0) st_agent(&slot[SW - 1], 2.0);
      }
    }
    if (wave == 1 && k >= rg0 && k < rg0 + LEAF_R) {
      double* rk = xaj + (long)LEAF_GMAX * SW;
      const double dv = lane <= W ? okl[lane] : 0.0;
      const bool esc = __ballot(lane <= W && is_xsent(dv)) != 0;
      if (lane <= W) st_agent(&rk[lane], dv);
      if (esc) __builtin_amdgcn_s_waitcnt(0);
      if (lane == 0) st_agent(&rk[W + 1], esc ? 2.0 : 1.0);
    }
    // 6. wave 0 waits for every slot's key and row index (one slot per
    // lane), chooses the pivot (the first workgroup holding the largest key;
    // workgroups are in row order), then loads the pivot row and the old row
    // k until they hold no sentinel, or until their state word says they
    // were published escaped (then they are loaded once more). One store and
    // two load round trips per column, no grid barrier.
    double* xsj = xs[j & 1];
    if (wave == 0) {
      const bool has = lane < G;
      double kv = -3.0, rv = 0.0;
      if (has) {
        kv = ld_agent(&xaj[(long)lane * SW]);
        rv = ld_agent(&xaj[(long)lane * SW + 1]);
      }
      unsigned spins = 0;
      for (;;) {
        const bool pend = has && (is_xsent(kv) || is_xsent(rv));
        if (__ballot(pend) == 0) break;
        if (++spins > SPIN_LIMIT || leaf_failed(a.err, spins)) {
          if (lane == 0) __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        if (has && is_xsent(kv)) kv = ld_agent(&xaj[(long)lane * SW]);
        if (has && is_xsent(rv)) rv = ld_agent(&xaj[(long)lane * SW + 1]);
      }
      double bv;
      const int pq = wave_first_max(kv, bv);
      const double brw = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(rv), pq),
                                          __builtin_amdgcn_readlane(__double2loint(rv), pq));
      const bool cand = bv > -2.0;  // some row >= k exists (always, as k < m)
      const double* pslot = xaj + (long)pq * SW;
      const double* rk = xaj + (long)LEAF_GMAX * SW;
      double pd = 0.0, od = 0.0, pst = 0.0, ost = 0.0;
      if (lane <= W) {
        if (cand) pd = ld_agent(&pslot[2 + lane]);
        od = ld_agent(&rk[lane]);
      }
      if (lane == 0) {
        if (cand) pst = ld_agent(&pslot[SW - 1]);
        ost = ld_agent(&rk[W + 1]);
      }
      bool pesc = false, oesc = false;  // escaped and complete
      for (;;) {
        pesc = __builtin_amdgcn_readfirstlane(__double2hiint(pst)) == __double2hiint(2.0) &&
               __builtin_amdgcn_readfirstlane(__double2loint(pst)) == 0;
        oesc = __builtin_amdgcn_readfirstlane(__double2hiint(ost)) == __double2hiint(2.0) &&
               __builtin_amdgcn_readfirstlane(__double2loint(ost)) == 0;
        const bool pend =
            lane <= W && ((cand && !pesc && is_xsent(pd)) || (!oesc && is_xsent(od)));
        if (__ballot(pend) == 0) break;
        if (++spins > SPIN_LIMIT || leaf_failed(a.err, spins)) {
          if (lane == 0) __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        if (lane <= W && cand && !pesc && is_xsent(pd)) pd = ld_agent(&pslot[2 + lane]);
        if (lane <= W && !oesc && is_xsent(od)) od = ld_agent(&rk[lane]);
        if (lane == 0 && cand && !pesc) pst = ld_agent(&pslot[SW - 1]);
        if (lane == 0 && !oesc) ost = ld_agent(&rk[W + 1]);
      }
      // escaped data may have been loaded before its last stores landed:
      // load it again, now that the state word has been seen
      if (lane <= W && cand && pesc) pd = ld_agent(&pslot[2 + lane]);
      if (lane <= W && oesc) od = ld_agent(&rk[lane]);
      if (lane <= W) {
        xsj[lane] = pd;
        xsj[W + 1 + lane] = od;
      }
      if (lane == 0) pinfo = cand ? (int)brw : k;
      // every workgroup has published this column's key, so all of them are
      // done reading the previous column: its area gets the sentinel back
      const double sv = __longlong_as_double((long long)XSENT);
      if (lane < SW) st_agent(&xareset[(long)g * SW + lane], sv);
      if (g == 0 && lane < W + 2) st_agent(&xareset[(long)LEAF_GMAX * SW + lane], sv);
    }
    __syncthreads();
    // 7. the pivot, the interchange in registers, the column of L and the
    // update of column j + 1 (the rest waits for step 3 of the next column)
    const int p = pinfo;
    if (t == 0) {
      pv[j] = p;
      if (g == 0) a.piv[k] = p;
    }
    const double* prow = xsj;          // the pivot row; prow[W] = its perm
    const double* orow = xsj + W + 1;  // the old row k
    if (p != k && (row == p || row == k)) {
      const double* src = row == p ? orow : prow;
#pragma unroll
      for (int c = 0; c < W; ++c) x[c] = src[c];
      pm = (int)src[W];
    }
    dpend = mine && row > k;
    if (dpend) {
      const double l = x[0] * (1.0 / prow[0]);
      x[0] = l;
      if (j + 1 < w) x[1] = fma(-l, prow[1], x[1]);
      dl = l;
    }
    // rotate: column j moves to the back
    const double x0 = x[0];
#pragma unroll
    for (int c = 0; c < W - 1; ++c) x[c] = x[c + 1];
    x[W - 1] = x0;
  }
  // after w rotations x[c] holds column (c + w) mod 32
  if (mine) {
#pragma unroll
    for (int c = 0; c < W; ++c) {
      const int col = (c + w) & (W - 1);
      if (col < w) T[c0 + col + (long)row * ldT] = x[c];
    }
    a.perm[row] = pm;
  }
  // The other panel columns: position d ends up with the row that was at
  // src(d), found by running the leaf's interchanges backwards.
  const int ncol = (a.pc1 - a.pc0) - w;
  if (ncol <= 0) return;
  __syncthreads();
  if (t < 2 * w) {
    const int d = (t & 1) ? pv[t >> 1] : c0 + (t >> 1);
    int y = d;
    for (int jj = w - 1; jj >= 0; --jj) {
      if (y == c0 + jj)
        y = pv[jj];
      else if (y == pv[jj])
        y = c0 + jj;
    }
    pair_dst[t] = d;
    pair_src[t] = y;
  }
  __syncthreads();
  // stage every moved value (agent-scope atomics, visible to all
  // workgroups), wait for all of them, then write the destinations
  const int items = 2 * w * ncol;
  const int stride = G * LEAF_R;
  for (int i2 = g * LEAF_R + t; i2 < items; i2 += stride) {
    const int pq = i2 / ncol;
    int c = a.pc0 + i2 % ncol;
    if (c >= c0) c += w;
    if (pair_src[pq] != pair_dst[pq]) st_agent(&a.stage[i2], T[c + (long)pair_src[pq] * ldT]);
  }
  target += (unsigned)G;
  grid_sync(a.ctr, target, a.err);
  for (int i2 = g * LEAF_R + t; i2 < items; i2 += stride) {
    const int pq = i2 / ncol;
    int c = a.pc0 + i2 % ncol;
    if (c >= c0) c += w;
    if (pair_src[pq] != pair_dst[pq]) T[c + (long)pair_dst[pq] * ldT] = ld_agent(&a.stage[i2]);
  }
}

// U block right of a leaf: for each column q of A in [c0+w, c1), solve
// L11 u = a (L11 unit lower, the leaf's w x w diagonal block; a = A[c0:c0+w,
// q]) in place. Thread = column; T layout (A[r][q] = T[q + r*ldT]).
__global__ __launch_bounds__(256) void k_unit_lower_solve(double* T, long ldT, int c0, int w,
                                                          int c1) {
  constexpr int W = LEAF_W;
  __shared__ double Ls[W][W + 1];  // Ls[i][j] = L11[i][j]
  for (int idx = threadIdx.x; idx < W * W; idx += 256) {
    const int i = idx / W, jj = idx % W;
    Ls[i][jj] = (i < w && jj < i) ? T[(c0 + jj) + (long)(c0 + i) * ldT] : 0.0;
  }
  __syncthreads();
  const int q = c0 + w + blockIdx.x * 256 + threadIdx.x;
  if (q >= c1) return;
  double u[W];
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const double v = T[q + (long)(c0 + min(i, w - 1)) * ldT];  // in bounds for any i
    u[i] = i < w ? v : 0.0;
  }
#pragma unroll
  for (int i = 0; i < W; ++i) {
    // make this step's LDS reads wait for u[i]: hoisting all 496 of them
    // ahead of the arithmetic would need twice the register file
    typedef __attribute__((address_space(3))) const double lds_double;
    lds_double* col = (lds_double*)&Ls[0][i];
    asm volatile("" : "+v"(col) : "v"(u[i]));
#pragma unroll
    for (int i2 = i + 1; i2 < W; ++i2) u[i2] = fma(-col[i2 * (W + 1)], u[i], u[i2]);
  }
#pragma unroll
  for (int i = 0; i < W; ++i)
    if (i < w) T[q + (long)(c0 + i) * ldT] = u[i];
}

__global__ void k_fill_int(int* p, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_iota(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}
__global__ void k_zero_u32(unsigned* p) { *p = 0u; }

__global__ void k_piv_to_i64(const int* p, long long* q, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    q[i] = p[i];
}

// Net row movement of one panel: inv0[original row] = position before the
// panel (rows k..m-1), then pairs (dst, src) for every position whose row
// changed; pairs[0] holds the count.
__global__ void k_inv_perm(const int* perm, int k, int m, int* inv0) {
  for (int r = k + blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x)
    inv0[perm[r]] = r;
}
__global__ void k_zero1(int* p) { p[0] = 0; }
__global__ void k_fill_u64_lu(unsigned long long* p, long n, unsigned long long v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_perm_pairs(const int* perm, const int* inv0, int k, int m, int* pairs,
                             int maxpairs) {
  for (int r = k + blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x) {
    const int src = inv0[perm[r]];
    if (src != r) {
      const int q = atomicAdd(&pairs[0], 1);
      if (q < maxpairs) {
        pairs[1 + 2 * q] = r;
        pairs[2 + 2 * q] = src;
      }
    }
  }
}

// Moves the rows listed in pairs (row dst of A receives old row src) in the
// columns [c_lo, c_hi) of A, i.e. T-rows: every workgroup takes AP_ROWS of
// them, reads all source values into LDS, then writes the destinations.
// AP_ROWS = 16, or 8 for panels wider than 256 columns (their up to 1024
// pairs keep the staging within 64 KiB).
template <typename D, int AP_ROWS>
__global__ __launch_bounds__(256) void k_apply_pairs(D* T, long ldT, const int* pairs,
                                                     int maxpairs, int c_lo, int c_hi) {
  extern __shared__ double vals_raw[];  // [maxpairs][AP_ROWS]
  D* vals = reinterpret_cast<D*>(vals_raw);
  const int np = min(pairs[0], maxpairs);
  const int cbase = c_lo + blockIdx.x * AP_ROWS;
  const int nc = min(AP_ROWS, c_hi - cbase);
  if (nc <= 0 || np == 0) return;
  for (int idx = threadIdx.x; idx < np * AP_ROWS; idx += 256) {
    const int q = idx / AP_ROWS, c = idx % AP_ROWS;
    if (c < nc) vals[idx] = T[cbase + c + (long)pairs[2 + 2 * q] * ldT];
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < np * AP_ROWS; idx += 256) {
    const int q = idx / AP_ROWS, c = idx % AP_ROWS;
    if (c < nc) T[cbase + c + (long)pairs[1 + 2 * q] * ldT] = vals[idx];
  }
}

// B_out[i][c] = B_in[perm[i]][c]  (gather rows of a column-major n x nrhs)
__global__ void k_gather_rows(int n, int nrhs, const int* perm, const double* B, long ldb,
                              double* out, long ldo) {
  const long total = (long)n * nrhs;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % n), c = (int)(t / n);
    out[i + (long)c * ldo] = B[perm[i] + (long)c * ldb];
  }
}

// out[perm[i]][c] = B[i][c]
__global__ void k_scatter_rows(int n, int nrhs, const int* perm, const double* B, long ldb,
                               double* out, long ldo) {
  const long total = (long)n * nrhs;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % n), c = (int)(t / n);
    out[perm[i] + (long)c * ldo] = B[i + (long)c * ldb];
  }
}

// permutation from a pivot sequence: perm[i] = row that ends at position i
// after applying the interchanges (k, piv[k]) for k = 0..npiv-1 in order
// (applyPivots, pivot.go:55-61). One workgroup, the index array in LDS when
// it fits, else in global memory.
__global__ __launch_bounds__(1024) void k_perm_from_piv(const long long* piv, int npiv, int n,
                                                       int* perm) {
  extern __shared__ int idx[];
  const bool in_lds = n <= 16384;
  int* a = in_lds ? idx : perm;
  for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < npiv; ++k) {
      const long long p = piv[k];
      if (p > 0 && p != k && p < n) {
        const int x = a[k];
        a[k] = a[p];
        a[p] = x;
      }
    }
  }
  __syncthreads();
  if (in_lds)
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = a[i];
}

inline int grid1(long total) {
  long b = (total + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

struct LuCtx {
  double* T;
  long ldT;
  int m, n;
  int* piv;
  int* perm;
  int* inv0;
  int* pairs[LU_NPAR];
  int maxpairs;
  double* linv[LU_NPAR];  // inverse of the unit upper T11 of a panel (nb x nb)
  double* tmp[2];   // U12 staging for the two streams
  double* xa;          // leaf exchange records (ring of LEAF_RING areas)
  unsigned col_count;  // leaf columns factored so far (ring position)
  double* stage;
  unsigned* ctr;
  unsigned sync_base;  // barrier arrivals queued so far on ctr
  int* err;
};

// Whether G workgroups of the leaf kernel fit on the current device at once
// (with nothing else running), cached per device.
int leaf_capacity_ok(int G) {
  static int cap[16] = {0};
  int d = 0;
  MOPS_CHECK_HIP(hipGetDevice(&d));
  if (d < 0 || d >= 16) d = 0;
  if (cap[d] == 0) {
    int per_cu = 0, ncu = 0;
    MOPS_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_getf2_leaf, LEAF_R, 0));
    MOPS_CHECK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, d));
    cap[d] = per_cu * ncu > 0 ? per_cu * ncu : -1;
  }
  if (cap[d] < G) {
    set_error("getrf: %d leaf workgroups do not fit on the device at once (capacity %d)", G,
              cap[d]);
    return ST_HIP;
  }
  return ST_OK;
}

int leaf(hipStream_t s, LuCtx& c, int c0, int w, int pc0, int pc1) {
  const int M = c.m - c0;
  const int G = ceil_div(M, LEAF_R);
  if (G > LEAF_GMAX) {
    set_error("getrf: more than %d rows are not supported", LEAF_GMAX * LEAF_R);
    return ST_BADARG;
  }
  if (pc1 - pc0 > LU_NBMAX) {
    set_error("getrf: panel wider than %d columns", LU_NBMAX);
    return ST_BADARG;
  }
  LeafArgs a;
  a.T = c.T;
  a.ldT = c.ldT;
  a.m = c.m;
  a.c0 = c0;
  a.w = w;
  a.pc0 = pc0;
  a.pc1 = pc1;
  a.piv = c.piv;
  a.perm = c.perm;
  a.xa = c.xa;
  a.ring0 = (int)(c.col_count % LEAF_RING);
  a.stage = c.stage;
  a.ctr = c.ctr;
  a.base = c.sync_base;
  a.err = c.err;
  // The leaf's workgroups wait for each other (exchange records, one grid
  // barrier), so all G of them must be resident at once. G <= LEAF_GMAX is
  // far below the device's capacity for this kernel (checked once per
  // process against the occupancy the kernel's resources allow), the
  // kernels running beside a leaf (GEMM tiles) never wait, so every
  // workgroup of the leaf is dispatched as they drain, and every wait is
  // bounded: a plain launch, which unlike a cooperative one does not
  // serialize with the trailing updates on the other streams.
  MOPS_TRY(leaf_capacity_ok(G));
  hipLaunchKernelGGL(k_getf2_leaf, dim3(G), dim3(LEAF_R), 0, s, a);
  MOPS_CHECK_HIP(hipGetLastError());
  // one grid barrier, for the gather, when the panel is wider than the leaf
  const int nsync = (pc1 - pc0) - w > 0 ? 1 : 0;
  c.sync_base += (unsigned)nsync * (unsigned)G;
  c.col_count += (unsigned)w;
  return ST_OK;
}

// Panel columns [pc0, pc1), rows pc0..m-1: leaves of LEAF_W columns; after
// each, the U block right of it (unit lower solve) and one GEMM for the
// rest of the panel.
int panel_factor(hipStream_t s, LuCtx& c, int pc0, int pc1) {
  const long ld = c.ldT;
  double* T = c.T;
  for (int c0 = pc0; c0 < pc1; c0 += LEAF_W) {
    const int w = min(LEAF_W, pc1 - c0);
    MOPS_TRY(leaf(s, c, c0, w, pc0, pc1));
    const int right = pc1 - (c0 + w);
    if (right <= 0) break;
    hipLaunchKernelGGL(k_unit_lower_solve, dim3(ceil_div(right, 256)), dim3(256), 0, s, T, ld, c0,
                       w, pc1);
    MOPS_CHECK_HIP(hipGetLastError());
    const int below = c.m - (c0 + w);
    if (below > 0)
      MOPS_TRY(gemm_d(s, false, false, right, below, w, -1.0, T + (c0 + w) + (long)c0 * ld, ld,
                      T + c0 + (long)(c0 + w) * ld, ld, 1.0, T + (c0 + w) + (long)(c0 + w) * ld,
                      ld));
  }
  return ST_OK;
}

// Panel [k, k+kb): factor it, derive its row movement (pairs[par]) and the
// inverse of its unit upper T11 (linv[par]).
int panel(hipStream_t s, LuCtx& c, int k, int kb, int par) {
  const int rows = c.m - k;
  hipLaunchKernelGGL(k_inv_perm, dim3(grid1(rows)), dim3(256), 0, s, c.perm, k, c.m, c.inv0);
  MOPS_CHECK_HIP(hipGetLastError());
  trace_mark_begin(s, TR_PANEL);
  MOPS_TRY(panel_factor(s, c, k, k + kb));
  trace_mark_end(s, TR_PANEL, 0.0);
  hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, s, c.pairs[par]);
  hipLaunchKernelGGL(k_perm_pairs, dim3(grid1(rows)), dim3(256), 0, s, c.perm, c.inv0, k, c.m,
                     c.pairs[par], c.maxpairs);
  MOPS_CHECK_HIP(hipGetLastError());
  const size_t mark = ws_mark();
  const int st = trtri_d(s, false, true, kb, c.T + k + (long)k * c.ldT, c.ldT, c.linv[par], kb);
  ws_release(mark);
  return st;
}

int apply_rows(hipStream_t s, LuCtx& c, int par, int c_lo, int c_hi) {
  if (c_hi <= c_lo) return ST_OK;
  if (c.maxpairs <= 512) {
    const size_t shm = (size_t)c.maxpairs * 16 * sizeof(double);
    hipLaunchKernelGGL((k_apply_pairs<double, 16>), dim3(ceil_div(c_hi - c_lo, 16)), dim3(256), shm,
                     s, c.T, c.ldT, c.pairs[par], c.maxpairs, c_lo, c_hi);
  } else {
    const size_t shm = (size_t)c.maxpairs * 8 * sizeof(double);
    hipLaunchKernelGGL((k_apply_pairs<double, 8>), dim3(ceil_div(c_hi - c_lo, 8)), dim3(256), shm,
                     s, c.T, c.ldT, c.pairs[par], c.maxpairs, c_lo, c_hi);
  }
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

// Columns [j0, j1) of A after panel [k, k+kb): row interchanges, U12 = L11^-1
// A12 (by the inverse), and the update A22 -= L21 U12 of rows k+kb..m-1.
int update_cols(hipStream_t s, LuCtx& c, int k, int kb, int par, int j0, int j1, double* tmp,
                bool gemm_trace) {
  if (j1 <= j0) return ST_OK;
  MOPS_TRY(apply_rows(s, c, par, j0, j1));
  const long ld = c.ldT;
  double* T = c.T;
  const int nc = j1 - j0;
  double* U = T + j0 + (long)k * ld;  // T rows j0..j1, T columns k..k+kb
  MOPS_TRY(copy_tile_d(s, nc, kb, U, ld, tmp, nc));
  MOPS_TRY(gemm_d(s, false, false, nc, kb, kb, 1.0, tmp, nc, c.linv[par], kb, 0.0, U, ld, GM_FULL,
                  TRI_B_UPPER));
  const int below = c.m - k - kb;
  if (below <= 0) return ST_OK;
  if (gemm_trace) trace_mark_begin(s, TR_GEMM_UPDATE);
  MOPS_TRY(gemm_d(s, false, false, nc, below, kb, -1.0, U, ld, T + k + (long)(k + kb) * ld, ld,
                  1.0, T + j0 + (long)(k + kb) * ld, ld));
  if (gemm_trace) trace_mark_end(s, TR_GEMM_UPDATE, 2.0 * nc * (double)below * kb);
  return ST_OK;
}

}  // namespace

namespace {

// Moves the trailing-update work of a factorization from s to the bulk
// stream (sr), after everything already queued on s; join_bulk makes s
// wait for it again. Without a bulk stream sr stays s.
int fork_bulk(hipStream_t s, hipStream_t& sr) {
  hipStream_t b = bulk_stream();
  hipEvent_t e = sync_event(4);
  if (!b || !e) return ST_OK;
  MOPS_CHECK_HIP(hipEventRecord(e, s));
  MOPS_CHECK_HIP(hipStreamWaitEvent(b, e, 0));
  sr = b;
  return ST_OK;
}

int join_bulk(hipStream_t s, hipStream_t sr) {
  if (sr == s) return ST_OK;
  hipEvent_t e = sync_event(5);
  if (!e) {
    set_error("getrf: no event to join the bulk stream");
    return ST_HIP;
  }
  MOPS_CHECK_HIP(hipEventRecord(e, sr));
  MOPS_CHECK_HIP(hipStreamWaitEvent(s, e, 0));
  return ST_OK;
}

}  // namespace

size_t getrf_ws_bytes(int m, int n, int nb) {
  if (m <= 0 || n <= 0) return 0;
  const int mn = m < n ? m : n;
  nb = lu_panel_width(nb, mn, LU_NBMAX);
  const int G = ceil_div(m, LEAF_R);
  size_t b = (size_t)m * n * sizeof(double);                   // T
  b += (size_t)LEAF_RING * LEAF_XAREA * sizeof(double) + 4096;   // exchange
  b += (size_t)2 * LEAF_W * LU_NBMAX * sizeof(double) + 256;        // gather staging
  b += (size_t)(mn + 3 * m + 64) * sizeof(int);                // piv, perm, inv0, ctr, err
  b += (size_t)LU_NPAR * (4 * nb + 2) * sizeof(int);           // pairs
  b += (size_t)LU_NPAR * nb * nb * sizeof(double);             // linv
  b += (size_t)2 * n * nb * sizeof(double);                    // tmp
  b += trtri_ws_bytes(nb) + trsm_ws_bytes(F_RIGHT, nb, nb) + 32 * 256;
  return b;
}

int getrf_d(hipStream_t s, int m, int n, int nb, double* A, long lda, long long* ipiv,
            int* perm_out, int* err_flag) {
  if (m <= 0 || n <= 0) return ST_OK;
  const int mn = m < n ? m : n;
  // DecomposeLU: unblocked iff min(m,n) <= nb || nb == 0 (lu.go:267). Every
  // blocking gives the same factorization (same pivots, to rounding), so the
  // GPU schedule picks its own panel width (lu_panel_width).
  nb = lu_panel_width(nb, mn, LU_NBMAX);
  const int G = ceil_div(m, LEAF_R);
  if (G > LEAF_GMAX) {
    set_error("getrf: more than %d rows are not supported", LEAF_GMAX * LEAF_R);
    return ST_BADARG;
  }
  LuCtx c;
  c.m = m;
  c.n = n;
  c.ldT = n;
  c.maxpairs = 2 * nb;
  c.T = (double*)ws_carve((size_t)m * n * sizeof(double));
  c.xa = (double*)ws_carve((size_t)LEAF_RING * LEAF_XAREA * sizeof(double));
  c.col_count = 0;
  c.stage = (double*)ws_carve((size_t)2 * LEAF_W * LU_NBMAX * sizeof(double));
  c.piv = (int*)ws_carve((size_t)mn * sizeof(int));
  c.perm = (int*)ws_carve((size_t)m * sizeof(int));
  c.inv0 = (int*)ws_carve((size_t)m * sizeof(int));
  bool slots_ok = true;
  for (int q = 0; q < LU_NPAR; ++q) {
    c.pairs[q] = (int*)ws_carve((size_t)(2 * c.maxpairs + 1) * sizeof(int));
    c.linv[q] = (double*)ws_carve((size_t)nb * nb * sizeof(double));
    slots_ok = slots_ok && c.pairs[q] && c.linv[q];
  }
  c.tmp[0] = (double*)ws_carve((size_t)n * nb * sizeof(double));
  c.tmp[1] = (double*)ws_carve((size_t)n * nb * sizeof(double));
  c.ctr = (unsigned*)ws_carve(256);
  c.err = err_flag ? err_flag : (int*)ws_carve(256);
  if (!c.T || !c.xa || !c.stage || !c.piv || !c.perm || !c.inv0 || !slots_ok ||
      !c.tmp[0] || !c.tmp[1] || !c.ctr || !c.err) {
    set_error("getrf: workspace too small");
    return ST_NOMEM;
  }
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, c.err, 1, 0);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, c.ctr);
  hipLaunchKernelGGL(k_fill_u64_lu, dim3(grid1(LEAF_RING * LEAF_XAREA)), dim3(256), 0, s,
                     (unsigned long long*)c.xa, (long)LEAF_RING * LEAF_XAREA, XSENT);
  c.sync_base = 0;
  hipLaunchKernelGGL(k_iota, dim3(grid1(m)), dim3(256), 0, s, c.perm, m);
  // T = A^T
  hipLaunchKernelGGL((k_transpose<double, double>), dim3(ceil_div(m, 64), ceil_div(n, 64)),
                     dim3(256), 0, s, m, n, A, lda, c.T, c.ldT);
  MOPS_CHECK_HIP(hipGetLastError());

  // Look-ahead schedule. Stream s2 factors panel k+1 right after it has
  // brought that panel's columns up to date with panel k ("next"), while
  // the caller's stream s applies panel k to everything else ("rest"):
  // the earlier columns (row interchanges only) and the columns right of
  // panel k+1 (interchanges, U12, trailing GEMM).
  hipStream_t s2 = aux_stream();
  hipEvent_t e_fork = sync_event(0), e_pan = sync_event(1), e_rest = sync_event(2),
             e_join = sync_event(3);
  if (!s2 || !e_fork || !e_pan || !e_rest || !e_join) {
    set_error("getrf: could not create the look-ahead stream");
    return ST_HIP;
  }
  // The rest updates go to the bulk stream, which leaves a few CUs to the
  // panel chain (its kernels are small and latency-bound: without free CUs
  // each of them waits for GEMM tiles to drain before it starts).
  hipStream_t sr = s;
  MOPS_TRY(fork_bulk(s, sr));
  MOPS_CHECK_HIP(hipEventRecord(e_fork, s));
  MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_fork, 0));
  MOPS_TRY(panel(s2, c, 0, min(nb, mn), 0));
  MOPS_CHECK_HIP(hipEventRecord(e_pan, s2));
  bool have_rest = false;
  int par = 0;
  for (int k = 0; k < mn; k += nb, par = (par + 1) % LU_NPAR) {
    const int kb = min(nb, mn - k);
    const int kb1 = min(nb, mn - k - kb);  // width of the next panel (0 at the end)
    const int j_next = k + kb, j_rest = k + kb + kb1;
    // s2 brings block column k+1 up to date only after step k-1 has
    if (have_rest) MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_rest, 0));
    MOPS_CHECK_HIP(hipStreamWaitEvent(sr, e_pan, 0));
    have_rest = false;
    if (j_rest < n) {
      // block column k+2 first: the next step's block column update needs
      // only it, so the panel chain does not wait for the whole rest
      const int j_far = min(n, j_rest + nb);
      MOPS_TRY(update_cols(sr, c, k, kb, par, j_rest, j_far, c.tmp[0], true));
      MOPS_CHECK_HIP(hipEventRecord(e_rest, sr));
      have_rest = true;
      MOPS_TRY(update_cols(sr, c, k, kb, par, j_far, n, c.tmp[0] + (long)(j_far - j_rest) * kb,
                           true));
    }
    MOPS_TRY(apply_rows(sr, c, par, 0, k));
    if (kb1 > 0) {
      MOPS_TRY(update_cols(s2, c, k, kb, par, j_next, j_next + kb1, c.tmp[1], false));
      MOPS_TRY(panel(s2, c, j_next, kb1, (par + 1) % LU_NPAR));
      MOPS_CHECK_HIP(hipEventRecord(e_pan, s2));
    }
  }
  MOPS_CHECK_HIP(hipEventRecord(e_join, s2));
  MOPS_CHECK_HIP(hipStreamWaitEvent(s, e_join, 0));
  MOPS_TRY(join_bulk(s, sr));

  // A = T^T
  hipLaunchKernelGGL((k_transpose<double, double>), dim3(ceil_div(n, 64), ceil_div(m, 64)),
                     dim3(256), 0, s, n, m, c.T, c.ldT, A, lda);
  MOPS_CHECK_HIP(hipGetLastError());
  if (ipiv) {
    hipLaunchKernelGGL(k_piv_to_i64, dim3(grid1(mn)), dim3(256), 0, s, c.piv, ipiv, mn);
    MOPS_CHECK_HIP(hipGetLastError());
  }
  if (perm_out) MOPS_CHECK_HIP(hipMemcpyAsync(perm_out, c.perm, (size_t)m * sizeof(int),
                                              hipMemcpyDeviceToDevice, s));
  return ST_OK;
}

// ---------------------------------------------------------------------------
// Mixed precision: the trailing matrix in FP32, the panels in FP64.
//
// T32 = A^T in FP32 holds the matrix between panels; every trailing update
// (row interchanges, U12 by the inverse of the panel's unit triangle, and
// A22 -= L21 U12) runs on FP32 data with the FP32 matrix cores, at twice the
// FP64 rate and half the bytes. Each panel is widened to FP64 and factored
// by the same code as getrf_d (pivot search, interchanges and the panel's
// own updates in FP64), then narrowed back. The factors are returned in
// FP64 (exactly the FP32 values) for the refinement solves.

namespace {

struct MixedCtx {
  bool bf16;         // trailing products A22 -= L21 U12 on the BF16 matrix cores
  unsigned* amax[2];  // max |U12| words of the BF16 products, one per stream
  float* T32;        // n x n, T32[c + r*n] = A[r][c]
  float* linv32[LU_NPAR];  // FP32 copies of the panels' unit-triangle inverses
  float* tmp32[2];   // U12 staging for the two streams
  double* P64;       // one FP64 panel: nb x n, P64[(c-k) + (r-k)*nb]
};

int apply_rows32(hipStream_t s, LuCtx& c, MixedCtx& x, int par, int c_lo, int c_hi) {
  if (c_hi <= c_lo) return ST_OK;
  if (c.maxpairs <= 512) {
    const size_t shm = (size_t)c.maxpairs * 16 * sizeof(float);
    hipLaunchKernelGGL((k_apply_pairs<float, 16>), dim3(ceil_div(c_hi - c_lo, 16)), dim3(256), shm,
                     s, x.T32, (long)c.n, c.pairs[par], c.maxpairs, c_lo, c_hi);
  } else {
    const size_t shm = (size_t)c.maxpairs * 8 * sizeof(float);
    hipLaunchKernelGGL((k_apply_pairs<float, 8>), dim3(ceil_div(c_hi - c_lo, 8)), dim3(256), shm,
                     s, x.T32, (long)c.n, c.pairs[par], c.maxpairs, c_lo, c_hi);
  }
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

int update_cols32(hipStream_t s, LuCtx& c, MixedCtx& x, int k, int kb, int par, int j0, int j1,
                  float* tmp, bool gemm_trace) {
  if (j1 <= j0) return ST_OK;
  MOPS_TRY(apply_rows32(s, c, x, par, j0, j1));
  const long ld = c.n;
  float* T = x.T32;
  const int nc = j1 - j0;
  float* U = T + j0 + (long)k * ld;
  MOPS_TRY(copy_tile_s(s, nc, kb, U, ld, tmp, nc));
  MOPS_TRY(gemm_s(s, false, false, nc, kb, kb, 1.0f, tmp, nc, x.linv32[par], kb, 0.0f, U, ld,
                  GM_FULL, TRI_B_UPPER));
  const int below = c.m - k - kb;
  if (below <= 0) return ST_OK;
  if (gemm_trace) trace_mark_begin(s, TR_GEMM_UPDATE);
  if (x.bf16)
    MOPS_TRY(gemm_bf16_s(s, nc, below, kb, -1.0f, U, ld, T + k + (long)(k + kb) * ld, ld,
                         T + j0 + (long)(k + kb) * ld, ld, x.amax[gemm_trace ? 0 : 1]));
  else
    MOPS_TRY(gemm_s(s, false, false, nc, below, kb, -1.0f, U, ld, T + k + (long)(k + kb) * ld,
                    ld, 1.0f, T + j0 + (long)(k + kb) * ld, ld));
  if (gemm_trace) trace_mark_end(s, TR_GEMM_UPDATE, 2.0 * nc * (double)below * kb);
  return ST_OK;
}

// Panel [k, k+kb) of T32: widen, factor in FP64 (panel() on a context whose
// T is the FP64 panel), narrow back; FP32 copy of the triangle's inverse.
int panel_mixed(hipStream_t s, LuCtx& c, MixedCtx& x, int k, int kb, int par) {
  const long ld = c.n;
  const int rows = c.m - k;
  MOPS_TRY(convert_s2d(s, kb, rows, x.T32 + k + (long)k * ld, ld, x.P64, kb));
  LuCtx pc = c;
  pc.T = x.P64 - k - (long)k * kb;  // only columns k..k+kb, rows k..m are touched
  pc.ldT = kb;
  MOPS_TRY(panel(s, pc, k, kb, par));
  c.sync_base = pc.sync_base;
  c.col_count = pc.col_count;
  MOPS_TRY(convert_d2s(s, kb, kb, c.linv[par], kb, x.linv32[par], kb));
  return convert_d2s(s, kb, rows, x.P64, kb, x.T32 + k + (long)k * ld, ld);
}

}  // namespace

size_t getrf_mixed_ws_bytes(int n, int nb) {
  if (n <= 0) return 0;
  nb = lu_panel_width(nb, n, LU_NB_MIXED);
  const int G = ceil_div(n, LEAF_R);
  size_t b = (size_t)n * n * sizeof(float);                    // T32
  b += (size_t)nb * n * sizeof(double);                        // P64
  b += (size_t)LEAF_RING * LEAF_XAREA * sizeof(double) + 4096;   // exchange
  b += (size_t)2 * LEAF_W * LU_NBMAX * sizeof(double) + 256;        // gather staging
  b += (size_t)(4 * n + 64) * sizeof(int);                     // piv, perm, inv0, ctr, err
  b += (size_t)LU_NPAR * (4 * nb + 2) * sizeof(int);           // pairs
  b += (size_t)LU_NPAR * nb * nb * (sizeof(double) + sizeof(float));  // linv
  b += (size_t)2 * n * nb * sizeof(float);                     // tmp32
  b += trtri_ws_bytes(nb) + 34 * 256;
  return b;
}

int getrf_mixed_d(hipStream_t s, int n, int nb, const double* A, long lda, double* LU, long ldlu,
                  long long* ipiv, int* perm_out, int* err_flag, bool bf16) {
  if (n <= 0) return ST_OK;
  nb = lu_panel_width(nb, n, LU_NB_MIXED);
  const int G = ceil_div(n, LEAF_R);
  if (G > LEAF_GMAX) {
    set_error("getrf: more than %d rows are not supported", LEAF_GMAX * LEAF_R);
    return ST_BADARG;
  }
  LuCtx c;
  MixedCtx x;
  c.m = n;
  c.n = n;
  c.ldT = n;
  c.T = nullptr;
  c.maxpairs = 2 * nb;
  x.T32 = (float*)ws_carve((size_t)n * n * sizeof(float));
  x.P64 = (double*)ws_carve((size_t)nb * n * sizeof(double));
  c.xa = (double*)ws_carve((size_t)LEAF_RING * LEAF_XAREA * sizeof(double));
  c.col_count = 0;
  c.stage = (double*)ws_carve((size_t)2 * LEAF_W * LU_NBMAX * sizeof(double));
  c.piv = (int*)ws_carve((size_t)n * sizeof(int));
  c.perm = (int*)ws_carve((size_t)n * sizeof(int));
  c.inv0 = (int*)ws_carve((size_t)n * sizeof(int));
  bool slots_ok = true;
  for (int q = 0; q < LU_NPAR; ++q) {
    c.pairs[q] = (int*)ws_carve((size_t)(2 * c.maxpairs + 1) * sizeof(int));
    c.linv[q] = (double*)ws_carve((size_t)nb * nb * sizeof(double));
    x.linv32[q] = (float*)ws_carve((size_t)nb * nb * sizeof(float));
    slots_ok = slots_ok && c.pairs[q] && c.linv[q] && x.linv32[q];
  }
  for (int q = 0; q < 2; ++q) {
    x.tmp32[q] = (float*)ws_carve((size_t)n * nb * sizeof(float));
    c.tmp[q] = nullptr;
  }
  c.ctr = (unsigned*)ws_carve(256);
  c.err = err_flag ? err_flag : (int*)ws_carve(256);
  x.bf16 = bf16;
  x.amax[0] = (unsigned*)ws_carve(256);
  x.amax[1] = (unsigned*)ws_carve(256);
  if (!x.T32 || !x.P64 || !c.xa || !c.stage || !c.piv || !c.perm || !c.inv0 ||
      !slots_ok || !x.tmp32[0] || !x.tmp32[1] || !c.ctr || !c.err || !x.amax[0] || !x.amax[1]) {
    set_error("getrf_mixed: workspace too small");
    return ST_NOMEM;
  }
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, c.err, 1, 0);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, c.ctr);
  hipLaunchKernelGGL(k_fill_u64_lu, dim3(grid1(LEAF_RING * LEAF_XAREA)), dim3(256), 0, s,
                     (unsigned long long*)c.xa, (long)LEAF_RING * LEAF_XAREA, XSENT);
  c.sync_base = 0;
  hipLaunchKernelGGL(k_iota, dim3(grid1(n)), dim3(256), 0, s, c.perm, n);
  hipLaunchKernelGGL((k_transpose<double, float>), dim3(ceil_div(n, 64), ceil_div(n, 64)),
                     dim3(256), 0, s, n, n, A, lda, x.T32, (long)n);
  MOPS_CHECK_HIP(hipGetLastError());

  // the look-ahead schedule of getrf_d, on FP32 trailing data
  hipStream_t s2 = aux_stream();
  hipEvent_t e_fork = sync_event(0), e_pan = sync_event(1), e_rest = sync_event(2),
             e_join = sync_event(3);
  if (!s2 || !e_fork || !e_pan || !e_rest || !e_join) {
    set_error("getrf_mixed: could not create the look-ahead stream");
    return ST_HIP;
  }
  hipStream_t sr = s;
  MOPS_TRY(fork_bulk(s, sr));
  MOPS_CHECK_HIP(hipEventRecord(e_fork, s));
  MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_fork, 0));
  MOPS_TRY(panel_mixed(s2, c, x, 0, nb, 0));
  MOPS_CHECK_HIP(hipEventRecord(e_pan, s2));
  bool have_rest = false;
  int par = 0;
  for (int k = 0; k < n; k += nb, par = (par + 1) % LU_NPAR) {
    const int kb = min(nb, n - k);
    const int kb1 = min(nb, n - k - kb);
    const int j_next = k + kb, j_rest = k + kb + kb1;
    if (have_rest) MOPS_CHECK_HIP(hipStreamWaitEvent(s2, e_rest, 0));
    MOPS_CHECK_HIP(hipStreamWaitEvent(sr, e_pan, 0));
    have_rest = false;
    if (j_rest < n) {
      // block column k+2 first (see getrf_d)
      const int j_far = min(n, j_rest + nb);
      MOPS_TRY(update_cols32(sr, c, x, k, kb, par, j_rest, j_far, x.tmp32[0], true));
      MOPS_CHECK_HIP(hipEventRecord(e_rest, sr));
      have_rest = true;
      MOPS_TRY(update_cols32(sr, c, x, k, kb, par, j_far, n,
                             x.tmp32[0] + (long)(j_far - j_rest) * kb, true));
    }
    MOPS_TRY(apply_rows32(sr, c, x, par, 0, k));
    if (kb1 > 0) {
      MOPS_TRY(update_cols32(s2, c, x, k, kb, par, j_next, j_next + kb1, x.tmp32[1], false));
      MOPS_TRY(panel_mixed(s2, c, x, j_next, kb1, (par + 1) % LU_NPAR));
      MOPS_CHECK_HIP(hipEventRecord(e_pan, s2));
    }
  }
  MOPS_CHECK_HIP(hipEventRecord(e_join, s2));
  MOPS_CHECK_HIP(hipStreamWaitEvent(s, e_join, 0));
  MOPS_TRY(join_bulk(s, sr));
  hipLaunchKernelGGL((k_transpose<float, double>), dim3(ceil_div(n, 64), ceil_div(n, 64)),
                     dim3(256), 0, s, n, n, x.T32, (long)n, LU, ldlu);
  MOPS_CHECK_HIP(hipGetLastError());
  if (ipiv) {
    hipLaunchKernelGGL(k_piv_to_i64, dim3(grid1(n)), dim3(256), 0, s, c.piv, ipiv, n);
    MOPS_CHECK_HIP(hipGetLastError());
  }
  if (perm_out) MOPS_CHECK_HIP(hipMemcpyAsync(perm_out, c.perm, (size_t)n * sizeof(int),
                                              hipMemcpyDeviceToDevice, s));
  return ST_OK;
}

size_t getrs_prep_ws_bytes(int n, int nrhs) {
  // two preparations, and trsm_d's own workspace when the solve is not chained
  return 2 * trsm_prep_ws_bytes(n, nrhs) + trsm_ws_bytes(F_LEFT, n, nrhs) + 512;
}

int getrs_prep_d(hipStream_t s, int n, int nrhs, const double* A, long lda, GetrsPrep& g) {
  g = GetrsPrep();
  if (n <= 0 || nrhs <= 0) return ST_OK;
  MOPS_TRY(trsm_prep_d(s, F_LEFT | F_LOWER | F_UNIT, n, nrhs, A, lda, nullptr, g.l));
  // one sentinel copy serves both solves (they run one after the other)
  return trsm_prep_d(s, F_LEFT | F_UPPER, n, nrhs, A, lda, g.l.W, g.u);
}

int getrs_prep_solve_d(hipStream_t s, int n, int nrhs, const double* A, long lda, const int* perm,
                       const double* R, long ldr, double* Out, long ldo, const GetrsPrep& g) {
  if (n <= 0 || nrhs <= 0) return ST_OK;
  // SolveLU (lu.go:320-330): pivots, then L (unit lower), then U
  hipLaunchKernelGGL(k_gather_rows, dim3(grid1((long)n * nrhs)), dim3(256), 0, s, n, nrhs, perm,
                     R, ldr, Out, ldo);
  MOPS_CHECK_HIP(hipGetLastError());
  MOPS_TRY(trsm_prep_solve_d(s, F_LEFT | F_LOWER | F_UNIT, n, nrhs, A, lda, Out, ldo, g.l));
  return trsm_prep_solve_d(s, F_LEFT | F_UPPER, n, nrhs, A, lda, Out, ldo, g.u);
}

size_t getrs_ws_bytes(int n, int nrhs) {
  return (size_t)n * nrhs * sizeof(double) + (size_t)n * sizeof(int) + 512 +
         2 * trsm_ws_bytes(F_LEFT, n, nrhs);
}

int getrs_d(hipStream_t s, int flags, int n, int nrhs, const double* A, long lda,
            const long long* ipiv, int npiv, const int* perm, double* B, long ldb) {
  if (n <= 0 || nrhs <= 0) return ST_OK;
  double* W = (double*)ws_carve((size_t)n * nrhs * sizeof(double));
  int* pm = (int*)ws_carve((size_t)n * sizeof(int));
  if (!W || !pm) {
    set_error("getrs: workspace too small");
    return ST_NOMEM;
  }
  if (!perm) {
    const size_t shm = n <= 16384 ? (size_t)n * sizeof(int) : 0;
    hipLaunchKernelGGL(k_perm_from_piv, dim3(1), dim3(1024), shm, s, ipiv, npiv, n, pm);
    MOPS_CHECK_HIP(hipGetLastError());
    perm = pm;
  }
  if (!(flags & F_TRANSA)) {
    // SolveLU (lu.go:320-330): pivots, then L (unit lower), then U
    hipLaunchKernelGGL(k_gather_rows, dim3(grid1((long)n * nrhs)), dim3(256), 0, s, n, nrhs,
                       perm, B, ldb, W, n);
    MOPS_CHECK_HIP(hipGetLastError());
    MOPS_TRY(trsm_d(s, F_LEFT | F_LOWER | F_UNIT, n, nrhs, 1.0, A, lda, W, n));
    MOPS_TRY(trsm_d(s, F_LEFT | F_UPPER, n, nrhs, 1.0, A, lda, W, n));
    return copy_tile_d(s, n, nrhs, W, n, B, ldb);
  }
  // A^T X = B with P A = L U: A^T = U^T L^T P, so X = P^T L^-T U^-T B.
  // (The reference applies the pivots first here, lu.go:322-326, which
  // does not solve A^T X = B; this path solves it.)
  MOPS_TRY(trsm_d(s, F_LEFT | F_UPPER | F_TRANSA, n, nrhs, 1.0, A, lda, B, ldb));
  MOPS_TRY(trsm_d(s, F_LEFT | F_LOWER | F_UNIT | F_TRANSA, n, nrhs, 1.0, A, lda, B, ldb));
  // X[perm[i]] = Y[i]
  MOPS_TRY(copy_tile_d(s, n, nrhs, B, ldb, W, n));
  hipLaunchKernelGGL(k_scatter_rows, dim3(grid1((long)n * nrhs)), dim3(256), 0, s, n, nrhs, perm,
                     W, n, B, ldb);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

}  // namespace mops
