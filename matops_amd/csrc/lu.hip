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
// Panel (M x w, rows k..m-1): recursive halving down to LEAF_W-wide leaves.
// A leaf is one cooperative kernel: G workgroups each hold up to 256 rows of
// the leaf in LDS; per column every workgroup finds its best candidate
// (wave shuffles), publishes it with the candidate row, and after one grid
// barrier all of them agree on the pivot, swap the two rows (in LDS for the
// leaf columns, in global memory for all other columns of A, sliced across
// the workgroups) and apply the rank-1 update to their rows. Between the
// halves, the U block is solved with the triangular-inverse TRSM and the
// lower rows are updated with one GEMM; after the panel the trailing matrix
// gets the TRSM of U12 and the GEMM A22 -= L21*U12 on the matrix cores.
#include <climits>
#include <cmath>

#include "mops_lu.hpp"
#include "mops_solvers.hpp"

namespace mops {
namespace {

constexpr int LEAF_W = 32;     // leaf width (columns)
constexpr int LEAF_R = 256;    // rows per workgroup in a leaf
constexpr int LEAF_T = 256;    // threads per leaf workgroup
constexpr unsigned SPIN_LIMIT = 1u << 26;

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

// pivot candidate: value key (NaN mapped), row; larger key wins, ties -> lower row
struct Cand {
  double v;
  int r;
};
__device__ __forceinline__ bool better(double v1, int r1, double v2, int r2) {
  return v1 > v2 || (v1 == v2 && r1 < r2);
}

__device__ __forceinline__ void st_agent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_agent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct LeafArgs {
  double* T;       // A^T, column-major, T[c + r*ldT] = A[r][c]
  long ldT;
  int n;           // columns of A (= rows of T)
  int m;           // rows of A
  int c0;          // first column of the leaf == first row of the panel
  int w;           // leaf width (<= LEAF_W)
  int* piv;        // global pivots, piv[c0 + j]
  int* perm;       // row permutation (perm[i] = original row now at i), or null
  double* xb;      // exchange: [2][G][LEAF_W + 3]
  double* rowk;    // exchange: [2][LEAF_W + 1]
  unsigned* ctr;   // grid barrier counter (zeroed before launch)
  int* err;        // set to 1 if a barrier wait timed out
};

// Grid barrier among the gridDim.x workgroups of a cooperative launch.
__device__ __forceinline__ void grid_sync(unsigned* ctr, unsigned target, int* err) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (++spins > SPIN_LIMIT) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(LEAF_T) void k_getf2_leaf(LeafArgs a) {
  constexpr int W = LEAF_W, LDW = LEAF_W + 1;
  __shared__ double tile[LEAF_R * LDW];
  __shared__ double prow[W + 1];
  __shared__ double okrow[W + 1];
  __shared__ int shp[2];
  __shared__ double wv[LEAF_T / 64];
  __shared__ int wr[LEAF_T / 64];
  __shared__ int lperm[LEAF_R];

  const int G = gridDim.x, g = blockIdx.x, t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int r0 = a.c0;  // panel top row
  const int M = a.m - r0;
  const int rg0 = r0 + g * LEAF_R;
  const int nr = max(0, min(LEAF_R, a.m - rg0));
  const int w = a.w;
  double* T = a.T;
  const long ldT = a.ldT;

  // load my rows of the leaf: tile[r][c] = A[rg0 + r][c0 + c]
  for (int idx = t; idx < LEAF_R * W; idx += LEAF_T) {
    const int r = idx / W, c = idx % W;
    double v = 0.0;
    if (r < nr && c < w) v = T[(a.c0 + c) + (long)(rg0 + r) * ldT];
    tile[r * LDW + c] = v;
  }
  if (t < LEAF_R) lperm[t] = (a.perm && t < nr) ? a.perm[rg0 + t] : rg0 + t;
  __syncthreads();
  (void)M;

  const int S = (a.n + G - 1) / G;  // slice of other columns swapped by this WG
  for (int j = 0; j < w; ++j) {
    const int k = r0 + j;  // diagonal row
    double* xbj = a.xb + (long)(j & 1) * G * (W + 3);
    double* rkj = a.rowk + (j & 1) * (W + 1);
    // ---- 1. local candidate among my rows >= k
    double v = -2.0;
    int r = INT_MAX;
    if (t < nr) {
      const int gr = rg0 + t;
      if (gr >= k) {
        const double x = fabs(tile[t * LDW + j]);
        v = isnan(x) ? (gr == k ? INFINITY : -1.0) : x;
        r = gr;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double v2 = __shfl_xor(v, off, 64);
      const int r2 = __shfl_xor(r, off, 64);
      if (better(v2, r2, v, r)) {
        v = v2;
        r = r2;
      }
    }
    if (lane == 0) {
      wv[wave] = v;
      wr[wave] = r;
    }
    __syncthreads();
    if (t == 0) {
      double bv = wv[0];
      int br = wr[0];
      for (int q = 1; q < LEAF_T / 64; ++q)
        if (better(wv[q], wr[q], bv, br)) {
          bv = wv[q];
          br = wr[q];
        }
      wv[0] = bv;
      wr[0] = br;
    }
    __syncthreads();
    // ---- 2. publish candidate (value, row, its leaf row, its perm entry)
    {
      const int br = wr[0];
      double* slot = xbj + (long)g * (W + 3);
      if (t == 0) {
        st_agent(&slot[0], wv[0]);
        st_agent(&slot[1], (double)br);
      }
      if (br != INT_MAX) {
        const int lr = br - rg0;
        if (t < W) st_agent(&slot[2 + t], tile[lr * LDW + t]);
        if (t == W) st_agent(&slot[2 + W], (double)lperm[lr]);
      }
      if (k >= rg0 && k < rg0 + nr) {
        const int lr = k - rg0;
        if (t < W) st_agent(&rkj[t], tile[lr * LDW + t]);
        if (t == W) st_agent(&rkj[W], (double)lperm[lr]);
      }
    }
    grid_sync(a.ctr, (unsigned)(j + 1) * (unsigned)G, a.err);
    // ---- 3. global pivot
    {
      double bv = -3.0;
      int br = INT_MAX;
      for (int q = t; q < G; q += LEAF_T) {
        const double* slot = xbj + (long)q * (W + 3);
        const double sv = ld_agent(&slot[0]);
        const int sr = (int)ld_agent(&slot[1]);
        if (better(sv, sr, bv, br)) {
          bv = sv;
          br = sr;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double v2 = __shfl_xor(bv, off, 64);
        const int r2 = __shfl_xor(br, off, 64);
        if (better(v2, r2, bv, br)) {
          bv = v2;
          br = r2;
        }
      }
      if (lane == 0) {
        wv[wave] = bv;
        wr[wave] = br;
      }
      __syncthreads();
      if (t == 0) {
        for (int q = 1; q < LEAF_T / 64; ++q)
          if (better(wv[q], wr[q], wv[0], wr[0])) {
            wv[0] = wv[q];
            wr[0] = wr[q];
          }
        int p = wr[0];
        if (p == INT_MAX) p = k;  // no candidate at all (cannot happen for k < m)
        shp[0] = p;
        if (g == 0) a.piv[k] = p;
      }
      __syncthreads();
      const int p = shp[0];
      const int gp = (p - r0) / LEAF_R;
      const double* ws = xbj + (long)gp * (W + 3);
      if (t <= W) {
        prow[t] = ld_agent(&ws[2 + t]);  // pivot row (+ its perm entry)
        okrow[t] = ld_agent(&rkj[t]);    // old row k (+ its perm entry)
      }
    }
    __syncthreads();
    const int p = shp[0];
    // ---- 4. swap rows k and p
    if (p != k) {
      if (p >= rg0 && p < rg0 + nr) {
        const int lr = p - rg0;
        if (t < W) tile[lr * LDW + t] = okrow[t];
        if (t == W) lperm[lr] = (int)okrow[W];
      }
      if (k >= rg0 && k < rg0 + nr) {
        const int lr = k - rg0;
        if (t < W) tile[lr * LDW + t] = prow[t];
        if (t == W) lperm[lr] = (int)prow[W];
      }
      // the same interchange in every other column of A (T rows), my slice
      const int cs = g * S, ce = min(a.n, cs + S);
      double* Tk = T + (long)k * ldT;
      double* Tp = T + (long)p * ldT;
      for (int c = cs + t; c < ce; c += LEAF_T) {
        if (c >= a.c0 && c < a.c0 + w) continue;
        const double x = Tk[c];
        Tk[c] = Tp[c];
        Tp[c] = x;
      }
    }
    __syncthreads();
    // ---- 5. column of L and rank-1 update of my rows below k
    const double pv = prow[j];
    if (t < nr) {
      const int gr = rg0 + t;
      if (gr > k) {
        double* row = &tile[t * LDW];
        const double l = row[j] / pv;
        row[j] = l;
        for (int c = j + 1; c < w; ++c) row[c] -= l * prow[c];
      }
    }
    __syncthreads();
  }
  // write my rows back
  for (int idx = t; idx < LEAF_R * W; idx += LEAF_T) {
    const int r = idx / W, c = idx % W;
    if (r < nr && c < w) T[(a.c0 + c) + (long)(rg0 + r) * ldT] = tile[r * LDW + c];
  }
  if (a.perm && t < nr) a.perm[rg0 + t] = lperm[t];
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
  hipStream_t s;
  double* T;
  long ldT;
  int m, n;
  int* piv;
  int* perm;
  double* xb;
  double* rowk;
  unsigned* ctr;
  int* err;
};

int leaf(LuCtx& c, int c0, int w) {
  const int M = c.m - c0;
  const int G = ceil_div(M, LEAF_R);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, c.s, c.ctr);
  MOPS_CHECK_HIP(hipGetLastError());
  LeafArgs a;
  a.T = c.T;
  a.ldT = c.ldT;
  a.n = c.n;
  a.m = c.m;
  a.c0 = c0;
  a.w = w;
  a.piv = c.piv;
  a.perm = c.perm;
  a.xb = c.xb;
  a.rowk = c.rowk;
  a.ctr = c.ctr;
  a.err = c.err;
  void* args[] = {&a};
  MOPS_CHECK_HIP(hipLaunchCooperativeKernel((const void*)k_getf2_leaf, dim3(G), dim3(LEAF_T),
                                            args, 0, c.s));
  return ST_OK;
}

// recursive factorization of the panel of columns [c0, c0+w), rows c0..m-1
int panel(LuCtx& c, int c0, int w) {
  if (w <= LEAF_W) return leaf(c, c0, w);
  int h = ((w / 2 + LEAF_W - 1) / LEAF_W) * LEAF_W;
  if (h >= w) h = w - LEAF_W;
  MOPS_TRY(panel(c, c0, h));
  const long ld = c.ldT;
  double* T = c.T;
  // U block of the right half: X = A[c0:c0+h, c0+h:c0+w] = L11^-1 X, i.e. in
  // T: X^T <- X^T * (L11^T)^-1 with L11^T unit upper in T
  MOPS_TRY(trsm_d(c.s, F_RIGHT | F_UPPER | F_UNIT, h, w - h, 1.0, T + c0 + (long)c0 * ld, ld,
                  T + (c0 + h) + (long)c0 * ld, ld));
  const int below = c.m - c0 - h;
  if (below > 0)
    MOPS_TRY(gemm_d(c.s, false, false, w - h, below, h, -1.0, T + (c0 + h) + (long)c0 * ld, ld,
                    T + c0 + (long)(c0 + h) * ld, ld, 1.0, T + (c0 + h) + (long)(c0 + h) * ld,
                    ld));
  if (c0 + h < c.m) MOPS_TRY(panel(c, c0 + h, w - h));
  return ST_OK;
}

}  // namespace

size_t getrf_ws_bytes(int m, int n, int nb) {
  if (m <= 0 || n <= 0) return 0;
  if (nb <= 0 || nb > (m < n ? m : n)) nb = m < n ? m : n;
  const int mn = m < n ? m : n;
  const int G = ceil_div(m, LEAF_R);
  size_t b = (size_t)m * n * sizeof(double);                     // T
  b += (size_t)2 * G * (LEAF_W + 3) * sizeof(double) + 4096;     // exchange
  b += (size_t)(mn + m + 64) * sizeof(int);                      // piv, perm, ctr, err
  b += trsm_ws_bytes(F_RIGHT, nb > LEAF_W ? nb : LEAF_W, n) + 8 * 256;
  return b;
}

int getrf_d(hipStream_t s, int m, int n, int nb, double* A, long lda, long long* ipiv,
            int* perm_out, int* err_flag) {
  if (m <= 0 || n <= 0) return ST_OK;
  const int mn = m < n ? m : n;
  // DecomposeLU: unblocked iff min(m,n) <= nb || nb == 0 (lu.go:267); both
  // orders produce the same factorization, here one panel covers it all
  if (nb <= 0 || nb >= mn) nb = mn;
  const int G = ceil_div(m, LEAF_R);
  LuCtx c;
  c.s = s;
  c.m = m;
  c.n = n;
  c.ldT = n;
  c.T = (double*)ws_carve((size_t)m * n * sizeof(double));
  c.xb = (double*)ws_carve((size_t)2 * G * (LEAF_W + 3) * sizeof(double));
  c.rowk = (double*)ws_carve(2 * (LEAF_W + 1) * sizeof(double));
  c.piv = (int*)ws_carve((size_t)mn * sizeof(int));
  c.perm = (int*)ws_carve((size_t)m * sizeof(int));
  c.ctr = (unsigned*)ws_carve(256);
  c.err = err_flag ? err_flag : (int*)ws_carve(256);
  if (!c.T || !c.xb || !c.rowk || !c.piv || !c.perm || !c.ctr || !c.err) {
    set_error("getrf: workspace too small");
    return ST_NOMEM;
  }
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, c.err, 1, 0);
  hipLaunchKernelGGL(k_iota, dim3(grid1(m)), dim3(256), 0, s, c.perm, m);
  // T = A^T
  hipLaunchKernelGGL((k_transpose<double, double>), dim3(ceil_div(m, 64), ceil_div(n, 64)),
                     dim3(256), 0, s, m, n, A, lda, c.T, c.ldT);
  MOPS_CHECK_HIP(hipGetLastError());
  for (int k = 0; k < mn; k += nb) {
    const int kb = min(nb, mn - k);
    trace_mark_begin(s, TR_PANEL);
    MOPS_TRY(panel(c, k, kb));
    trace_mark_end(s, TR_PANEL, 0.0);
    if (k + kb < n) {
      double* T = c.T;
      const long ld = c.ldT;
      // U12: A[k:k+kb, k+kb:n] = L11^-1 A12
      MOPS_TRY(trsm_d(s, F_RIGHT | F_UPPER | F_UNIT, kb, n - k - kb, 1.0, T + k + (long)k * ld,
                      ld, T + (k + kb) + (long)k * ld, ld));
      if (k + kb < m) {
        // A22 -= L21 U12  <=>  T22 -= T12 * T21
        trace_mark_begin(s, TR_GEMM_UPDATE);
        MOPS_TRY(gemm_d(s, false, false, n - k - kb, m - k - kb, kb, -1.0,
                        T + (k + kb) + (long)k * ld, ld, T + k + (long)(k + kb) * ld, ld, 1.0,
                        T + (k + kb) + (long)(k + kb) * ld, ld));
        trace_mark_end(s, TR_GEMM_UPDATE, 2.0 * (double)(n - k - kb) * (m - k - kb) * kb);
      }
    }
  }
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
