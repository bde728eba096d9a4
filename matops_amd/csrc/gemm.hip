// Column-major GEMM on the gfx950 matrix cores.
//
// Replaces the blocked dot-product GEMM of calgo/dmmat_gemm.c:364-416
// (dmult_mm_blocked3 -> _dblock_mult_gepp :309 -> _dblock_ddot_sse :14) and
// its beta pre-scaling dscale_tile (calgo/vec_scale.c:114). The same kernel,
// restricted to one triangle of C, is the SYRK of calgo/dmmat_rank.c:62-126.
//
// Design (one workgroup per BMxBN tile of C; 4 or 8 wave64s):
//  * k is walked in BK-deep slabs. A slab of op(A) (BM x BK) and op(B)
//    (BK x BN) is staged in LDS as [k][m] and [k][n] rows, double buffered;
//    the next slab is fetched into registers while the current one feeds
//    the matrix cores, so global latency hides behind MFMA work.
//  * Each wave owns a WM x WN sub-tile made of 16x16 MFMA blocks
//    (v_mfma_f64_16x16x4_f64, or v_mfma_f32_16x16x4_f32 for FP32). The
//    instruction computes D = op(B)^T * op(A)^T, i.e. C transposed, so lane
//    l holds C row l%16 and the 16 lanes of a group write one contiguous
//    run of a column: stores to column-major C coalesce.
//  * Workgroup ids are remapped so that each of the 8 XCDs receives a
//    contiguous band of tiles (its own L2 then serves neighbouring tiles),
//    and tiles are grouped GROUP_M tile-rows at a time for reuse of op(B).
#include "mops_kernels.hpp"

#include <cmath>

namespace mops {
namespace {

template <typename T>
struct Mfma;
// Operand layout (both types): lane l supplies A[l%16][l/16] and
// B[l/16][l%16] of the 16x4 and 4x16 operands. Result layout: register r of
// lane l holds D[drow(l/16, r)][l%16]; the 64-bit result of the FP64
// instruction interleaves rows (row = l/16 + 4r), the FP32 one keeps four
// consecutive rows per lane group (row = 4*(l/16) + r).
template <>
struct Mfma<double> {
  typedef v4d acc_t;
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int drow(int g, int r) { return g + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef v4f acc_t;
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int drow(int g, int r) { return 4 * g + r; }
};

template <typename T>
struct GemmArgs {
  int M, N, K;
  const T* A;
  long lda;
  const T* B;
  long ldb;
  T* C;
  long ldc;
  T alpha, beta;
  int mode, tri;
  int tiles_m, tiles_n;
  long sA, sB, sC;     // element strides between batch entries (blockIdx.y)
  long sA2, sB2, sC2;  // second batch dimension (blockIdx.z)
  // split of k (kchunk > 0, never with batches): blockIdx.y = chunk; the
  // chunks before the last leave their partial tiles in part and raise
  // kflags; the last chunk adds them in chunk order and writes C
  int kchunk;
  int order;  // TILE_* visiting order of the output tiles
  T* part;
  int* kflags;
  int kready;   // value of a set flag in this launch
  int* status;  // raised (STATUS_GEMM_WAIT) when a wait for a chunk times out
  unsigned spin_limit;  // polls of one chunk flag before giving up
};

constexpr int GROUP_M = 8;
constexpr int SMALL_KS = 2;  // k shares of the 64x64-tile kernel for full outputs (8 waves)
// Super-tile edge (in tiles) of the triangle-only updates. 4 rather than 8
// or 16: a diagonal super-tile leaves SG(SG-1)/2 of its SG^2 slots empty,
// and the Cholesky n = 8192 measured 7.11 (4), 7.13 (2), 7.22 (8) and
// 7.72 ms (16).
constexpr int SG = 4;
constexpr int NXCD = 8;

// Visiting order of the output tiles. TILE_GROUPED: groups of GROUP_M tile
// rows, renumbered so that each XCD gets a contiguous range (L2 reuse of
// the row panels). The others serve products whose tiles have very
// different k ranges (triangular operands): heaviest tiles first, and the
// ids go to the XCDs round robin as the hardware hands them out, so every
// XCD gets its share of heavy and light tiles.
//   TILE_ROWS_UP / _DOWN     row by row from the top / bottom (work
//                            depends on the tile row)
//   TILE_COLS_LEFT / _RIGHT  column by column from the left / right
//   TILE_SHELLS              square grids: shells d = max(tm, tn) = 0, 1, ...
//                            (work decreasing in max(tm, tn))
//   TILE_LOWER_ROWS          lower triangle tm >= tn, row by row from the
//                            top (work decreasing in tm)
//   TILE_LOWER_BANDS         lower triangle by diagonals tm - tn from the
//                            farthest (work increasing in tm - tn)
//   TILE_LOWER_TRAP          the tiles of a lower trapezoid (GM_LOWER_T)
//                            that hold elements with row >= column, row by
//                            row from the top: min(tm + 1, tiles_n) in row
//                            tm (the other tiles are not launched)
enum : int {
  TILE_GROUPED = 0,
  TILE_ROWS_UP = 1,
  TILE_ROWS_DOWN = 2,
  TILE_COLS_LEFT = 3,
  TILE_COLS_RIGHT = 4,
  TILE_SHELLS = 5,
  TILE_LOWER_ROWS = 6,
  TILE_LOWER_BANDS = 7,
  TILE_LOWER_TRAP = 8
};

// r with r (r + 1) / 2 <= x < (r + 1) (r + 2) / 2
__device__ __forceinline__ int tri_root(int x) {
  long r = (long)((sqrtf(8.0f * (float)x + 1.0f) - 1.0f) * 0.5f);
  while ((r + 1) * (r + 2) / 2 <= x) ++r;
  while (r * (r + 1) / 2 > x) --r;
  return (int)r;
}

// Hardware hands workgroup ids to the XCDs round robin (id % 8). Renumber so
// that XCD x sees a contiguous range of logical tile indices.
__device__ __forceinline__ int xcd_linear(int bid, int nblk) {
  int x = bid % NXCD;
  int q = nblk / NXCD, r = nblk % NXCD;
  return x * q + (x < r ? x : r) + bid / NXCD;
}

// TRI: instance for the triangle-only modes (GM_LOWER / GM_UPPER), which
// keeps the symmetric rank-k updates apart from the other GEMMs in profiles.
// AL: every tile lies inside M x N and every slab inside K (M % BM, N % BN,
// K % BK and the k chunk % BK all zero): no bounds checks are compiled in
// (DGEMM 1024^3 67.3 -> 63 us).
// KS: waves per output sub-tile, splitting each slab's k steps between them
// (their partial tiles are added through LDS at the end): more waves per
// SIMD for small tiles without more workgroups.
template <typename T, int BM, int BN, int BK, int WGM, int WGN, int KS, bool TA, bool TB, bool TRI,
          bool AL>
__global__ __launch_bounds__(64 * WGM * WGN * KS) void k_gemm(GemmArgs<T> g) {
  constexpr int NTS = 64 * WGM * WGN;  // threads of one k share
  constexpr int NT = NTS * KS;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int FM = WM / 16, FN = WN / 16;
  // Row pitch = 128 bytes modulo 256 (doubles), so the two k-rows read by
  // lanes 0-31 of one ds_read_b64 fall in disjoint bank halves.
  constexpr int PAD = 16;
  constexpr int LDS_A = BM + PAD, LDS_B = BN + PAD;
  // Column swizzle of the staged slabs: element (k, i) sits at column
  // i ^ (4 * ((k >> 1) & 7)) of row k. An operand that is contiguous along k
  // is staged as (k, k+1) pairs: eight neighbouring lanes write one column
  // at rows 0, 2, ..., 14, which the pitch alone maps to one bank (an 8-way
  // conflict on every slab store); the swizzle gives them 8 banks. It only
  // flips bits 2-4 of the column, so element pairs stay contiguous, a row
  // keeps its footprint, and the MFMA reads (16 consecutive columns of rows
  // 4r..4r+3; rows 4r and 4r+1 share the swizzle) keep their bank halves.
  // When neither operand is staged as k pairs (A not transposed, B
  // transposed: the Cholesky's SYRK and block-column updates) nothing is
  // swizzled, so the MFMA reads of a slab share base registers and use
  // immediate offsets (SYRK n = 7680, k = 512: 46.4 -> 47.9 TFLOP/s).
  // Dropping it for one operand of the other forms made the LU slower.
  constexpr bool SW = TA || !TB;
  auto swz = [](int kk, int i) { return SW ? i ^ (4 * ((kk >> 1) & 7)) : i; };
  auto posA = [&](int kk, int i) { return kk * LDS_A + swz(kk, i); };
  auto posB = [&](int kk, int i) { return kk * LDS_B + swz(kk, i); };
  // operands move in pairs of consecutive elements along their contiguous
  // dimension (m or n when not transposed for A / transposed for B, else k)
  constexpr int A_PP = BM * BK / 2 / NT, B_PP = BN * BK / 2 / NT;
  static_assert(A_PP * NT * 2 == BM * BK && B_PP * NT * 2 == BN * BK, "tile");
  typedef typename Mfma<T>::acc_t acc_t;

  static_assert(BK % (4 * KS) == 0, "k split");
  // 128x128 tiles and the instances with two k shares keep two slabs in
  // flight: with 8 waves per workgroup they run few workgroups per CU, and
  // a slab's loads need more than one slab of MFMAs to arrive (DGEMM 1024^3
  // 68 -> 65 us). The 4-wave triangle instance does not: more registers
  // made the SYRK of the Cholesky slower.
  constexpr bool PF2 = (BM == 128 && BN == 128 && KS == 1) || KS == 2;
  constexpr int SMEM_AB = 2 * BK * (LDS_A + LDS_B);
  constexpr int SMEM_RED = (KS - 1) * BM * BN;
  __shared__ T smem[SMEM_AB > SMEM_RED ? SMEM_AB : SMEM_RED];
  T(*As)[BK * LDS_A] = reinterpret_cast<T(*)[BK * LDS_A]>(smem);
  T(*Bs)[BK * LDS_B] = reinterpret_cast<T(*)[BK * LDS_B]>(smem + 2 * BK * LDS_A);

  // ---- tile selection ----
  const int lin = g.order == TILE_GROUPED ? xcd_linear(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  int tm, tn;
  if (g.order != TILE_GROUPED) {
    const int Tm = g.tiles_m, Tn = g.tiles_n;
    switch (g.order) {
      case TILE_ROWS_UP: tm = lin / Tn; tn = lin % Tn; break;
      case TILE_ROWS_DOWN: tm = Tm - 1 - lin / Tn; tn = lin % Tn; break;
      case TILE_COLS_LEFT: tn = lin / Tm; tm = lin % Tm; break;
      case TILE_COLS_RIGHT: tn = Tn - 1 - lin / Tm; tm = lin % Tm; break;
      case TILE_SHELLS: {
        long d = (long)sqrtf((float)lin);
        while ((d + 1) * (d + 1) <= lin) ++d;
        while (d * d > lin) --d;
        const int o = (int)(lin - d * d);  // shell d: (d, 0..d), then (0..d-1, d)
        if (o <= d) {
          tm = (int)d;
          tn = o;
        } else {
          tm = o - (int)d - 1;
          tn = (int)d;
        }
        break;
      }
      case TILE_LOWER_ROWS: {
        const int r = tri_root(lin);
        tm = r;
        tn = (int)(lin - (long)r * (r + 1) / 2);
        break;
      }
      case TILE_LOWER_TRAP: {
        // rows 0 .. Tn - 1 hold a triangle of Tn (Tn + 1) / 2 tiles, the rows
        // below Tn tiles each
        const long tri = (long)Tn * (Tn + 1) / 2;
        if (lin < tri) {
          const int r = tri_root(lin);
          tm = r;
          tn = (int)(lin - (long)r * (r + 1) / 2);
        } else {
          tm = Tn + (int)((lin - tri) / Tn);
          tn = (int)((lin - tri) % Tn);
        }
        break;
      }
      default: {  // TILE_LOWER_BANDS: band r holds the r + 1 tiles of tm - tn = Tm - 1 - r
        const int r = tri_root(lin);
        tn = (int)(lin - (long)r * (r + 1) / 2);
        tm = tn + Tm - 1 - r;
        break;
      }
    }
    if (tm >= Tm || tn >= Tn) return;
  } else if (g.mode == GM_FULL || g.mode == GM_LOWER_T || g.mode == GM_UPPER_T) {
    const int per_group = GROUP_M * g.tiles_n;
    const int grp = lin / per_group;
    const int first = grp * GROUP_M;
    const int gsz = min(GROUP_M, g.tiles_m - first);
    const int in = lin - grp * per_group;
    tm = first + in % gsz;
    tn = in / gsz;
  } else {
    // Triangle-only update: tiles are visited in SGxSG super-tiles (so the
    // WGs that run together share row panels of op(A) and op(B) in L2);
    // super-tiles are enumerated row by row over the lower triangle. Slots
    // that fall outside the triangle exit at once.
    const int sidx = lin / (SG * SG), w = lin % (SG * SG);
    int r = (int)((sqrtf(8.0f * (float)sidx + 1.0f) - 1.0f) * 0.5f);
    while ((r + 1) * (r + 2) / 2 <= sidx) ++r;
    while (r * (r + 1) / 2 > sidx) --r;
    const int c = sidx - r * (r + 1) / 2;
    if (g.mode == GM_LOWER) {
      tm = r * SG + w % SG;
      tn = c * SG + w / SG;
      if (tm >= g.tiles_m || tn > tm) return;
    } else {
      tm = c * SG + w % SG;
      tn = r * SG + w / SG;
      if (tn >= g.tiles_m || tm > tn) return;
    }
  }
  const int m0 = tm * BM, n0 = tn * BN;
  // trapezoids: skip tiles entirely on the other side of the diagonal
  if (g.mode == GM_LOWER_T && m0 + BM <= n0) return;
  if (g.mode == GM_UPPER_T && m0 >= n0 + BN) return;

  int kbeg = 0, kend = g.K;
  switch (g.tri) {
    case TRI_A_LOWER: kend = min(g.K, m0 + BM); break;
    case TRI_A_UPPER: kbeg = m0; break;
    case TRI_B_LOWER: kbeg = n0; break;
    case TRI_B_UPPER: kend = min(g.K, n0 + BN); break;
    case TRI_AB_LOWER:
      kbeg = n0;
      kend = min(g.K, m0 + BM);
      break;
    case TRI_AU_BL: kbeg = max(m0, n0); break;
    default: break;
  }
  kbeg = (kbeg / BK) * BK;
  if (g.kchunk > 0) {
    kbeg = blockIdx.y * g.kchunk;
    kend = min(g.K, kbeg + g.kchunk);
  }

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int kh = wave / (WGM * WGN), wsp = wave % (WGM * WGN);  // k share, sub-tile
  const int wm = wsp % WGM, wn = wsp / WGM;
  const int stid = wsp * 64 + lane;  // thread index within the k share
  const bool lead = kh == 0;         // the k share that finishes the tile
  const int lr = lane & 15, lk = lane >> 4;

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = T(0);

  typedef T pair_t __attribute__((ext_vector_type(2), aligned(sizeof(T))));
  pair_t ra[A_PP], rb[B_PP];
  const T* __restrict__ A = g.A + blockIdx.y * g.sA + blockIdx.z * g.sA2;
  const T* __restrict__ B = g.B + blockIdx.y * g.sB + blockIdx.z * g.sB2;
  const int M = g.M, N = g.N, K = g.K;
  const long lda = g.lda, ldb = g.ldb;
  // tiles inside M x N take slabs inside K without bounds checks
  const bool inner_mn = AL || (m0 + BM <= M && n0 + BN <= N);

  // element e of this thread's share of a slab: pair start (i, kk) in the
  // operand's (m or n, k) coordinates; the pair runs along i when the
  // operand is contiguous in i (CI), else along kk
  auto coords = [&](int e, int rows, bool ci, int& i, int& kk) {
    const int idx = tid + e * NT;
    if (ci) {
      i = 2 * (idx % (rows / 2));
      kk = idx / (rows / 2);
    } else {
      kk = 2 * (idx % (BK / 2));
      i = idx / (BK / 2);
    }
  };
  auto fetch = [&](const T* X, long ldx, bool ci, int i0, int lim, int i, int k, bool inner) {
    const T* p = ci ? X + (i0 + i) + (long)k * ldx : X + k + (long)(i0 + i) * ldx;
    if (inner) return *(const pair_t*)p;
    pair_t v = {T(0), T(0)};
    const long step = ci ? 1 : ldx;
    if (ci) {
      if (k < K) {
        if (i0 + i < lim) v.x = p[0];
        if (i0 + i + 1 < lim) v.y = p[step];
      }
    } else if (i0 + i < lim) {
      if (k < K) v.x = p[0];
      if (k + 1 < K) v.y = p[1];
    }
    return v;
  };
  // offsets of this thread's pairs at k = 0 (for full slabs: the slab's
  // start then adds one uniform term instead of a 64-bit product per pair)
  long offA[A_PP], offB[B_PP];
#pragma unroll
  for (int e = 0; e < A_PP; ++e) {
    int i, kk;
    coords(e, BM, !TA, i, kk);
    offA[e] = !TA ? (m0 + i) + (long)kk * lda : kk + (long)(m0 + i) * lda;
  }
#pragma unroll
  for (int e = 0; e < B_PP; ++e) {
    int j, kk;
    coords(e, BN, TB, j, kk);
    offB[e] = TB ? (n0 + j) + (long)kk * ldb : kk + (long)(n0 + j) * ldb;
  }
  auto load_to = [&](int k0, pair_t(&xa)[A_PP], pair_t(&xb)[B_PP]) {
    const bool inner = AL || (inner_mn && k0 + BK <= K);
    if (inner) {
      const T* Ak = A + (!TA ? (long)k0 * lda : (long)k0);
      const T* Bk = B + (TB ? (long)k0 * ldb : (long)k0);
#pragma unroll
      for (int e = 0; e < A_PP; ++e) xa[e] = *(const pair_t*)(Ak + offA[e]);
#pragma unroll
      for (int e = 0; e < B_PP; ++e) xb[e] = *(const pair_t*)(Bk + offB[e]);
      return;
    }
#pragma unroll
    for (int e = 0; e < A_PP; ++e) {
      int i, kk;
      coords(e, BM, !TA, i, kk);
      xa[e] = fetch(A, lda, !TA, m0, M, i, k0 + kk, inner);
    }
#pragma unroll
    for (int e = 0; e < B_PP; ++e) {
      int j, kk;
      coords(e, BN, TB, j, kk);
      xb[e] = fetch(B, ldb, TB, n0, N, j, k0 + kk, inner);
    }
  };
  auto store_from = [&](int buf, const pair_t(&xa)[A_PP], const pair_t(&xb)[B_PP]) {
#pragma unroll
    for (int e = 0; e < A_PP; ++e) {
      int i, kk;
      coords(e, BM, !TA, i, kk);
      if (!TA) {
        *(pair_t*)&As[buf][posA(kk, i)] = xa[e];
      } else {
        As[buf][posA(kk, i)] = xa[e].x;
        As[buf][posA(kk + 1, i)] = xa[e].y;
      }
    }
#pragma unroll
    for (int e = 0; e < B_PP; ++e) {
      int j, kk;
      coords(e, BN, TB, j, kk);
      if (TB) {
        *(pair_t*)&Bs[buf][posB(kk, j)] = xb[e];
      } else {
        Bs[buf][posB(kk, j)] = xb[e].x;
        Bs[buf][posB(kk + 1, j)] = xb[e].y;
      }
    }
  };
  auto load = [&](int k0) { load_to(k0, ra, rb); };
  auto store = [&](int buf) { store_from(buf, ra, rb); };
  // the MFMAs of one slab (this k share's k steps) from LDS buffer buf
  auto compute = [&](int buf) {
    const T* as = &As[buf][0];
    const T* bs = &Bs[buf][0];
#pragma unroll
    for (int kk = kh * (BK / KS); kk < (kh + 1) * (BK / KS); kk += 4) {
      T af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = as[posA(kk + lk, wm * WM + i * 16 + lr)];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = bs[posB(kk + lk, wn * WN + j * 16 + lr)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<T>::op(bf[j], af[i], acc[i][j]);
    }
  };

  if constexpr (PF2) {
    // two slabs in flight: slab s+2 is fetched into one register set while
    // the MFMAs run on slab s (LDS) and the other set (slab s+1) goes to
    // the other LDS buffer
    pair_t ra2[A_PP], rb2[B_PP];
    if (kbeg < kend) {
      load_to(kbeg, ra, rb);
      store_from(0, ra, rb);
      if (kbeg + BK < kend) load_to(kbeg + BK, ra2, rb2);
      __syncthreads();
      // Pairs of slabs (LDS buffer 0, then 1) while both fetches of an
      // iteration exist: no fetch sits under a condition there (the compiler
      // cannot count the loads younger than a conditionally fetched set, and
      // its wait then also covers the set fetched just before: one slab in
      // flight instead of two), and the loop has one exit (an exit between
      // the halves made it copy the accumulators every iteration). The last
      // one to three slabs follow without fetches past the end (fetching a
      // slab again instead slowed the short-k products of the LU panel).
      // DGEMM 1024^3 62 -> 56 us, n = 7680, k = 512: 49 -> 54 TFLOP/s.
      // Scheduling barriers that pin each half to fetch, MFMAs, LDS stores
      // were slower (61.5 us).
      int k0 = kbeg;
      for (; k0 + 3 * BK < kend; k0 += 2 * BK) {
        load_to(k0 + 2 * BK, ra, rb);
        compute(0);
        store_from(1, ra2, rb2);
        __syncthreads();
        load_to(k0 + 3 * BK, ra2, rb2);
        compute(1);
        store_from(0, ra, rb);
        __syncthreads();
      }
      if (k0 + 2 * BK < kend) {  // three slabs: k0 (buffer 0), k0 + BK (ra2), k0 + 2 BK
        load_to(k0 + 2 * BK, ra, rb);
        compute(0);
        store_from(1, ra2, rb2);
        __syncthreads();
        compute(1);
        store_from(0, ra, rb);
        __syncthreads();
        compute(0);
      } else if (k0 + BK < kend) {  // two slabs
        compute(0);
        store_from(1, ra2, rb2);
        __syncthreads();
        compute(1);
      } else {
        compute(0);
      }
    }
  } else if (kbeg < kend) {
    load(kbeg);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
      const bool more = k0 + BK < kend;
      if (more) load(k0 + BK);
      compute(buf);
      if (more) store(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }

  if constexpr (KS > 1) {
    // the k shares' partial tiles through LDS (free after the loop's last
    // barrier), added by the lead share
    __syncthreads();
    if (!lead) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            smem[((((kh - 1) * FM + i) * FN + j) * 4 + r) * NTS + stid] = acc[i][j][r];
    }
    __syncthreads();
    if (lead) {
#pragma unroll
      for (int h = 1; h < KS; ++h)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              acc[i][j][r] += smem[((((h - 1) * FM + i) * FN + j) * 4 + r) * NTS + stid];
    }
  }

  if (g.kchunk > 0) {
    // partial tiles in the register layout (coalesced both ways); the
    // reducing chunk is the last in dispatch order, so the chunks it waits
    // for were all dispatched before it and never wait themselves
    const int z = blockIdx.y, S = gridDim.y;
    constexpr int TILE = BM * BN;
    const long tid_tile = (long)tm * g.tiles_n + tn;  // slot of this output tile
    const long ntile = (long)g.tiles_m * g.tiles_n;
    if (z < S - 1) {
      T* P = g.part + ((long)z * ntile + tid_tile) * TILE;
      if (lead) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) P[((i * FN + j) * 4 + r) * NTS + stid] = acc[i][j][r];
      }
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (tid == 0)
        __hip_atomic_store(&g.kflags[tid_tile * S + z], g.kready, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    if (tid == 0) {
      for (int zz = 0; zz < S - 1; ++zz) {
        for (unsigned spins = 0;; ++spins) {
          if (spins >= g.spin_limit) {
            // give up rather than hang; the host reads the status word
            raise_status(g.status, STATUS_GEMM_WAIT);
            break;
          }
          if (__hip_atomic_load(&g.kflags[tid_tile * S + zz], __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT) == g.kready)
            break;
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    if (lead) {
      for (int zz = 0; zz < S - 1; ++zz) {
        const T* P = g.part + ((long)zz * ntile + tid_tile) * TILE;
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] += P[((i * FN + j) * 4 + r) * NTS + stid];
      }
    }
  }

  if (!lead) return;  // no barrier follows

  // ---- epilogue: lane holds C[m0+..+lr][n0+..+drow(lk, r)] ----
  // Processed EPI row blocks at a time: all loads of C for the chunk are
  // issued before its first store, so their latencies overlap (a load
  // after a store to the same array would otherwise wait for it).
  const T alpha = g.alpha, beta = g.beta;
  T* __restrict__ C = g.C + blockIdx.y * g.sC + blockIdx.z * g.sC2;
  const long ldc = g.ldc;
  const bool lower = g.mode == GM_LOWER || g.mode == GM_LOWER_T;
  const bool upper = g.mode == GM_UPPER || g.mode == GM_UPPER_T;
  constexpr int EPI = FM >= 2 ? 2 : 1;
#pragma unroll
  for (int i0 = 0; i0 < FM; i0 += EPI) {
    bool ok[EPI][FN][4];
    T cv[EPI][FN][4];
#pragma unroll
    for (int ii = 0; ii < EPI; ++ii) {
      const int gm = m0 + wm * WM + (i0 + ii) * 16 + lr;
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gn = n0 + wn * WN + j * 16 + Mfma<T>::drow(lk, r);
          ok[ii][j][r] = (AL || (gm < M && gn < N)) && !(lower && gm < gn) && !(upper && gm > gn);
          cv[ii][j][r] = T(0);
          if (beta != T(0) && ok[ii][j][r]) cv[ii][j][r] = C[gm + (long)gn * ldc];
        }
    }
#pragma unroll
    for (int ii = 0; ii < EPI; ++ii) {
      const int gm = m0 + wm * WM + (i0 + ii) * 16 + lr;
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gn = n0 + wn * WN + j * 16 + Mfma<T>::drow(lk, r);
          if (ok[ii][j][r]) C[gm + (long)gn * ldc] = alpha * acc[i0 + ii][j][r] + beta * cv[ii][j][r];
        }
    }
  }
}

__global__ void k_zero_flags_i32(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

template <typename T>
__global__ void k_scale_tile(int M, int N, T beta, T* C, long ldc) {
  const long total = (long)M * N;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % M), j = (int)(t / M);
    T* p = C + i + (long)j * ldc;
    *p = beta == T(0) ? T(0) : beta * *p;
  }
}

template <typename S, typename D>
__global__ void k_copy_tile(int M, int N, const S* A, long lda, D* B, long ldb) {
  const long total = (long)M * N;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % M), j = (int)(t / M);
    B[i + (long)j * ldb] = (D)A[i + (long)j * lda];
  }
}

inline int grid_for(long total) {
  long b = (total + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

// Scratch of a split-k launch: partial tiles and chunk flags. The partial
// tiles are carved from the calling thread's workspace and the flags taken
// from its flag pool (ready marks a set flag; nothing to clear); with a
// SplitScratchOverride in scope both come from its buffer, whose flags the
// launch clears first (ready = 1, clear = true). False when they do not fit.
struct SplitScratch {
  void* part;
  int* flags;
  int ready;
  bool clear;
};
thread_local void* g_split_base = nullptr;
thread_local size_t g_split_bytes = 0;

bool split_scratch(long part_bytes, long nflags, SplitScratch& sc) {
  const size_t pb = align_up((size_t)part_bytes, 256), fb = align_up((size_t)nflags * 4, 256);
  if (g_split_base) {
    if (pb + fb > g_split_bytes) return false;
    sc.part = g_split_base;
    sc.flags = (int*)((char*)g_split_base + pb);
    sc.ready = 1;
    sc.clear = true;
    return true;
  }
  sc.part = ws_carve(pb);
  sc.clear = false;
  return sc.part && signal_flags(nflags, &sc.flags, &sc.ready);
}

// Tile order of a product (the k range of a tile follows from the known
// zeros of its operands, so the heavy tiles sit in known rows, columns or
// shells).
int tile_order(int mode, int tri) {
  if (mode == GM_FULL) {
    switch (tri) {
      case TRI_A_LOWER: return TILE_ROWS_DOWN;   // k < m0 + BM
      case TRI_A_UPPER: return TILE_ROWS_UP;     // k >= m0
      case TRI_B_LOWER: return TILE_COLS_LEFT;   // k >= n0
      case TRI_B_UPPER: return TILE_COLS_RIGHT;  // k < n0 + BN
      case TRI_AU_BL: return TILE_SHELLS;        // k >= max(m0, n0)
      default: return TILE_GROUPED;
    }
  }
  if (mode == GM_LOWER) {
    if (tri == TRI_A_UPPER) return TILE_LOWER_ROWS;    // k >= m0
    if (tri == TRI_AB_LOWER) return TILE_LOWER_BANDS;  // n0 <= k < m0 + BM
  }
  return TILE_GROUPED;
}

template <typename T, int BM, int BN, int BK, int WGM, int WGN, int KS>
int launch_gemm(hipStream_t s, bool ta, bool tb, GemmArgs<T> g, int batch, int batch2) {
  g.tiles_m = ceil_div(g.M, BM);
  g.tiles_n = ceil_div(g.N, BN);
  // shells and the lower-triangle orders need a square tile grid
  if (g.order >= TILE_SHELLS && g.order <= TILE_LOWER_BANDS && g.tiles_m != g.tiles_n) {
    if (g.order != TILE_SHELLS) {
      set_error("gemm: triangle order on a non-square output");
      return ST_BADARG;
    }
    g.order = TILE_GROUPED;
  }
  // a lower trapezoid whose launch would be mostly tiles above the diagonal
  // (which exit at once, and all in the first XCDs' share of the grouped
  // order): launch only the tiles below it
  if (g.order == TILE_GROUPED && g.mode == GM_LOWER_T) {
    const long all = (long)g.tiles_m * g.tiles_n;
    const long tm = g.tiles_m, tn = g.tiles_n;
    const long work = tm <= tn ? tm * (tm + 1) / 2 : tn * (tn + 1) / 2 + (tm - tn) * tn;
    if (4 * all > 5 * work) g.order = TILE_LOWER_TRAP;
  }
  long nblk;
  if (g.order == TILE_LOWER_ROWS || g.order == TILE_LOWER_BANDS) {
    nblk = (long)g.tiles_m * (g.tiles_m + 1) / 2;
  } else if (g.order == TILE_LOWER_TRAP) {
    const long tm = g.tiles_m, tn = g.tiles_n;
    nblk = tm <= tn ? tm * (tm + 1) / 2 : tn * (tn + 1) / 2 + (tm - tn) * tn;
  } else if (g.order != TILE_GROUPED || g.mode == GM_FULL || g.mode == GM_LOWER_T ||
             g.mode == GM_UPPER_T) {
    nblk = (long)g.tiles_m * g.tiles_n;
  } else {
    const long ts = ceil_div(g.tiles_m, SG);
    nblk = ts * (ts + 1) / 2 * SG * SG;
  }
  if (nblk <= 0) return ST_OK;
  if (nblk > 0x7fffffff) {
    set_error("gemm: grid too large");
    return ST_BADARG;
  }
  dim3 grid((unsigned)nblk, (unsigned)batch, (unsigned)batch2), block(64 * WGM * WGN * KS);
  const bool tri = g.order == TILE_GROUPED && (g.mode == GM_LOWER || g.mode == GM_UPPER);
  const bool al = g.M % BM == 0 && g.N % BN == 0 && g.K % BK == 0 && g.kchunk % BK == 0;
#define MOPS_GEMM_KERNEL(TA_, TB_, TRI_, AL_) k_gemm<T, BM, BN, BK, WGM, WGN, KS, TA_, TB_, TRI_, AL_>
#define MOPS_GEMM_LAUNCH(TA_, TB_)                                                               \
  if (tri && al)                                                                                \
    hipLaunchKernelGGL((MOPS_GEMM_KERNEL(TA_, TB_, true, true)), grid, block, 0, s, g);         \
  else if (tri)                                                                                 \
    hipLaunchKernelGGL((MOPS_GEMM_KERNEL(TA_, TB_, true, false)), grid, block, 0, s, g);        \
  else if (al)                                                                                  \
    hipLaunchKernelGGL((MOPS_GEMM_KERNEL(TA_, TB_, false, true)), grid, block, 0, s, g);        \
  else                                                                                          \
    hipLaunchKernelGGL((MOPS_GEMM_KERNEL(TA_, TB_, false, false)), grid, block, 0, s, g);
  if (!ta && !tb) {
    MOPS_GEMM_LAUNCH(false, false)
  } else if (ta && !tb) {
    MOPS_GEMM_LAUNCH(true, false)
  } else if (!ta && tb) {
    MOPS_GEMM_LAUNCH(false, true)
  } else {
    MOPS_GEMM_LAUNCH(true, true)
  }
#undef MOPS_GEMM_LAUNCH
#undef MOPS_GEMM_KERNEL
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

template <typename T>
int gemm_impl(hipStream_t s, bool ta, bool tb, int M, int N, int K, T alpha, const T* A,
              long lda, const T* B, long ldb, T beta, T* C, long ldc, int mode, int tri,
              int batch = 1, long sA = 0, long sB = 0, long sC = 0, int batch2 = 1,
              long sA2 = 0, long sB2 = 0, long sC2 = 0) {
  if (M < 0 || N < 0 || K < 0) {
    set_error("gemm: negative dimension");
    return ST_BADARG;
  }
  if (M == 0 || N == 0 || batch <= 0 || batch2 <= 0) return ST_OK;
  if (batch > 65535 || batch2 > 65535) {
    set_error("gemm: batch too large");
    return ST_BADARG;
  }
  if ((mode == GM_LOWER || mode == GM_UPPER) && M != N) {
    set_error("gemm: triangular update needs a square C");
    return ST_BADARG;
  }
  // Only an empty inner dimension skips the product: with alpha == 0 the
  // product is still formed, so Inf/NaN in op(A) or op(B) reach C as they
  // do through calgo's dmult_mm_blocked3 (calgo/dmmat_gemm.c:297,413).
  if (K == 0) {
    if (beta == T(1)) return ST_OK;
    if (batch != 1 || batch2 != 1) {
      set_error("gemm: batched scale-only call is not supported");
      return ST_BADARG;
    }
    if (mode != GM_FULL) {
      // triangle-only scale: rare, fall back to per-column launches
      const bool lo = mode == GM_LOWER || mode == GM_LOWER_T;
      for (int j = 0; j < N; ++j) {
        int r0 = lo ? j : 0;
        int rows = lo ? M - j : j + 1;
        if (lo && j >= M) break;
        if (!lo && rows > M) rows = M;
        hipLaunchKernelGGL(k_scale_tile<T>, dim3(grid_for(rows)), dim3(256), 0, s, rows, 1,
                           beta, C + r0 + (long)j * ldc, ldc);
      }
    } else {
      hipLaunchKernelGGL(k_scale_tile<T>, dim3(grid_for((long)M * N)), dim3(256), 0, s, M, N,
                         beta, C, ldc);
    }
    MOPS_CHECK_HIP(hipGetLastError());
    return ST_OK;
  }
  GemmArgs<T> g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.C = C;
  g.ldc = ldc;
  g.alpha = alpha;
  g.beta = beta;
  g.mode = mode;
  g.tri = tri;
  g.sA = sA;
  g.sB = sB;
  g.sC = sC;
  g.sA2 = sA2;
  g.sB2 = sB2;
  g.sC2 = sC2;
  g.kchunk = 0;
  g.order = TILE_GROUPED;
  g.part = nullptr;
  g.kflags = nullptr;
  g.kready = 0;
  g.status = nullptr;
  g.spin_limit = 0;
  // Few output tiles and a long k: split k so that the grid still covers
  // the GPU (the partial tiles are added in a fixed order, so the result
  // does not depend on timing). Needs workspace; without it, no split.
  int S = 0, kc = 0;
  if ((mode == GM_FULL || mode == GM_LOWER || mode == GM_UPPER) && tri == TRI_NONE &&
      batch == 1 && batch2 == 1 && gemm_split_k(M, N, K, S, kc)) {
    // skinny products (the residual of the refinement, 16 right-hand
    // sides): 128x16 tiles with 32-deep slabs, two k shares and two slabs in
    // flight keep 64 KiB of A in flight per CU (one workgroup per CU); k is
    // split in S2 chunks so that the tm x S2 workgroups fill whole rounds of
    // the CUs. The plan is made before the flags are taken, so that exactly
    // the flags this launch polls are taken (and, in a caller's buffer,
    // cleared).
    int S2 = 0, kc2 = 0;
    if (mode == GM_FULL && gemm_skinny(M, N, K)) {
      int dev = 0, ncu = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          ncu <= 0)
        ncu = 256;
      S2 = gemm_skinny_chunks(M, K, S, ncu, kc2);
    }
    long part_elems = 0, nflags = 0;
    gemm_split_need(M, N, K, part_elems, nflags);
    const long used_flags = S2 >= 2 ? (long)ceil_div(M, 128) * S2
                                    : (long)ceil_div(M, 64) * ceil_div(N, 64) * S;
    const size_t mark = ws_mark();
    SplitScratch sc;
    int* status = active_status();
    if (split_scratch(part_elems * (long)sizeof(T), nflags, sc) && status &&
        used_flags <= nflags) {
      if (sc.clear) {
        hipLaunchKernelGGL(k_zero_flags_i32, dim3(grid_for(used_flags)), dim3(256), 0, s,
                           sc.flags, (int)used_flags);
        MOPS_CHECK_HIP(hipGetLastError());
      }
      g.part = (T*)sc.part;
      g.kflags = sc.flags;
      g.kready = sc.ready;
      g.status = status;
      g.spin_limit = wait_spin_limit(1u << 26);
      int st;
      if (S2 >= 2) {
        g.kchunk = kc2;
        st = launch_gemm<T, 128, 16, 32, 2, 1, 2>(s, ta, tb, g, S2, 1);
      } else {
        g.kchunk = kc;
        st = (mode == GM_FULL) ? launch_gemm<T, 64, 64, 16, 2, 2, SMALL_KS>(s, ta, tb, g, S, 1)
                               : launch_gemm<T, 64, 64, 16, 2, 2, 1>(s, ta, tb, g, S, 1);
      }
      ws_release(mark);
      return st;
    }
    ws_release(mark);
  }
  // Products with a triangular operand: heaviest tiles first, spread over
  // the XCDs. A lower-triangle output with a triangular operand is a
  // trapezoid launch over the lower tiles (element mask on the diagonal).
  g.order = tile_order(mode, tri);
  if (g.order == TILE_LOWER_ROWS || g.order == TILE_LOWER_BANDS) {
    g.mode = GM_LOWER_T;
    if (gemm_big_tiles(M, N, (long)batch * batch2))
      return launch_gemm<T, 128, 128, 16, 2, 4, 1>(s, ta, tb, g, batch, batch2);
    return launch_gemm<T, 64, 64, 16, 2, 2, SMALL_KS>(s, ta, tb, g, batch, batch2);
  }
  // Large outputs: 128x128 tiles (8 MFMA blocks per wave, 8 waves). Smaller
  // ones, and the triangle-only updates (whose diagonal tiles do wasted
  // work and whose grids balance better in finer pieces), use 64x64 tiles;
  // full and trapezoid outputs with two k shares (8 waves per tile).
  if (mode != GM_LOWER && mode != GM_UPPER && gemm_big_tiles(M, N, (long)batch * batch2))
    return launch_gemm<T, 128, 128, 16, 2, 4, 1>(s, ta, tb, g, batch, batch2);
  // triangle-only updates keep one k share: with 4 workgroups of 4 waves per
  // CU their large grids are faster than with 2 of 8 (measured at 4000-8000)
  if (mode == GM_LOWER || mode == GM_UPPER)
    return launch_gemm<T, 64, 64, 16, 2, 2, 1>(s, ta, tb, g, batch, batch2);
  return launch_gemm<T, 64, 64, 16, 2, 2, SMALL_KS>(s, ta, tb, g, batch, batch2);
}

}  // namespace

SplitScratchOverride::SplitScratchOverride(void* base, size_t bytes)
    : prev_base_(g_split_base), prev_bytes_(g_split_bytes) {
  g_split_base = base;
  g_split_bytes = bytes;
}

SplitScratchOverride::~SplitScratchOverride() {
  g_split_base = prev_base_;
  g_split_bytes = prev_bytes_;
}

int gemm_d(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha, const double* A,
           long lda, const double* B, long ldb, double beta, double* C, long ldc, int mode,
           int tri) {
  return gemm_impl<double>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mode, tri);
}

int gemm_s(hipStream_t s, bool ta, bool tb, int M, int N, int K, float alpha, const float* A,
           long lda, const float* B, long ldb, float beta, float* C, long ldc, int mode,
           int tri) {
  return gemm_impl<float>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, mode, tri);
}

int gemm_batched_d(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                   const double* A, long lda, long sA, const double* B, long ldb, long sB,
                   double beta, double* C, long ldc, long sC, int batch, int tri) {
  return gemm_impl<double>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, GM_FULL, tri,
                           batch, sA, sB, sC);
}

int gemm_batched2_d(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                    const double* A, long lda, long sA, long sA2, const double* B, long ldb,
                    long sB, long sB2, double beta, double* C, long ldc, long sC, long sC2,
                    int batch, int batch2, int tri) {
  return gemm_impl<double>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, GM_FULL, tri,
                           batch, sA, sB, sC, batch2, sA2, sB2, sC2);
}

int gemm_batched_s(hipStream_t s, bool ta, bool tb, int M, int N, int K, float alpha,
                   const float* A, long lda, long sA, const float* B, long ldb, long sB,
                   float beta, float* C, long ldc, long sC, int batch, int tri) {
  return gemm_impl<float>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, GM_FULL, tri,
                          batch, sA, sB, sC);
}

int scale_tile_d(hipStream_t s, int M, int N, double beta, double* C, long ldc) {
  if (M <= 0 || N <= 0 || beta == 1.0) return ST_OK;
  hipLaunchKernelGGL(k_scale_tile<double>, dim3(grid_for((long)M * N)), dim3(256), 0, s, M, N,
                     beta, C, ldc);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

int copy_tile_d(hipStream_t s, int M, int N, const double* A, long lda, double* B, long ldb) {
  if (M <= 0 || N <= 0) return ST_OK;
  hipLaunchKernelGGL((k_copy_tile<double, double>), dim3(grid_for((long)M * N)), dim3(256), 0, s,
                     M, N, A, lda, B, ldb);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

int copy_tile_s(hipStream_t s, int M, int N, const float* A, long lda, float* B, long ldb) {
  if (M <= 0 || N <= 0) return ST_OK;
  hipLaunchKernelGGL((k_copy_tile<float, float>), dim3(grid_for((long)M * N)), dim3(256), 0, s,
                     M, N, A, lda, B, ldb);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

int convert_d2s(hipStream_t s, int M, int N, const double* A, long lda, float* B, long ldb) {
  if (M <= 0 || N <= 0) return ST_OK;
  hipLaunchKernelGGL((k_copy_tile<double, float>), dim3(grid_for((long)M * N)), dim3(256), 0, s,
                     M, N, A, lda, B, ldb);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

int convert_s2d(hipStream_t s, int M, int N, const float* A, long lda, double* B, long ldb) {
  if (M <= 0 || N <= 0) return ST_OK;
  hipLaunchKernelGGL((k_copy_tile<float, double>), dim3(grid_for((long)M * N)), dim3(256), 0, s,
                     M, N, A, lda, B, ldb);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

}  // namespace mops
