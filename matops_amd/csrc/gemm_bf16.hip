// FP32-in, BF16-multiply GEMM for the BF16 form of the mixed-precision LU:
//
//   C (M x N, FP32) += alpha * A (M x K, FP32) * B (K x N, FP32)
//
// with A scaled by a power of two 2^-e before it is rounded to BF16 (e from
// the largest |A| of the call, read on the device, so the product of a
// panel's U12 block keeps its full 8-bit mantissa range whatever its
// magnitude) and the product multiplied back by 2^e in the epilogue. B (the
// panel's L21, |entries| <= 1 under partial pivoting) is rounded unscaled.
// Products run on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// Tiles of 128 x 128, k slabs of 32, 256 threads (2 x 2 waves of 64 x 64,
// 4 x 4 MFMA blocks each). A slab is loaded as FP32 (each thread 16
// consecutive k of one row of A or one column of B: coalesced along the
// rows of A, one 64-byte run per column of B), rounded to BF16 and staged
// in LDS as [row][k] and [col][k] with an 80-byte pitch, so that the
// 16-byte fragment reads and writes of 16 lanes hit all 64 banks once.
#include "mops_kernels.hpp"

namespace mops {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int HB_M = 128, HB_N = 128, HB_K = 32, HB_T = 256;
constexpr int HB_P = HB_K + 8;  // LDS pitch in bf16 (80 bytes)

__device__ __forceinline__ int hb_xcd_linear(int bid, int nblk) {
  const int x = bid % 8, q = nblk / 8, r = nblk % 8;
  return x * q + (x < r ? x : r) + bid / 8;
}

// *amax_bits holds max |A| as the bits of a non-negative float (or 0)
__global__ void k_absmax_f32(int M, int K, const float* __restrict__ A, long lda,
                             unsigned* amax_bits) {
  float m = 0.f;
  const long total = (long)M * K;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const float v = fabsf(A[(e % M) + (e / M) * lda]);
    m = (v > m || v != v) ? v : m;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const float w = __shfl_xor(m, o);
    m = (w > m || w != w) ? w : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(amax_bits, __float_as_uint(m));
}

__global__ void k_zero_u32_hb(unsigned* p) { *p = 0u; }

__global__ __launch_bounds__(HB_T) void k_gemm_bf16(int M, int N, int K, float alpha,
                                                    const float* __restrict__ A, long lda,
                                                    const unsigned* __restrict__ amax_bits,
                                                    const float* __restrict__ B, long ldb,
                                                    float* __restrict__ C, long ldc) {
  __shared__ __bf16 As[HB_M * HB_P];
  __shared__ __bf16 Bs[HB_N * HB_P];
  const int tiles_m = (M + HB_M - 1) / HB_M;
  const int lin = hb_xcd_linear(blockIdx.x, gridDim.x);
  const int tm = lin % tiles_m, tn = lin / tiles_m;
  const int m0 = tm * HB_M, n0 = tn * HB_N;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  // scale of A: 2^-e with e the exponent of max |A| (no scaling for 0,
  // Inf or NaN, which then propagate as they are)
  float sa = 1.f, inv_sa = 1.f;
  {
    const float amax = __uint_as_float(*amax_bits);
    if (amax > 0.f && amax <= 3.0e38f) {
      int e;
      frexpf(amax, &e);
      sa = ldexpf(1.f, -e);
      inv_sa = ldexpf(1.f, e);
    }
  }

  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  // staging: thread t owns row (A) / column (B) t % 128 and k half t / 128
  const int sr = t % 128, sk = (t / 128) * 16;
  float ra[16], rb[16];
  auto load = [&](int k0) {
    const int gm = m0 + sr, gn = n0 + sr;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int k = k0 + sk + e;
      ra[e] = (gm < M && k < K) ? A[gm + (long)k * lda] : 0.f;
      rb[e] = (gn < N && k < K) ? B[k + (long)gn * ldb] : 0.f;
    }
  };
  auto stage = [&]() {
    bf16x8 a0, a1, b0, b1;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      a0[e] = (__bf16)(ra[e] * sa);
      a1[e] = (__bf16)(ra[8 + e] * sa);
      b0[e] = (__bf16)rb[e];
      b1[e] = (__bf16)rb[8 + e];
    }
    *(bf16x8*)&As[sr * HB_P + sk] = a0;
    *(bf16x8*)&As[sr * HB_P + sk + 8] = a1;
    *(bf16x8*)&Bs[sr * HB_P + sk] = b0;
    *(bf16x8*)&Bs[sr * HB_P + sk + 8] = b1;
  };

  const int fr = lane & 15, fk = (lane >> 4) * 8;  // fragment row / column, k group
  load(0);
  for (int k0 = 0; k0 < K; k0 += HB_K) {
    __syncthreads();  // the previous slab's reads are done
    stage();
    __syncthreads();
    if (k0 + HB_K < K) load(k0 + HB_K);  // next slab in flight during the MFMAs
    bf16x8 af[4], bfr[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) af[i] = *(const bf16x8*)&As[(wm * 64 + i * 16 + fr) * HB_P + fk];
#pragma unroll
    for (int j = 0; j < 4; ++j) bfr[j] = *(const bf16x8*)&Bs[(wn * 64 + j * 16 + fr) * HB_P + fk];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
  }

  // lane holds D[4 (lane / 16) + r][lane % 16] of each 16 x 16 block: four
  // consecutive rows of one column of C
  const float f = alpha * inv_sa;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gn = n0 + wn * 64 + j * 16 + fr;
      const int gm = m0 + wm * 64 + i * 16 + (lane >> 4) * 4;
      if (gn >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (gm + r < M) C[gm + r + (long)gn * ldc] += f * acc[i][j][r];
    }
}

}  // namespace

int gemm_bf16_s(hipStream_t s, int M, int N, int K, float alpha, const float* A, long lda,
                const float* B, long ldb, float* C, long ldc, unsigned* amax_word) {
  if (M <= 0 || N <= 0 || K <= 0) return ST_OK;
  hipLaunchKernelGGL(k_zero_u32_hb, dim3(1), dim3(1), 0, s, amax_word);
  const long total = (long)M * K;
  long g = (total + 255) / 256;
  if (g > 2048) g = 2048;
  hipLaunchKernelGGL(k_absmax_f32, dim3((unsigned)g), dim3(256), 0, s, M, K, A, lda, amax_word);
  const long tiles = (long)((M + HB_M - 1) / HB_M) * ((N + HB_N - 1) / HB_N);
  if (tiles > 0x7fffffff) {
    set_error("gemm_bf16: grid too large");
    return ST_BADARG;
  }
  hipLaunchKernelGGL(k_gemm_bf16, dim3((unsigned)tiles), dim3(HB_T), 0, s, M, N, K, alpha, A,
                     lda, amax_word, B, ldb, C, ldc);
  MOPS_CHECK_HIP(hipGetLastError());
  return ST_OK;
}

}  // namespace mops

extern "C" int mops_d_gemm_bf16(int M, int N, int K, float alpha, const float* A, int lda,
                                const float* B, int ldb, float* C, int ldc, void* stream) {
  using namespace mops;
  clear_error();
  if (M < 0) return -1;
  if (N < 0) return -2;
  if (K < 0) return -3;
  if (lda < (M > 1 ? M : 1)) return -6;
  if (ldb < (K > 1 ? K : 1)) return -8;
  if (ldc < (M > 1 ? M : 1)) return -10;
  if (M == 0 || N == 0 || K == 0) return ST_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  WsUse wu(s);
  MOPS_TRY(ws_reserve(256));
  const size_t mark = ws_mark();
  unsigned* w = (unsigned*)ws_carve(256);
  if (!w) {
    set_error("gemm_bf16: workspace too small");
    return ST_NOMEM;
  }
  const int st = gemm_bf16_s(s, M, N, K, alpha, A, lda, B, ldb, C, ldc, w);
  ws_release(mark);
  return st;
}
