// Host-side sweep of the split-k sizing rule of the GEMM (gemm_plan.hpp):
// for every M in [1, mmax] (step 1), the N values of the skinny and tile
// boundaries, every K of a grid and CU counts 8..512, the flags and partial
// tiles the chosen instance uses must fit in what gemm_split_need reports
// (what the entry points reserve and the kernel clears).
// Prints "ok <cases>" or the first violation and exits non-zero.
#include <cstdio>
#include <cstdlib>

#include "../../matops_amd/csrc/gemm_plan.hpp"

using namespace mops;

int main(int argc, char** argv) {
  if (argc == 6 && argv[1][0] == 'c') {
    // "case M N K ncu": the plan of one shape
    const long M = atol(argv[2]), N = atol(argv[3]), K = atol(argv[4]);
    const int ncu = atoi(argv[5]);
    int S = 0, kc = 0, kc2 = 0, S2 = 0;
    if (gemm_split_k(M, N, K, S, kc) && gemm_skinny(M, N, K))
      S2 = gemm_skinny_chunks(M, K, S, ncu, kc2);
    long pe = 0, nf = 0;
    gemm_split_need(M, N, K, pe, nf);
    const long tiles = ((M + 63) / 64) * ((N + 63) / 64), tm = (M + 127) / 128;
    printf("S=%d kc=%d S2=%d kc2=%d flags64=%ld flags_skinny=%ld need_flags=%ld need_part=%ld\n",
           S, kc, S2, kc2, tiles * S, tm * S2, nf, pe);
    return 0;
  }
  const long mmax = argc > 1 ? atol(argv[1]) : 40000;
  const long Ns[] = {1, 2, 5, 15, 16, 17, 31, 64, 65, 128, 300, 511};
  const long Ks[] = {500, 512, 1000, 2047, 2048, 2049, 4096, 5000, 8192, 16384, 20000, 40000};
  const int cus[] = {8, 64, 80, 104, 228, 256, 304, 512};
  long cases = 0;
  for (long M = 1; M <= mmax; ++M)
    for (long N : Ns)
      for (long K : Ks) {
        int S = 0, kc = 0;
        if (!gemm_split_k(M, N, K, S, kc)) continue;
        long pe = 0, nf = 0;
        gemm_split_need(M, N, K, pe, nf);
        const long tiles = ((M + 63) / 64) * ((N + 63) / 64);
        if ((long)(S - 1) * tiles * 4096 > pe || tiles * S > nf) {
          printf("64x64 split exceeds its sizing: M=%ld N=%ld K=%ld S=%d\n", M, N, K, S);
          return 1;
        }
        if ((long)S * kc < K || (long)(S - 1) * kc >= K) {
          printf("chunks do not cover k exactly once: M=%ld N=%ld K=%ld S=%d kc=%d\n", M, N, K,
                 S, kc);
          return 1;
        }
        ++cases;
        if (!gemm_skinny(M, N, K)) continue;
        for (int ncu : cus) {
          int kc2 = 0;
          const int S2 = gemm_skinny_chunks(M, K, S, ncu, kc2);
          if (S2 < 2) continue;
          const long tm = (M + 127) / 128;
          if ((long)(S2 - 1) * tm * 128 * 16 > pe || tm * S2 > nf) {
            printf("skinny split exceeds its sizing: M=%ld N=%ld K=%ld S=%d S2=%d cus=%d "
                   "flags %ld > %ld\n", M, N, K, S, S2, ncu, tm * S2, nf);
            return 1;
          }
          if ((long)S2 * kc2 < K || (long)(S2 - 1) * kc2 >= K || kc2 % 32) {
            printf("skinny chunks do not cover k: M=%ld K=%ld S2=%d kc2=%d\n", M, K, S2, kc2);
            return 1;
          }
          ++cases;
        }
      }
  printf("ok %ld\n", cases);
  return 0;
}
