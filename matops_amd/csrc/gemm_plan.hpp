// Launch plan of the GEMM (tile size, split of k, skinny instance) and the
// workspace a split needs. Plain C++ with no HIP dependency, so that the
// sizing rules can be compiled and swept on the host.
#pragma once

namespace mops {

// Tile configuration rule of gemm_d / gemm_s: 128x128 tiles when the output
// has at least 512 of them, else 64x64 tiles (so the grid still covers the
// 256 CUs).
inline bool gemm_big_tiles(long M, long N, long batch = 1) {
  return ((M + 127) / 128) * ((N + 127) / 128) * batch >= 512;
}

// Split of k for outputs of few 64x64 tiles and a long k (fewer than 128
// tiles and k >= 512, or fewer than 512 tiles and k >= 2048): S chunks of
// kc (a multiple of 16, at least 128) so that S x tiles comes near 1024
// workgroups. Returns false when no split applies.
inline bool gemm_split_k(long M, long N, long K, int& S, int& kc) {
  const long tiles = ((M + 63) / 64) * ((N + 63) / 64);
  if (tiles <= 0 || tiles >= 512 || K < 512) return false;
  if (tiles >= 128 && K < 2048) return false;
  long s = (1024 + tiles - 1) / tiles;
  if (s > K / 128) s = K / 128;
  if (s > 32) s = 32;
  if (s < 2) return false;
  long c = (K + s - 1) / s;
  c = (c + 15) / 16 * 16;
  S = (int)((K + c - 1) / c);
  kc = (int)c;
  return S >= 2;
}

// Shapes whose split uses the skinny instance (128x16 tiles, 32-deep slabs):
// full outputs with N <= 16, M >= 2048 and K >= 2048.
inline bool gemm_skinny(long M, long N, long K) { return N <= 16 && M >= 2048 && K >= 2048; }

// Number of k chunks S2 of the skinny instance, given the split S of the
// 64x64 instance and the CU count: the tm = ceil(M/128) row tiles times S2
// chunks fill whole rounds of the CUs (cost: rounds x CUs / workgroups, plus
// 1% per chunk), 2 <= S2 <= 2 S and S2 <= K / 64; each chunk is a multiple
// of 32 deep. kc2 receives the chunk depth. Returns S2 (< 2: no split).
inline int gemm_skinny_chunks(long M, long K, int S, int ncu, int& kc2) {
  const long tm = (M + 127) / 128;
  if (ncu <= 0) ncu = 256;
  int S2 = 2;
  double best = 1e30;
  for (int c = 2; c <= 2 * S && c <= K / 64; ++c) {
    const long wg = tm * c;
    const double cost = (double)((wg + ncu - 1) / ncu) * ncu / (double)wg + 0.01 * c;
    if (cost < best) {
      best = cost;
      S2 = c;
    }
  }
  const long per = (K + S2 - 1) / S2;
  kc2 = (int)((per + 31) / 32 * 32);
  return (int)((K + kc2 - 1) / kc2);
}

// Partial-tile elements and flags the split of an M x N x K product may use,
// whichever instance runs it: the 64x64 instance uses (S-1) x tiles partial
// tiles of 64x64 and tiles x S flags; the skinny one (S2-1) x tm partial
// tiles of 128x16 and tm x S2 flags with S2 <= 2 S. Both are covered.
inline void gemm_split_need(long M, long N, long K, long& part_elems, long& nflags) {
  part_elems = 0;
  nflags = 0;
  int S = 0, kc = 0;
  if (!gemm_split_k(M, N, K, S, kc)) return;
  const long tiles = ((M + 63) / 64) * ((N + 63) / 64);
  part_elems = (long)(S - 1) * tiles * 64 * 64;
  nflags = tiles * S;
  if (gemm_skinny(M, N, K)) {
    const long tm = (M + 127) / 128;
    const long p2 = (long)(2 * S - 1) * tm * 128 * 16, f2 = tm * 2 * S;
    if (p2 > part_elems) part_elems = p2;
    if (f2 > nflags) nflags = f2;
  }
}

// Workspace the split of gemm_d / gemm_s may carve (bytes, for elements of
// elem_bytes, with room for the alignment of two carves); 0 when the shape
// is not split.
inline unsigned long gemm_split_ws_bytes(long M, long N, long K, unsigned long elem_bytes) {
  long pe = 0, nf = 0;
  gemm_split_need(M, N, K, pe, nf);
  if (nf == 0) return 0;
  return (unsigned long)pe * elem_bytes + (unsigned long)nf * sizeof(int) + 1024;
}

}  // namespace mops
