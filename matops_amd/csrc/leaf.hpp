// Register-resident kernels for one 64x64 diagonal block: Cholesky of the
// block and inverse of a triangular factor.
//
// Layout: the block's column c belongs to 4 consecutive lanes of one wave
// (thread t: column t / 4, lane-in-column sub = t % 4); lane sub holds rows
// sub, sub+4, ..., sub+60 in registers. Loops over the 64 columns are fully
// unrolled so every register index is a compile-time constant; values move
// inside a column's lanes with wave shuffles. The triangular inverse needs no
// barrier at all; the factorization needs one barrier per column, to publish
// the scaled column through LDS.
#pragma once

#include <cmath>

#include "mops_common.hpp"

namespace mops {
namespace leaf {

constexpr int NL = 64;   // block order
constexpr int LPC = 4;   // lanes per column
constexpr int RPL = 16;  // rows per lane
constexpr int LDL = NL + 1;

// Column c (rows sub + 4q) of inv(L), L lower triangular in LDS (row pitch
// LDL), order n <= 64. dinv[k] = 1 / L[k][k] (ignored when unit).
__device__ __forceinline__ void inv_lower_column(const double* Ls, const double* dinv, int n,
                                                 int c, int sub, bool unit, double (&x)[RPL]) {
  const int lane = threadIdx.x & 63;
  const int base = lane & ~(LPC - 1);
#pragma unroll
  for (int q = 0; q < RPL; ++q) x[q] = (sub + LPC * q == c) ? 1.0 : 0.0;
#pragma unroll
  for (int K = 0; K < RPL; ++K) {
#pragma unroll
    for (int sp = 0; sp < LPC; ++sp) {
      const int k = LPC * K + sp;
      if (k < n) {
        double xk = __shfl(x[K], base | sp, 64);
        xk = k < c ? 0.0 : (unit ? xk : xk * dinv[k]);
        if (sub == sp) x[K] = xk;
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
          const int r = sub + LPC * q;
          if (r > k && r < n) x[q] -= Ls[r * LDL + k] * xk;
        }
      }
    }
  }
}

}  // namespace leaf
}  // namespace mops
