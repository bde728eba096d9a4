// Register-resident kernels for one small diagonal block (NL <= 128):
// Cholesky of the block and inverse of a triangular factor.
//
// Layout: the block's column c belongs to LPC consecutive lanes of one wave
// (thread t: column t / LPC, lane-in-column sub = t % LPC); lane sub holds
// rows sub, sub+LPC, sub+2*LPC, ... in registers (RPL = NL / LPC of them).
// Values inside a column move between its lanes with wave shuffles, so the
// triangular inverse needs no barrier at all and the factorization needs
// one barrier per column (to publish the scaled column through LDS).
#pragma once

#include <cmath>

#include "mops_common.hpp"

namespace mops {
namespace leaf {

template <int RPL>
__device__ __forceinline__ double pick(const double (&v)[RPL], int q) {
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < RPL; ++i)
    if (i == q) r = v[i];
  return r;
}

template <int RPL>
__device__ __forceinline__ void put(double (&v)[RPL], int q, double x) {
#pragma unroll
  for (int i = 0; i < RPL; ++i)
    if (i == q) v[i] = x;
}

// Column c (rows sub + LPC*q) of inv(L), L lower triangular in LDS (row
// stride LDL). unit: the diagonal is taken as 1 and not read. n: order.
// Forward substitution L x = e_c with the solved component broadcast inside
// the column's lanes (divisions as in the reference's substitution).
template <int NL, int LPC, int LDL>
__device__ __forceinline__ void inv_lower_column(const double* Ls, int n, int c, int sub,
                                                 bool unit, double (&x)[NL / LPC]) {
  constexpr int RPL = NL / LPC;
  const int lane = threadIdx.x & 63;
  const int base = lane & ~(LPC - 1);
#pragma unroll
  for (int q = 0; q < RPL; ++q) x[q] = (sub + LPC * q == c) ? 1.0 : 0.0;
  for (int k = 0; k < n; ++k) {
    double xk = __shfl(pick<RPL>(x, k / LPC), base | (k % LPC), 64);
    if (k < c)
      xk = 0.0;  // above the diagonal of column c
    else if (!unit)
      xk = xk / Ls[k * LDL + k];
    if (sub == k % LPC) put<RPL>(x, k / LPC, xk);
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int r = sub + LPC * q;
      if (r > k && r < n) x[q] -= Ls[r * LDL + k] * xk;
    }
  }
}

}  // namespace leaf
}  // namespace mops
