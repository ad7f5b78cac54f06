// Device helpers shared by the eigensolver kernels (tridiag.hip, wide256.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace oakg {

__device__ __forceinline__ double wave_sum_d(double s) { return wave_sum_fast(s); }

// 1/sqrt(h) to full double precision: the hardware estimate y (relative
// error below 2^-22) and one third-order correction, y (1 + e/2 + 3e^2/8)
// with e = 1 - h y^2 (the neglected term 5e^3/16 is below 2^-64)
__device__ __forceinline__ double rsqrt_acc(double h) {
  const double y = __builtin_amdgcn_rsq(h);
  const double e = fma(-h, y * y, 1.0);
  return fma(y * e, fma(0.375, e, 0.5), y);
}
// 1/x the same way: y (1 + e + e^2) with e = 1 - x y
__device__ __forceinline__ double rcp_acc(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  const double e = fma(-x, y, 1.0);
  return fma(fma(e, e, e), y, y);
}

// Padded index of row r in the LDS row vectors of the Householder
// reductions: one pad entry per 8 rows (register tiles start every 8 rows).
__host__ __device__ constexpr int tri_rx(int r) { return r + (r >> 3); }

// ---------------------------------------------------------------------------
// Rotation log.  The zones of one 64-zone group share cap * 64 records; record
// k of the zone in lane l sits at (k / 4) * 256 + 4 l + (k % 4), so that one
// store instruction of the QL wave stays within a few KB.  A chain (one QL
// sweep on rows l..m) starts with a header {l + 65536 m, count} followed by
// `count` records {c, s} for the rotation indices i = m-1, m-2, ...
// ---------------------------------------------------------------------------
__device__ __forceinline__ size_t rlog_off(int zl, int64_t cap) {
  return (size_t)(zl >> 6) * (size_t)cap * 64 + (size_t)(zl & 63) * 4;
}
__device__ __forceinline__ size_t rlog_rec(int64_t k) { return (size_t)(k >> 2) * 256 + (size_t)(k & 3); }

}  // namespace oakg
