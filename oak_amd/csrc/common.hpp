// Shared device helpers for the gfx950 local-analysis kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Device-side error bits, OR-ed into a word the host reads after each call.
#define OAKG_DERR_RMSE 1u   // rmse^2 <= 0 or not finite
#define OAKG_DERR_NAN 2u    // NaN amplitude (rrsqrt.F90:145-149)
#define OAKG_DERR_CAP 4u    // selection list overflow
#define OAKG_DERR_IDX 8u    // observation/state index out of range
#define OAKG_DERR_NOCONV 16u  // QL did not finish a zone that has no fallback (N > 128)

// v_mfma_f64_16x16x4f64 operand layout on gfx950 (checked on the device):
//   A (16x4):  lane l holds A[l % 16][l / 16]
//   B (4x16):  lane l holds B[l / 16][l % 16]
//   D (16x16): lane l, register r holds D[l / 16 + 4 r][l % 16]
__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also
// waits for every outstanding global load and store of the wave (its fence
// covers global memory), which exposes the full memory latency at each
// barrier of a loop that streams results to HBM or prefetches from it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Column j of a row of the internal observation-anomaly matrix HS (m x NP)
// is stored at hs_col<NP>(j).  For NP = 128 the 8 columns 16 J + c (J = 0..7)
// that lane c of the Gram MFMA needs are contiguous (8 c + J), so a lane
// reads them with 16-byte loads; for NP <= 64 the order is the natural one.
template <int NP>
__host__ __device__ constexpr int hs_col(int j) {
  return NP == 128 ? 8 * (j & 15) + (j >> 4) : j;
}

__device__ __forceinline__ double shfl_xor_d(double v, int mask) {
  return __shfl_xor(v, mask, 64);
}

__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }

// Lane permutation of a double by DPP (two 32-bit moves, a few cycles; a
// __shfl goes through ds_bpermute, an LDS round trip per 32-bit half).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// st(e, ld(e)) for e = e0, e0 + step, ... < n with B loads in flight per
// thread: in a plain loop each load is waited for before its store.
template <int B, typename Ld, typename St>
__device__ __forceinline__ void copy_batched(int e0, int n, int step, Ld ld, St st) {
  for (int eb = e0; eb < n; eb += B * step) {
    double v[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int e = eb + u * step;
      v[u] = e < n ? ld(e) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int e = eb + u * step;
      if (e < n) st(e, v[u]);
    }
  }
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wave, returned (uniform) to every lane: quad
// swaps and row rotations leave each 16-lane row's sum in its lanes, then the
// four row sums are read into scalars and added.
__device__ __forceinline__ double wave_sum_fast(double v) {
  v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]: xor 1
  v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]: xor 2
  v += dpp_d<0x124>(v);  // row_ror 4: the next quad's sum
  v += dpp_d<0x128>(v);  // row_ror 8: the other half row
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

// OAK's distance (assimilation.F90:3635-3672), p0 = observation, p1 = zone.
// No contraction: the products and sums round exactly as OAK evaluates them,
// so the inclusive cutoff test d <= maxLength selects the same observations.
__device__ __forceinline__ double oak_distance(int metric, double p0x, double p0y, double p1x,
                                               double p1y) {
#pragma clang fp contract(off)
  const double pi = 3.141592653589793238462643383279502884197;
  const double R = 6378137.0;
  if (metric == 0) {
    double dx = p1x - p0x, dy = p1y - p0y;
    double s = dx * dx;
    s = s + dy * dy;
    return sqrt(s);
  } else if (metric == 2) {
    double coeff = pi * R / 180.;
    double a = coeff * cos((p0y + p1y) * (pi / 360.)) * (p1x - p0x);
    double b = coeff * (p1y - p0y);
    double s = a * a;
    s = s + b * b;
    return sqrt(s);
  } else {
    double d2r = pi / 180.;
    double a = p0y * d2r, b = p1y * d2r, C = (p1x - p0x) * d2r;
    double coeff = sin(b) * sin(a) + cos(b) * cos(a) * cos(C);
    coeff = fmax(fmin(coeff, 1.), -1.);
    return R * acos(coeff);
  }
}

// exp(-(d/L)^2), assimilation.F90:3767
__device__ __forceinline__ double oak_taper(double d, double L) {
#pragma clang fp contract(off)
  double q = d / L;
  return exp(-(q * q));
}
