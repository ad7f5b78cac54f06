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

// v_mfma_f64_16x16x4f64 operand layout on gfx950 (checked on the device):
//   A (16x4):  lane l holds A[l % 16][l / 16]
//   B (4x16):  lane l holds B[l / 16][l % 16]
//   D (16x16): lane l, register r holds D[l / 16 + 4 r][l % 16]
__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double shfl_xor_d(double v, int mask) {
  return __shfl_xor(v, mask, 64);
}

__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }

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
