// Mixed-precision ensemble analysis (FP32 state): the same steps as the FP64
// path (assimilation.F90:3107-3134, rrsqrt.F90:100-190, assimilation.F90:
// 3299-3320) with FP32 storage of the ensemble, of Ea and of S = H A', and
// the ensemble update A' W on FP32 MFMA.  The N x N eigensolve, the
// eigenvalue functions (1+lambda)^-1 and (1+lambda)^-1/2, the transform T,
// the innovation scaling and the accumulation of S^T R^-1 S stay in FP64.
//
//   k_state_mean_f32   xbar (FP64 sums of FP32 members)
//   k_spmm_obs_f32     HE = H E + Hshift (FP64 accumulation)
//   k_obs_anom_f32     HS = (HE - ybar)/sqrt(N-1) stored FP32, innov/rho FP64
//   k_gram<NP, float>  (analysis.hip / analysis128.hip) M = I + S^T diag(r) S
//                      from the FP32 S on FP64 MFMA: products of FP32 values
//                      are exact in FP64, so only the rounding of S to FP32
//                      remains.  An FP32-accumulated Gram has errors of size
//                      eps32 |C| in every direction, including the weakly
//                      observed ones where the analysis keeps the forecast
//                      spread (T ~ I); with |C| up to ~1e5 (precise SSH
//                      observations) that is ~1e-3 of the spread.  Rounding
//                      S itself perturbs those directions only at second order.
//   k_update_f32       (N <= 64; analysis128.hip's k_xform128f for NP = 128)
//                      Ea = xa + sqrt(N-1) infl (Sf T) on v_mfma_f32_16x16x4f32
//                      with T rounded to FP32; xa = xbar + clamp(Sf a) in FP64
//                      (the FP32 sum of N terms Sf_k a_k cancels strongly)
#include "analysis.hpp"
#include "common.hpp"

namespace oakg {

// v_mfma_f32_16x16x4f32 operand layout on gfx950 (checked on the device):
//   A (16x4):  lane l holds A[l % 16][l / 16]
//   B (4x16):  lane l holds B[l / 16][l % 16]
//   D (16x16): lane l, register r holds D[4 (l / 16) + r][l % 16]
// (the FP64 instruction places D rows at l / 16 + 4 r instead)
__device__ __forceinline__ f4 mfma_f32(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__global__ void k_state_mean_f32(int64_t n, int N, const float* __restrict__ E, int64_t ld,
                                 double* __restrict__ xbar) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0;
  for (int k = 0; k < N; ++k) s += (double)E[i + (int64_t)k * ld];
  xbar[i] = s / N;
}

// The same with four consecutive elements per thread (16-byte loads; E 16-byte
// aligned, ld and n multiples of 4): a quarter of the load instructions, the
// same sums.
__global__ void k_state_mean_f32x4(int64_t n4, int N, const float4* __restrict__ E4, int64_t ld4,
                                   double* __restrict__ xbar) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n4) return;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int k = 0; k < N; ++k) {
    const float4 v = E4[i + (int64_t)k * ld4];
    s0 += (double)v.x;
    s1 += (double)v.y;
    s2 += (double)v.z;
    s3 += (double)v.w;
  }
  double2* x2 = reinterpret_cast<double2*>(xbar + 4 * i);
  x2[0] = make_double2(s0 / N, s1 / N);
  x2[1] = make_double2(s2 / N, s3 / N);
}

__global__ void k_spmm_obs_f32(int64_t m, int64_t n, int N, const int64_t* __restrict__ row_ptr,
                               const int32_t* __restrict__ col, const double* __restrict__ val,
                               const float* __restrict__ E, int64_t ld,
                               const float* __restrict__ Hshift, double* __restrict__ HEc,
                               unsigned* __restrict__ err) {
  int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (l >= m) return;
  const float* Ek = E + (int64_t)k * ld;
  double s = 0;
  for (int64_t e = row_ptr[l]; e < row_ptr[l + 1]; ++e) {
    int32_t j = col[e];
    if (j < 0 || j >= n) { atomicOr(err, OAKG_DERR_IDX); continue; }
    s = fma(val[e], (double)Ek[j], s);
  }
  HEc[l + (int64_t)k * m] = s + (double)Hshift[l];
}

template <int NP, int TR>
__global__ __launch_bounds__(256) void k_obs_anom_f32(int64_t m, int N, const double* __restrict__ HEc,
                                                      const double* __restrict__ yo,
                                                      const double* __restrict__ rmse2,
                                                      const uint8_t* __restrict__ excl,
                                                      double scaling, float* __restrict__ HS,
                                                      double* __restrict__ innov,
                                                      double* __restrict__ rho,
                                                      unsigned* __restrict__ err) {
  __shared__ double tile[TR][NP + 1];
  __shared__ double ybar[TR];
  const int64_t l0 = (int64_t)blockIdx.x * TR;
  for (int e = threadIdx.x; e < TR * NP; e += 256) {
    int k = e / TR, li = e % TR;
    int64_t l = l0 + li;
    tile[li][k] = (l < m && k < N) ? HEc[l + (int64_t)k * m] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x < TR) {
    int li = threadIdx.x;
    int64_t l = l0 + li;
    double s = 0;
    for (int k = 0; k < N; ++k) s += tile[li][k];
    double yb = s / N;
    ybar[li] = yb;
    if (l < m) {
      double r2 = rmse2[l];
      if (!(r2 > 0) || !isfinite(r2)) atomicOr(err, OAKG_DERR_RMSE);
      innov[l] = yo[l] - yb;
      rho[l] = (excl && excl[l]) ? 0.0 : 1.0 / r2;
    }
  }
  __syncthreads();
  const double inv = 1.0 / scaling;
  for (int e = threadIdx.x; e < TR * NP; e += 256) {
    int li = e / NP, k = e % NP;
    int64_t l = l0 + li;
    if (l < m) HS[l * NP + hs_col<NP>(k)] = (k < N) ? (float)((tile[li][k] - ybar[li]) * inv) : 0.0f;
  }
}

// Ensemble update for FP32 members: Sf = (E - xbar)/sqrt(N-1) on the fly,
// xa = xbar + clamp(Sf a), Ea = xa + sqrt(N-1) infl (Sf T).  The MFMA computes
// the transposed product (A = [T|a]^T from LDS, B = Sf^T) so that each
// 16-row strip is stored as contiguous runs of the column-major Ea.
template <int NP>
__global__ __launch_bounds__(256) void k_update_f32(int64_t z0, const int32_t* __restrict__ zorder, int N,
                                                    const int64_t* __restrict__ zone_ptr,
                                                    const double* __restrict__ mat,
                                                    const double* __restrict__ vec,
                                                    const double* __restrict__ xbar,
                                                    const float* __restrict__ E, int64_t ldE,
                                                    float* __restrict__ Out, int64_t ldO,
                                                    double scaling, double infl, double maxcorr,
                                                    const double* __restrict__ maxcorr_v) {
  constexpr int NB = NP / 16;
  constexpr int LDW = NP + 17;  // the 4 k-rows of one MFMA step land on distinct banks
  __shared__ float Ws[NP * LDW];
  __shared__ double as[NP];
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  const int64_t i0 = zone_ptr[z];
  const int64_t sz = zone_ptr[z + 1] - i0;
  const double* T = mat + (size_t)zl * NP * NP;
  const double* a = vec + (size_t)zl * 4 * NP + NP;
  for (int e = threadIdx.x; e < NP * NP; e += 256) Ws[(e / NP) * LDW + e % NP] = (float)T[e];
  if (threadIdx.x < NP) as[threadIdx.x] = a[threadIdx.x] / scaling;
  __syncthreads();
  const float inv = (float)(1.0 / scaling);
  const double sc_infl = scaling * infl;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, kk = lane >> 4;
  for (int64_t rb = 16 * wave; rb < sz; rb += 64) {
    f4 acc[NB];
#pragma unroll
    for (int J = 0; J < NB; ++J) acc[J] = f4{0, 0, 0, 0};
    const int64_t row = rb + li;
    const bool rok = row < sz;
    const double xbd = rok ? xbar[i0 + row] : 0.0;
    const float xb = (float)xbd;
    const float* Er = E + (i0 + (rok ? row : 0));
    double dxp = 0.0;  // (E - xbar) a / sqrt(N-1) over this lane's members, FP64
#pragma unroll 4
    for (int k0 = 0; k0 < NP; k0 += 4) {
      const int kc = k0 + kk;
      float bval = 0.0f;
      if (rok && kc < N) {
        const float ev = Er[(int64_t)kc * ldE];
        bval = (ev - xb) * inv;
        dxp = fma((double)ev - xbd, as[kc], dxp);
      }
#pragma unroll
      for (int J = 0; J < NB; ++J) acc[J] = mfma_f32(Ws[kc * LDW + 16 * J + li], bval, acc[J]);
    }
    dxp += __shfl_xor(dxp, 16, 64);
    dxp += __shfl_xor(dxp, 32, 64);
    if (rok) {
      double xar = xbd + dxp;
      const double mc = maxcorr_v ? maxcorr_v[i0 + row] : maxcorr;
      if (xar - mc > xbd) xar = xbd + mc;
      if (xar < xbd - mc) xar = xbd - mc;
      float* Orow = Out + (i0 + row);
#pragma unroll
      for (int J = 0; J < NB; ++J) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int col = 16 * J + 4 * kk + r;
          if (col < N) Orow[(int64_t)col * ldO] = (float)(xar + sc_infl * (double)acc[J][r]);
        }
      }
    }
  }
}

static inline unsigned nblk32(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <int NP>
static void prep_ens_f32_t(const PrepEnsArgs& A, hipStream_t st) {
  if (A.xbar) {  // NULL: observation side only (the report)
    if ((uintptr_t)A.E32 % 16 == 0 && A.ldE % 4 == 0 && A.n % 4 == 0)
      hipLaunchKernelGGL(k_state_mean_f32x4, dim3(nblk32(A.n / 4, 256)), dim3(256), 0, st, A.n / 4,
                         A.N, reinterpret_cast<const float4*>(A.E32), A.ldE / 4, A.xbar);
    else
      hipLaunchKernelGGL(k_state_mean_f32, dim3(nblk32(A.n, 256)), dim3(256), 0, st, A.n, A.N,
                         A.E32, A.ldE, A.xbar);
  }
  if (A.m == 0) return;
  hipLaunchKernelGGL(k_spmm_obs_f32, dim3(nblk32(A.m, 256), A.N), dim3(256), 0, st, A.m, A.n, A.N,
                     A.row_ptr, A.col, A.val, A.E32, A.ldE, A.Hshift32, A.HEc, A.err);
  constexpr int TR = NP > 128 ? 16 : (NP > 64 ? 32 : 64);
  hipLaunchKernelGGL((k_obs_anom_f32<NP, TR>), dim3(nblk32(A.m, TR)), dim3(256), 0, st, A.m, A.N,
                     A.HEc, A.yo, A.rmse2, A.excl, A.scaling, A.HS32, A.innov, A.rho, A.err);
}

void launch_prep_ens_f32(int NP, const PrepEnsArgs& A, hipStream_t st) {
  switch (NP) {
    case 16: prep_ens_f32_t<16>(A, st); break;
    case 32: prep_ens_f32_t<32>(A, st); break;
    case 64: prep_ens_f32_t<64>(A, st); break;
    case 128: prep_ens_f32_t<128>(A, st); break;
    case 256: prep_ens_f32_t<256>(A, st); break;
  }
}

void launch_update_f32(int NP, const ChunkArgs& A, hipStream_t st) {
#define OAK_UPD32(P)                                                                             \
  hipLaunchKernelGGL(k_update_f32<P>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.N, A.zone_ptr, A.mat, \
                     A.vec, A.X0, A.S32, A.ldS, A.Out32, A.ldO, A.scaling, A.infl, A.maxcorr, \
                     A.maxcorr_v)
  switch (NP) {
    case 16: OAK_UPD32(16); break;
    case 32: OAK_UPD32(32); break;
    case 64: OAK_UPD32(64); break;
  }
#undef OAK_UPD32
}

}  // namespace oakg
