// Innovation negative log-likelihood of the local analysis and its adjoint.
//
// For zone z with local observations q (weights w_q = exp(-(d_q/L_z)^2),
// error variances sigma_q^2, r_q = w_q^2 / sigma_q^2, inflated observation
// anomalies s_q = alpha S_q, innovations d_q) the analysis solves with
// M = I + S'^T diag(r) S', b = S'^T (r o d), a = M^-1 b (rrsqrt.F90:135-142).
// The local innovation covariance D = S' S'^T + diag(1/r) gives, through the
// Woodbury identity and det D = det R det M,
//
//   J_z = 1/2 [ sum_q r_q d_q^2 - b^T a + sum_q log(1/r_q) + log det M ]
//
// (= 1/2 [d^T D^-1 d + log det D]).  Its derivatives, with t_q = s'_q . a and
// u_q = s'_q^T M^-1 s'_q:
//
//   dJ_z/dr_q   = 1/2 [ (d_q - t_q)^2 + u_q - 1/r_q ]
//   dJ_z/dalpha = ( sum_k (1 - 1/mu_k) - a^T a ) / alpha
//
// and r_q = w_q^2 exp(-2 log sigma_q), so dr_q/dlog sigma_q = -2 r_q and
// dr_q/dL_z = r_q (-4 log w_q) / L_z.  M^-1 = Y diag(mu^-1/2) Y^T with
// Y = U diag(mu^-1/4) as left by the eigensolve, so u_q = sum_k
// mu_k^-1/2 (Y^T s'_q)_k^2: one N x N product per observation, the same
// work as the Gram.  The observations with r_q = 0 (excluded) take no part.
#include "analysis.hpp"
#include "common.hpp"

namespace oakg {

__device__ __forceinline__ double nll_wave_sum(double s) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) s += shfl_xor_d(s, o);
  return s;
}

// One workgroup (4 waves) per zone; wave w takes the observations w, w+4, ...
// and lane j holds entries j, j+64, ... of the N-vectors.
template <int NP>
__global__ __launch_bounds__(256) void k_nll_adjoint(int64_t z0, const int32_t* __restrict__ zorder, int N,
                                                     const int64_t* __restrict__ sel_ptr,
                                                     const int32_t* __restrict__ sel_idx,
                                                     const double* __restrict__ sel_w,
                                                     const double* __restrict__ rho,
                                                     const double* __restrict__ innov,
                                                     const double* __restrict__ HS,
                                                     const double* __restrict__ ymat,
                                                     const double* __restrict__ vec,
                                                     const double* __restrict__ corr, double infl,
                                                     double* __restrict__ Jz,
                                                     double* __restrict__ gL,
                                                     double* __restrict__ gA,
                                                     double* __restrict__ gth) {
  constexpr int KPL = NP > 64 ? NP / 64 : 1;  // vector entries per lane
  constexpr bool YLDS = NP <= 64;
  __shared__ double Ys[YLDS ? NP * NP : 1];
  __shared__ double sw[4][NP];
  __shared__ double red[4][3];
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  const int64_t p0 = sel_ptr[z];
  const int64_t p = sel_ptr[z + 1] - p0;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  if (p == 0) {
    if (tid == 0) {
      Jz[z] = 0.0;
      gL[z] = 0.0;
      gA[z] = 0.0;
    }
    return;
  }
  const double* V = vec + (size_t)zl * 4 * NP;
  const double* Yg = ymat + (size_t)zl * NP * NP;
  if (YLDS) {
    copy_batched<16>(tid, NP * NP, 256, [&](int e) { return Yg[e]; },
                     [&](int e, double x) { Ys[e] = x; });
    __syncthreads();
  }
  const double* Y = YLDS ? Ys : Yg;
  double ak[KPL], wk[KPL];
#pragma unroll
  for (int c = 0; c < KPL; ++c) {
    const int k = lane + 64 * c;
    ak[c] = k < NP ? V[NP + k] : 0.0;
    wk[c] = k < NP ? 1.0 / sqrt(V[3 * NP + k]) : 0.0;  // mu_k^-1/2
  }
  const double Lz = corr[z];
  double accL = 0.0, accR = 0.0, accLog = 0.0;
  for (int64_t q = wave; q < p; q += 4) {
    const int64_t l = sel_idx[p0 + q];
    const double w = sel_w[p0 + q];
    const double r = (w * w) * rho[l];
    const double* srow = HS + (size_t)l * NP;
    double t = 0.0;
#pragma unroll
    for (int c = 0; c < KPL; ++c) {
      const int j = lane + 64 * c;
      const double s = j < NP ? srow[hs_col<NP>(j)] : 0.0;
      if (j < NP) sw[wave][j] = s;
      t = fma(s, ak[c], t);
    }
    __builtin_amdgcn_wave_barrier();
    t = nll_wave_sum(t);
    double u = 0.0;
#pragma unroll
    for (int c = 0; c < KPL; ++c) {
      const int k = lane + 64 * c;
      if (k < NP) {
        double y = 0.0;
        for (int j = 0; j < NP; ++j) y = fma(Y[(size_t)j * NP + k], sw[wave][j], y);
        u = fma(wk[c] * y, y, u);
      }
    }
    u = nll_wave_sum(u);
    __builtin_amdgcn_wave_barrier();
    if (r > 0.0) {
      const double d = innov[l];
      const double e = d - t;
      const double g = 0.5 * (fma(e, e, u) - 1.0 / r);
      if (lane == 0) atomicAdd(gth + l, -2.0 * r * g);
      accL = fma(g * r, -4.0 * log(w) / Lz, accL);
      accR = fma(r * d, d, accR);
      accLog -= log(r);
    }
  }
  if (lane == 0) {
    red[wave][0] = accL;
    red[wave][1] = accR;
    red[wave][2] = accLog;
  }
  __syncthreads();
  if (wave == 0) {
    // sum_k log mu_k, sum_k (1 - 1/mu_k), a^T a, b^T a over the NP entries
    // (padding entries have mu = 1, a = b = 0)
    double slog = 0, sinv = 0, aa = 0, ba = 0;
    for (int k = lane; k < NP; k += 64) {
      const double mu = V[3 * NP + k], ak2 = V[NP + k];
      slog += log(mu);
      sinv += 1.0 - 1.0 / mu;
      aa = fma(ak2, ak2, aa);
      ba = fma(V[k], ak2, ba);
    }
    slog = nll_wave_sum(slog);
    sinv = nll_wave_sum(sinv);
    aa = nll_wave_sum(aa);
    ba = nll_wave_sum(ba);
    if (lane == 0) {
      const double sL = ((red[0][0] + red[1][0]) + red[2][0]) + red[3][0];
      const double sR = ((red[0][1] + red[1][1]) + red[2][1]) + red[3][1];
      const double sG = ((red[0][2] + red[1][2]) + red[2][2]) + red[3][2];
      Jz[z] = 0.5 * (sR - ba + sG + slog);
      gL[z] = sL;
      gA[z] = (sinv - aa) / infl;
    }
  }
}

void launch_nll_adjoint(int NP, const ChunkArgs& A, hipStream_t st) {
#define OAK_NLL(P)                                                                                \
  hipLaunchKernelGGL(k_nll_adjoint<P>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.N, A.sel_ptr,       \
                     A.sel_idx, A.sel_w, A.rho, A.innov, A.HS, A.ymat, A.vec, A.corr, A.infl,     \
                     A.nll_J, A.nll_gL, A.nll_gA, A.nll_gth)
  switch (NP) {
    case 16: OAK_NLL(16); break;
    case 32: OAK_NLL(32); break;
    case 64: OAK_NLL(64); break;
    case 128: OAK_NLL(128); break;
  }
#undef OAK_NLL
}

// Sums of the per-zone values (fixed order: one block, strided partial sums
// then a tree) -> out[0] = sum J_z, out[1] = sum dJ/dalpha.
__global__ __launch_bounds__(1024) void k_nll_totals(int64_t nz, const double* __restrict__ Jz,
                                                     const double* __restrict__ gA,
                                                     double* __restrict__ out) {
  __shared__ double sj[1024], sa[1024];
  double a = 0, b = 0;
  for (int64_t z = threadIdx.x; z < nz; z += 1024) {
    a += Jz[z];
    b += gA[z];
  }
  sj[threadIdx.x] = a;
  sa[threadIdx.x] = b;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sj[threadIdx.x] += sj[threadIdx.x + s];
      sa[threadIdx.x] += sa[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sj[0];
    out[1] = sa[0];
  }
}

void launch_nll_totals(int64_t nz, const double* Jz, const double* gA, double* out,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_nll_totals, dim3(1), dim3(1024), 0, st, nz, Jz, gA, out);
}

}  // namespace oakg
