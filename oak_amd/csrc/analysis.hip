// FP64 per-zone local analysis on gfx950: rrsqrt.F90:100-190 for every zone.
//
//   k_gram    C = S^T diag(r) S and b = S^T (r o d) on v_mfma_f64_16x16x4f64,
//             stored as M = I + C (rrsqrt.F90:135, 142; r = w^2/rmse^2,
//             covariance.F90:612-619)
//   k_eig     one-sided (Hestenes) Jacobi on M, one matrix row per lane held in
//             registers; columns converge to x_j = mu_j u_j with mu = 1 + lambda
//             (the eigen-decomposition of rrsqrt.F90:136-140).  Epilogue builds
//             a = (I+C)^-1 b (rrsqrt.F90:142), v ~ (I+C)^1/2 1 (:176-177) and
//             Y = X diag(mu^-5/4) with Y Y^T = (I+C)^-1/2
//   k_tmat    T = (I+C)^-1/2 Omega, Omega = RotateVector(w, v) (rrsqrt.F90:185,
//             737-745; perpSpace matoper.F90:509-535) as identity + rank-5 terms
//   k_update  [xa - xf | Sa] = Sf_z [a | T] on MFMA, with the ensemble-mode
//             epilogue (inflation, maxCorrection, Ea = xa + sqrt(N-1) Sa;
//             assimilation.F90:3299-3320)
#include "analysis.hpp"
#include "common.hpp"

namespace oakg {

// ---------------------------------------------------------------------------
// Observation-space preparation
// ---------------------------------------------------------------------------
// locAnalysis mode: HSf (m x N, col-major) -> HS row-major m x NP (zero padded);
// innov = yo - Hxf; rho = 1/rmse^2 (0 for excluded observations).
template <int NP, int TR = (NP > 64 ? 32 : 64)>
__global__ __launch_bounds__(256) void k_prep_loc(int64_t m, int N, const double* __restrict__ HSf,
                                                  int64_t ld, const double* __restrict__ Hxf,
                                                  const double* __restrict__ yo,
                                                  const double* __restrict__ rmse2,
                                                  const uint8_t* __restrict__ excl,
                                                  double* __restrict__ HS,
                                                  double* __restrict__ innov,
                                                  double* __restrict__ rho,
                                                  unsigned* __restrict__ err, double scale) {
  __shared__ double tile[TR][NP + 1];
  const int64_t l0 = (int64_t)blockIdx.x * TR;
  for (int e = threadIdx.x; e < TR * NP; e += 256) {
    int k = e / TR, li = e % TR;
    int64_t l = l0 + li;
    tile[li][k] = (l < m && k < N) ? HSf[l + (int64_t)k * ld] : 0.0;
  }
  if (threadIdx.x < TR) {
    int64_t l = l0 + threadIdx.x;
    if (l < m) {
      double r2 = rmse2[l];
      if (!(r2 > 0) || !isfinite(r2)) atomicOr(err, OAKG_DERR_RMSE);
      innov[l] = yo[l] - Hxf[l];
      rho[l] = (excl && excl[l]) ? 0.0 : 1.0 / r2;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < TR * NP; e += 256) {
    int li = e / NP, k = e % NP;
    int64_t l = l0 + li;
    if (l < m) HS[l * NP + hs_col<NP>(k)] = tile[li][k] * scale;
  }
}

// Ensemble mode, state mean: xbar = sum(E,2)/N (assimilation.F90:3127).
__global__ void k_state_mean(int64_t n, int N, const double* __restrict__ E, int64_t ld,
                             double* __restrict__ xbar) {
#pragma clang fp contract(off)
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0;
  for (int k = 0; k < N; ++k) s = s + E[i + (int64_t)k * ld];
  xbar[i] = s / N;
}

// The same with two consecutive elements per thread (16-byte loads; E 16-byte
// aligned, ld and n even): half the load instructions, the same sums.
__global__ void k_state_mean_x2(int64_t n2, int N, const double2* __restrict__ E2, int64_t ld2,
                                double* __restrict__ xbar) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n2) return;
  double s0 = 0, s1 = 0;
  for (int k = 0; k < N; ++k) {
    const double2 v = E2[i + (int64_t)k * ld2];
    s0 = s0 + v.x;
    s1 = s1 + v.y;
  }
  reinterpret_cast<double2*>(xbar)[i] = make_double2(s0 / N, s1 / N);
}

// Ensemble mode, HE(:,k) = H.x.E(:,k) + Hshift (assimilation.F90:3112-3114;
// COO accumulation order of matoper_inc.F90:238-240 kept within each row).
__global__ void k_spmm_obs(int64_t m, int64_t n, int N, const int64_t* __restrict__ row_ptr,
                           const int32_t* __restrict__ col, const double* __restrict__ val,
                           const double* __restrict__ E, int64_t ld,
                           const double* __restrict__ Hshift, double* __restrict__ HEc,
                           unsigned* __restrict__ err) {
#pragma clang fp contract(off)
  int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (l >= m) return;
  const double* Ek = E + (int64_t)k * ld;
  double s = 0;
  for (int64_t e = row_ptr[l]; e < row_ptr[l + 1]; ++e) {
    int32_t j = col[e];
    if (j < 0 || j >= n) { atomicOr(err, OAKG_DERR_IDX); continue; }
    double t = val[e] * Ek[j];
    s = s + t;
  }
  HEc[l + (int64_t)k * m] = s + Hshift[l];
}

// Ensemble mode, observation anomalies: ybar = sum(HE,2)/N, HSf = (HE - ybar)/sqrt(N-1)
// (assimilation.F90:3128-3134), written row-major m x NP; innov = yo - ybar.
template <int NP, int TR = (NP > 64 ? 32 : 64)>
__global__ __launch_bounds__(256) void k_obs_anom(int64_t m, int N, const double* __restrict__ HEc,
                                                  const double* __restrict__ yo,
                                                  const double* __restrict__ rmse2,
                                                  const uint8_t* __restrict__ excl, double scaling,
                                                  double* __restrict__ HS,
                                                  double* __restrict__ innov,
                                                  double* __restrict__ rho,
                                                  unsigned* __restrict__ err) {
#pragma clang fp contract(off)
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
    for (int k = 0; k < N; ++k) s = s + tile[li][k];
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
  for (int e = threadIdx.x; e < TR * NP; e += 256) {
    int li = e / NP, k = e % NP;
    int64_t l = l0 + li;
    if (l < m) HS[l * NP + hs_col<NP>(k)] = (k < N) ? (tile[li][k] - ybar[li]) / scaling : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Gram matrix per zone (rrsqrt.F90:135 and the S^T R^-1 d factor of :142).
// One workgroup (4 waves) per zone; the waves split the local observations in
// steps of 4 and each wave accumulates every lower 16x16 tile; the partial
// tiles are summed across waves in a fixed order (deterministic).  TS is the
// stored type of S (double, or float for the mixed path: products of FP32
// values are exact in FP64, so the FP64 accumulation adds no FP32 error).
// ---------------------------------------------------------------------------
template <int NP, typename TS>
__global__ __launch_bounds__(256) void k_gram(int64_t z0, const int32_t* __restrict__ zorder, const int64_t* __restrict__ sel_ptr,
                                              const int32_t* __restrict__ sel_idx,
                                              const double* __restrict__ sel_w, int64_t m,
                                              const double* __restrict__ rho,
                                              const double* __restrict__ innov,
                                              const TS* __restrict__ HS,
                                              double* __restrict__ mat, double* __restrict__ vec,
                                              unsigned* __restrict__ err) {
  constexpr int NB = NP / 16;
  constexpr int NT = NB * (NB + 1) / 2;
  __shared__ double red[4][256];
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  const int64_t p0 = sel_ptr[z];
  const int64_t p = sel_ptr[z + 1] - p0;
  double* M = mat + (size_t)zl * NP * NP;
  double* bv_out = vec + (size_t)zl * 4 * NP;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int kq = lane >> 4, jl = lane & 15;
  if (p == 0) {
    for (int e = threadIdx.x; e < NP * NP; e += 256) M[e] = (e / NP == e % NP) ? 1.0 : 0.0;
    for (int e = threadIdx.x; e < NP; e += 256) bv_out[e] = 0.0;
    return;
  }
  d4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = d4{0, 0, 0, 0};
  double bacc[NB];
#pragma unroll
  for (int J = 0; J < NB; ++J) bacc[J] = 0;
  const int64_t nsteps = (p + 3) >> 2;
  // The steps of a wave (s = wave, wave + 4, ...) are software-pipelined as
  // in gram128_wave (analysis128.hip): the index and weight two steps ahead
  // and the gathers of the next step are in flight while a step's MFMAs run;
  // the step is branch-free, and a compiler barrier keeps the prefetched
  // loads from being folded into the next iteration.
  struct Idx {
    int64_t l;
    double w;
    bool ok;
  };
  bool bad = false;  // an index outside 0..m-1 (reported once, after the loop)
  auto fetch_idx = [&](int64_t s) {
    Idx x;
    const int64_t q = 4 * s + kq;
    x.ok = q < p;
    const int64_t qc = x.ok ? q : p - 1;  // p >= 1 here
    x.l = sel_idx[p0 + qc];
    x.w = sel_w[p0 + qc];
    return x;
  };
  auto fetch_row = [&](Idx& x, double& rh, double& inn, TS* rv) {
    const bool inr = x.l >= 0 && x.l < m;
    bad = bad | (x.ok & !inr);
    x.ok = x.ok & inr;
    x.l = x.ok ? x.l : 0;
    rh = rho[x.l];
    inn = innov[x.l];
    const TS* row = HS + (size_t)x.l * NP + jl;
#pragma unroll
    for (int J = 0; J < NB; ++J) rv[J] = row[16 * J];
  };
  Idx i0 = fetch_idx(wave), i1 = fetch_idx(wave + 4);
  double rh0, inn0;
  TS rv0[NB];
  fetch_row(i0, rh0, inn0, rv0);
  for (int64_t s = wave; s < nsteps; s += 4) {
    const Idx i2 = fetch_idx(s + 8);
    double rh1, inn1;
    TS rv1[NB];
    fetch_row(i1, rh1, inn1, rv1);
    asm volatile("" ::: "memory");  // compiler barrier: keeps the loads above here
    const bool ok = i0.ok;
    const double r = ok ? (i0.w * i0.w) * rh0 : 0.0;
    const double rd = ok ? r * inn0 : 0.0;
    double bv[NB];
#pragma unroll
    for (int J = 0; J < NB; ++J) bv[J] = ok ? (double)rv0[J] : 0.0;
    int t = 0;
#pragma unroll
    for (int I = 0; I < NB; ++I) {
      double a = r * bv[I];
#pragma unroll
      for (int J = 0; J <= I; ++J) {
        acc[t] = mfma_f64(a, bv[J], acc[t]);
        ++t;
      }
    }
#pragma unroll
    for (int J = 0; J < NB; ++J) bacc[J] = fma(rd, bv[J], bacc[J]);
    i0 = i1;
    i1 = i2;
    rh0 = rh1;
    inn0 = inn1;
#pragma unroll
    for (int J = 0; J < NB; ++J) rv0[J] = rv1[J];
  }
  if (bad) atomicOr(err, OAKG_DERR_IDX);
  // b: sum the four k-groups of the wave, then the waves
#pragma unroll
  for (int J = 0; J < NB; ++J) {
    bacc[J] += shfl_xor_d(bacc[J], 16);
    bacc[J] += shfl_xor_d(bacc[J], 32);
  }
  if (lane < 16) {
#pragma unroll
    for (int J = 0; J < NB; ++J) red[wave][16 * J + lane] = bacc[J];
  }
  __syncthreads();
  if (threadIdx.x < NP) {
    int e = threadIdx.x;
    bv_out[e] = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
  }
  __syncthreads();
  int t = 0;
#pragma unroll
  for (int I = 0; I < NB; ++I) {
#pragma unroll
    for (int J = 0; J <= I; ++J) {
#pragma unroll
      for (int r = 0; r < 4; ++r) red[wave][lane * 4 + r] = acc[t][r];
      __syncthreads();
      {
        const int tid = threadIdx.x;
        const double v = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
        const int ll = tid >> 2, rr = tid & 3;
        const int i = (ll >> 4) + 4 * rr, j = ll & 15;
        const int gi = 16 * I + i, gj = 16 * J + j;
        if (I != J || i >= j) {
          const double mv = v + (gi == gj ? 1.0 : 0.0);
          M[gi * NP + gj] = mv;
          M[gj * NP + gi] = mv;
        }
      }
      __syncthreads();
      ++t;
    }
  }
}

// ---------------------------------------------------------------------------
// One-sided Jacobi eigensolver, NP <= 64: a group of NP lanes per zone, lane r
// holds row r of X (initially M = I + C).  Each round rotates the H = NP/2
// column pairs (i, NP-1-i) at fixed register slots, then cycles slots
// 1..NP-1 by one (round-robin "circle" ordering), so NP-1 rounds pair every
// two columns once and restore the slot order.  Column sums over the group's
// rows go through LDS (each lane writes its products, each lane then adds one
// column over half of the rows; one xor-shuffle joins the halves).  Squared
// column norms are tracked per slot in LDS and recomputed every sweep.
// ---------------------------------------------------------------------------
template <int NP>
struct EigLds {
  static constexpr int H = NP / 2;
  static constexpr int GPW = 64 / NP;
  double red[64][H + 1];     // products, one row per lane
  double cs[GPW][H][2];      // rotation (c, s) of each pair
  double nrm[GPW][NP];       // squared column norm per slot
  double bc[GPW][2][NP];     // broadcasts of the epilogue
};

template <int NP>
__device__ __forceinline__ double group_sum(double s) {
#pragma unroll
  for (int o = 1; o < NP; o <<= 1) s += shfl_xor_d(s, o);
  return s;
}

// Sum over the group's rows of H values per lane, values given by f(k).
// Returns, in lane r, the total of column (r % H) (lanes r and r ^ H agree).
template <int NP, typename F>
__device__ __forceinline__ double colsum_half(EigLds<NP>& L, int lane, int gb, int r, F f) {
  constexpr int H = NP / 2;
#pragma unroll
  for (int k = 0; k < H; ++k) L.red[lane][k] = f(k);
  __builtin_amdgcn_wave_barrier();
  const int k = r & (H - 1), h = r / H;
  double s = 0;
#pragma unroll
  for (int i = 0; i < H; ++i) s += L.red[gb + h * H + i][k];
  __builtin_amdgcn_wave_barrier();
  return s + shfl_xor_d(s, H);
}

// Column sums of NP values per lane (two passes of H columns); lane r gets
// the total of column r.
template <int NP, typename F>
__device__ __forceinline__ double colsum_full(EigLds<NP>& L, int lane, int gb, int r, F f) {
  constexpr int H = NP / 2;
  const double s0 = colsum_half<NP>(L, lane, gb, r, [&](int k) { return f(k); });
  const double s1 = colsum_half<NP>(L, lane, gb, r, [&](int k) { return f(k + H); });
  return r < H ? s0 : s1;
}

template <int NP>
__global__ __launch_bounds__(256) void k_eig(int64_t z0, const int32_t* __restrict__ zorder, int nzc, int N,
                                             const int64_t* __restrict__ sel_ptr,
                                             const double* __restrict__ mat,
                                             double* __restrict__ ymat, double* __restrict__ vec,
                                             double* __restrict__ ampl,
                                             int32_t* __restrict__ sweeps_out,
                                             unsigned* __restrict__ err, double tol,
                                             int max_sweeps,
                                             const int32_t* __restrict__ only) {
  constexpr int GPW = 64 / NP;
  constexpr int H = NP / 2;
  __shared__ EigLds<NP> lds[4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  EigLds<NP>& L = lds[wave];
  const int grp = lane / NP, r = lane % NP, gb = grp * NP;
  const int zl = (blockIdx.x * 4 + wave) * GPW + grp;
  // with `only`, zones whose flag has no bit in 3 are left untouched
  bool valid = zl < nzc && (!only || (only[zl] & 3));
  int64_t z = valid ? zone_at(zorder, z0 + zl) : 0;
  if (z < 0) {  // solved by the other pipeline
    valid = false;
    z = 0;
  }
  // a wave with no zone to solve leaves at once (the fallback pass visits
  // every zone and redoes only the flagged ones); the waves share no barrier
  if (__ballot(valid) == 0) return;
  const bool active = valid && (sel_ptr[z + 1] > sel_ptr[z]);
  const double* M = mat + (size_t)(valid ? zl : 0) * NP * NP;
  double x[NP];
  if (active) {
#pragma unroll
    for (int j = 0; j < NP; ++j) x[j] = M[r * NP + j];
  } else {
#pragma unroll
    for (int j = 0; j < NP; ++j) x[j] = 0.0;
    x[0] = 1.0;  // any orthogonal start: the result of inactive zones is not used
  }

  const uint64_t gmask = (NP == 64) ? ~0ull : (((1ull << NP) - 1) << gb);
  bool done = !active;
  int sweeps = 0;
  for (int sw = 0; sw < max_sweeps; ++sw) {
    if (__ballot(!done) == 0) break;
    L.nrm[grp][r] = colsum_full<NP>(L, lane, gb, r, [&](int j) { return x[j] * x[j]; });
    bool rotated = false;
    for (int rd = 0; rd < NP - 1; ++rd) {
      const double gam = colsum_half<NP>(L, lane, gb, r, [&](int i) { return x[i] * x[NP - 1 - i]; });
      const int k = r & (H - 1);
      const double alpha = L.nrm[grp][k];
      const double beta = L.nrm[grp][NP - 1 - k];
      double c = 1.0, s = 0.0, t = 0.0;
      if (!done && fabs(gam) > tol * sqrt(alpha * beta)) {
        const double zeta = (beta - alpha) / (2.0 * gam);
        t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        c = 1.0 / sqrt(1.0 + t * t);
        s = c * t;
        rotated = true;
      }
      __builtin_amdgcn_wave_barrier();
      if (r < H) {
        L.cs[grp][k][0] = c;
        L.cs[grp][k][1] = s;
        // norms after the rotation, stored at the slots the columns move to
        const int sp = (k == 0) ? 0 : k + 1;
        const int sq = (k == 0) ? 1 : NP - k;
        L.nrm[grp][sp] = alpha - t * gam;
        L.nrm[grp][sq] = beta + t * gam;
      }
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int i = 0; i < H; ++i) {
        const double ci = L.cs[grp][i][0], si = L.cs[grp][i][1];
        const double xp = x[i], xq = x[NP - 1 - i];
        x[i] = ci * xp - si * xq;
        x[NP - 1 - i] = si * xp + ci * xq;
      }
      const double tl = x[NP - 1];
#pragma unroll
      for (int j = NP - 1; j >= 2; --j) x[j] = x[j - 1];
      x[1] = tl;
    }
    const uint64_t rot = __ballot(rotated) & gmask;
    if (!done) ++sweeps;
    if (rot == 0) done = true;
  }

  // Epilogue: mu_j = ||x_j|| (eigenvalues of I + C), clamped at 1 as
  // "where (lambda < 0) lambda = 0" (rrsqrt.F90:137).
  const double nrm2 = colsum_full<NP>(L, lane, gb, r, [&](int j) { return x[j] * x[j]; });
  const double mun = sqrt(nrm2);
  const double muc = fmax(mun, 1.0);
  const double bval = active ? vec[(size_t)zl * 4 * NP + r] : 0.0;
  const double one = (r < N) ? 1.0 : 0.0;
  const double c1 = colsum_full<NP>(L, lane, gb, r, [&](int j) { return x[j] * bval; });
  const double c2 = colsum_full<NP>(L, lane, gb, r, [&](int j) { return x[j] * one; });
  // a = U diag(1/mu) U^T b ; v = U diag(mu^1/2) U^T 1 ; with U = X diag(1/mun)
  const double sa = 1.0 / (nrm2 * muc);
  L.bc[grp][0][r] = c1 * sa;
  L.bc[grp][1][r] = c2 * sqrt(muc) / nrm2;
  __builtin_amdgcn_wave_barrier();
  double ar = 0, vr = 0;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    ar = fma(x[j], L.bc[grp][0][j], ar);
    vr = fma(x[j], L.bc[grp][1][j], vr);
  }
  __builtin_amdgcn_wave_barrier();
  // One step of iterative refinement against the unmodified M: a += M^-1 (b - M a).
  // The eigenvector cross-talk left by the stopping tolerance would otherwise
  // reach a with a factor mu_max / mu_min.
  L.bc[grp][0][r] = ar;
  __builtin_amdgcn_wave_barrier();
  double ma = 0;
  if (active) {
#pragma unroll 8
    for (int j = 0; j < NP; ++j) ma = fma(M[r * NP + j], L.bc[grp][0][j], ma);
  }
  const double res = bval - ma;
  const double c3 = colsum_full<NP>(L, lane, gb, r, [&](int j) { return x[j] * res; });
  __builtin_amdgcn_wave_barrier();
  L.bc[grp][0][r] = c3 * sa;
  __builtin_amdgcn_wave_barrier();
  double da = 0;
#pragma unroll
  for (int j = 0; j < NP; ++j) da = fma(x[j], L.bc[grp][0][j], da);
  ar += da;
  const double nv = group_sum<NP>(vr * vr);
  vr = vr / sqrt(nv);
  __builtin_amdgcn_wave_barrier();
  // Y = X diag(muc^-1/4 / mun): Y Y^T = U diag(mu^-1/2) U^T
  L.bc[grp][0][r] = 1.0 / (sqrt(sqrt(muc)) * mun);
  __builtin_amdgcn_wave_barrier();
  if (valid) {
    double* Yz = ymat + (size_t)zl * NP * NP;
#pragma unroll
    for (int j = 0; j < NP; ++j) Yz[r * NP + j] = x[j] * L.bc[grp][0][j];
    double* V = vec + (size_t)zl * 4 * NP;
    V[NP + r] = active ? ar : 0.0;
    V[2 * NP + r] = vr;
    V[3 * NP + r] = muc;
    if (ampl && active && r < N) ampl[(size_t)z * N + r] = ar;
    if (active && ar != ar) atomicOr(err, OAKG_DERR_NAN);
    if (r == 0) sweeps_out[zl] = sweeps;
  }
}

// ---------------------------------------------------------------------------
// T = (Y Y^T) Omega, Omega = RotateVector(w, v) = v w^T + P(v) P(w)^T with
// P(u) = [I; 0] + alpha_u (u + sign(u_N) e_N) u(1:N-1)^T, alpha_u = -1/(|u_N|+1):
//   Omega = I - e e^T + v w^T + alpha_w w~ z_w^T + alpha_v z_v v~^T
//           + alpha_v alpha_w (v~.w~) z_v z_w^T
// (u~ = u with u_N zeroed, z_u = u + sign(u_N) e, e = e_N, N the true size).
// Zones without observations get T = I exactly (Sa = Sf, rrsqrt.F90:326,371).
// ---------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(256) void k_tmat(int64_t z0, const int32_t* __restrict__ zorder, int N, const int64_t* __restrict__ sel_ptr,
                                              const double* __restrict__ ymat,
                                              double* __restrict__ mat,
                                              const double* __restrict__ vec) {
  constexpr int NB = NP / 16;
  constexpr int LD = NP + 1;
  __shared__ double Ys[NP * LD];
  __shared__ double T0[NP * LD];
  __shared__ double U[4][NP];
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  double* M = mat + (size_t)zl * NP * NP;
  const bool active = sel_ptr[z + 1] > sel_ptr[z];
  if (!active) {
    for (int e = threadIdx.x; e < NP * NP; e += 256) M[e] = (e / NP == e % NP) ? 1.0 : 0.0;
    return;
  }
  const double* v = vec + (size_t)zl * 4 * NP + 2 * NP;
  const double* Yz = ymat + (size_t)zl * NP * NP;
  for (int e = threadIdx.x; e < NP * NP; e += 256) Ys[(e / NP) * LD + e % NP] = Yz[e];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, kk = lane >> 4;
  for (int tile = wave; tile < NB * NB; tile += 4) {
    const int I = tile / NB, J = tile % NB;
    d4 acc = {0, 0, 0, 0};
#pragma unroll 4
    for (int j0 = 0; j0 < NP; j0 += 4) {
      const double a = Ys[(16 * I + li) * LD + j0 + kk];
      const double b = Ys[(16 * J + li) * LD + j0 + kk];
      acc = mfma_f64(a, b, acc);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) T0[(16 * I + (lane >> 4) + 4 * r) * LD + 16 * J + li] = acc[r];
  }
  __syncthreads();
  const double wN = 1.0 / sqrt((double)N);
  const double vN = v[N - 1];
  const double sgv = vN >= 0 ? 1.0 : -1.0;
  const double alv = -1.0 / (fabs(vN) + 1.0);
  const double alw = -1.0 / (wN + 1.0);
  // v~ . w~
  double vw = 0;
  for (int s = 0; s < N - 1; ++s) vw = fma(v[s], wN, vw);
  if (threadIdx.x < NP) {
    const int r = threadIdx.x;
    double u1 = 0, u2 = 0, u3 = 0;
    for (int q = 0; q < NP; ++q) {
      const double t = T0[r * LD + q];
      const double vq = q < N ? v[q] : 0.0;
      u1 = fma(t, vq, u1);
      u2 = fma(t, q < N - 1 ? wN : 0.0, u2);
      u3 = fma(t, vq + (q == N - 1 ? sgv : 0.0), u3);
    }
    U[0][r] = T0[r * LD + N - 1];
    U[1][r] = u1;
    U[2][r] = u2;
    U[3][r] = u3;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NP * NP; e += 256) {
    const int r = e / NP, s = e % NP;
    double t = T0[r * LD + s];
    if (s < N) {
      const double ws = wN;
      const double zw = wN + (s == N - 1 ? 1.0 : 0.0);
      const double vt = (s < N - 1) ? v[s] : 0.0;
      if (s == N - 1) t -= U[0][r];
      t += U[1][r] * ws + alw * U[2][r] * zw + alv * U[3][r] * vt + alv * alw * vw * U[3][r] * zw;
    }
    M[e] = t;
  }
}

// ---------------------------------------------------------------------------
// Update: out = Sf_z [T | a] for the zone's rows.  Computed transposed on the
// MFMA (A = [T|a]^T from LDS, B = Sf_z^T from global) so that the stores of
// each 16-row strip are contiguous in the column-major outputs.
//   MODE_LOC: xa = xf + Sf a, Sa = Sf T            (rrsqrt.F90:151,185,463)
//   MODE_ENS: Sf = (E - xbar)/sqrt(N-1) on the fly; xa = xbar + clamp(Sf a);
//             Ea = xa + sqrt(N-1) * (inflation * Sf T)  (assimilation.F90:3301-3320)
// ---------------------------------------------------------------------------
template <int NP, int MODE>
__global__ __launch_bounds__(256) void k_update(int64_t z0, const int32_t* __restrict__ zorder, int N, const int64_t* __restrict__ zone_ptr,
                                                const double* __restrict__ mat,
                                                const double* __restrict__ vec,
                                                const double* __restrict__ X0,  // xf or xbar
                                                const double* __restrict__ S, int64_t ldS,  // Sf or E
                                                double* __restrict__ xa, double* __restrict__ Out,
                                                int64_t ldO, double scaling, double infl,
                                                double maxcorr,
                                                const double* __restrict__ maxcorr_v) {
#pragma clang fp contract(off)
  constexpr int NB = NP / 16;
  constexpr int LDW = NP + 17;
  __shared__ double Ws[NP * LDW];
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  const int64_t i0 = zone_ptr[z];
  const int64_t sz = zone_ptr[z + 1] - i0;
  const double* T = mat + (size_t)zl * NP * NP;
  const double* a = vec + (size_t)zl * 4 * NP + NP;
  for (int e = threadIdx.x; e < NP * (NP + 16); e += 256) {
    const int k = e / (NP + 16), c = e % (NP + 16);
    double v = 0.0;
    if (c < NP) v = T[k * NP + c];
    else if (c == NP) v = a[k];
    Ws[k * LDW + c] = v;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, kk = lane >> 4;
  for (int64_t rb = 16 * wave; rb < sz; rb += 64) {
    d4 acc[NB + 1];
#pragma unroll
    for (int J = 0; J <= NB; ++J) acc[J] = d4{0, 0, 0, 0};
    const int64_t row = rb + li;
    const bool rok = row < sz;
    const double xb = (MODE == 1 && rok) ? X0[i0 + row] : 0.0;
    for (int k0 = 0; k0 < NP; k0 += 4) {
      const int kc = k0 + kk;
      double bval = 0.0;
      if (rok && kc < N) {
        const double sv = S[(i0 + row) + (int64_t)kc * ldS];
        bval = (MODE == 1) ? (sv - xb) / scaling : sv;
      }
#pragma unroll
      for (int J = 0; J <= NB; ++J) acc[J] = mfma_f64(Ws[kc * LDW + 16 * J + li], bval, acc[J]);
    }
    // xa - xf of row (rb + li) is column 0 of the last tile: lanes 0..15, register 0
    const double dx = shfl_d(acc[NB][0], li);
    double xar = 0.0;
    if (rok) {
      const double x0 = X0[i0 + row];
      xar = x0 + dx;
      if (MODE == 1) {
        const double mc = maxcorr_v ? maxcorr_v[i0 + row] : maxcorr;
        if (xar - mc > x0) xar = x0 + mc;
        if (xar < x0 - mc) xar = x0 - mc;
      }
      if (MODE == 0 && kk == 0) xa[i0 + row] = xar;
    }
#pragma unroll
    for (int J = 0; J < NB; ++J) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = 16 * J + kk + 4 * r;
        if (rok && col < N) {
          double o = acc[J][r];
          if (MODE == 1) {
            double sa = (infl != 1.0) ? infl * o : o;
            o = xar + sa * scaling;
          }
          Out[(i0 + row) + (int64_t)col * ldO] = o;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <int NP>
static void prep_loc_t(const PrepLocArgs& A, hipStream_t st) {
  constexpr int TR = NP > 128 ? 16 : (NP > 64 ? 32 : 64);
  hipLaunchKernelGGL((k_prep_loc<NP, TR>), dim3(nblk(A.m, TR)), dim3(256), 0, st, A.m, A.N, A.HSf,
                     A.ldHSf, A.Hxf, A.yo, A.rmse2, A.excl, A.HS, A.innov, A.rho, A.err,
                     A.hs_scale);
}

void launch_prep_loc(int NP, const PrepLocArgs& A, hipStream_t st) {
  if (A.m == 0) return;
  switch (NP) {
    case 16: prep_loc_t<16>(A, st); break;
    case 32: prep_loc_t<32>(A, st); break;
    case 64: prep_loc_t<64>(A, st); break;
    case 128: prep_loc_t<128>(A, st); break;
    case 256: prep_loc_t<256>(A, st); break;
  }
}

template <int NP>
static void prep_ens_t(const PrepEnsArgs& A, hipStream_t st) {
  if (A.xbar) {  // NULL: observation side only (the report)
    if ((uintptr_t)A.E % 16 == 0 && A.ldE % 2 == 0 && A.n % 2 == 0)
      hipLaunchKernelGGL(k_state_mean_x2, dim3(nblk(A.n / 2, 256)), dim3(256), 0, st, A.n / 2, A.N,
                         reinterpret_cast<const double2*>(A.E), A.ldE / 2, A.xbar);
    else
      hipLaunchKernelGGL(k_state_mean, dim3(nblk(A.n, 256)), dim3(256), 0, st, A.n, A.N, A.E,
                         A.ldE, A.xbar);
  }
  if (A.m == 0) return;
  hipLaunchKernelGGL(k_spmm_obs, dim3(nblk(A.m, 256), A.N), dim3(256), 0, st, A.m, A.n, A.N,
                     A.row_ptr, A.col, A.val, A.E, A.ldE, A.Hshift, A.HEc, A.err);
  constexpr int TR = NP > 128 ? 16 : (NP > 64 ? 32 : 64);
  hipLaunchKernelGGL((k_obs_anom<NP, TR>), dim3(nblk(A.m, TR)), dim3(256), 0, st, A.m, A.N, A.HEc, A.yo,
                     A.rmse2, A.excl, A.scaling, A.HS, A.innov, A.rho, A.err);
}

void launch_prep_ens(int NP, const PrepEnsArgs& A, hipStream_t st) {
  if (A.E32) {
    launch_prep_ens_f32(NP, A, st);
    return;
  }
  switch (NP) {
    case 16: prep_ens_t<16>(A, st); break;
    case 32: prep_ens_t<32>(A, st); break;
    case 64: prep_ens_t<64>(A, st); break;
    case 128: prep_ens_t<128>(A, st); break;
    case 256: prep_ens_t<256>(A, st); break;
  }
}

// The eigensolver of a chunk (M in A.mat, b in A.vec) for NP <= 64: a, v,
// mu into A.vec and Y = U diag(mu^-1/4) into A.ymat.
template <int NP>
static void eig_t(const ChunkArgs& A, hipStream_t st) {
  constexpr int GPW = 64 / NP;
  if (A.eig_method == 1) {
    launch_tql_pipeline(NP, A, st);
    // zones the QL could not finish fall back to Jacobi
    hipLaunchKernelGGL(k_eig<NP>, dim3(nblk(A.nzc, 4 * GPW)), dim3(256), 0, st, A.z0, A.zorder, A.nzc,
                       A.N, A.sel_ptr, A.mat, A.ymat, A.vec, A.ampl, A.sweeps, A.err, A.tol,
                       A.max_sweeps, (const int32_t*)A.tql_status);
  } else {
    hipLaunchKernelGGL(k_eig<NP>, dim3(nblk(A.nzc, 4 * GPW)), dim3(256), 0, st, A.z0, A.zorder, A.nzc,
                       A.N, A.sel_ptr, A.mat, A.ymat, A.vec, A.ampl, A.sweeps, A.err, A.tol,
                       A.max_sweeps, (const int32_t*)nullptr);
  }
}

void launch_eig_small(int NP, const ChunkArgs& A, hipStream_t st) {
  switch (NP) {
    case 16: eig_t<16>(A, st); break;
    case 32: eig_t<32>(A, st); break;
    case 64: eig_t<64>(A, st); break;
  }
}

template <int NP>
static void chunk_t(const ChunkArgs& A, hipStream_t st, PhaseTimer* tm) {
  if (tm) tm->mark(PH_GRAM);
  if (A.prec == 1)
    hipLaunchKernelGGL((k_gram<NP, float>), dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                       A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS32, A.mat, A.vec, A.err);
  else
    hipLaunchKernelGGL((k_gram<NP, double>), dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                       A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS, A.mat, A.vec, A.err);
  if (tm) tm->mark(PH_EIG);
  eig_t<NP>(A, st);
  if (A.mode == 2) {  // innovation likelihood: adjoint instead of transform + update
    if (tm) tm->mark(PH_UPDATE);
    launch_nll_adjoint(NP, A, st);
    if (tm) tm->mark(PH_END);
    return;
  }
  if (tm) tm->mark(PH_TMAT);
  hipLaunchKernelGGL(k_tmat<NP>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.N, A.sel_ptr, A.ymat,
                     A.mat, A.vec);
  if (tm) tm->mark(PH_UPDATE);
  if (A.prec == 1)
    launch_update_f32(NP, A, st);
  else if (A.mode == 0)
    hipLaunchKernelGGL((k_update<NP, 0>), dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.N, A.zone_ptr,
                       A.mat, A.vec, A.X0, A.S, A.ldS, A.xa, A.Out, A.ldO, A.scaling, A.infl,
                       A.maxcorr, A.maxcorr_v);
  else
    hipLaunchKernelGGL((k_update<NP, 1>), dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.N, A.zone_ptr,
                       A.mat, A.vec, A.X0, A.S, A.ldS, A.xa, A.Out, A.ldO, A.scaling, A.infl,
                       A.maxcorr, A.maxcorr_v);
  if (tm) tm->mark(PH_END);
}

void launch_chunk(int NP, const ChunkArgs& A, hipStream_t st, PhaseTimer* tm) {
  switch (NP) {
    case 16: chunk_t<16>(A, st, tm); break;
    case 32: chunk_t<32>(A, st, tm); break;
    case 64: chunk_t<64>(A, st, tm); break;
    case 128: launch_chunk128(A, st, tm); break;
    case 256: launch_chunk256(A, st, tm); break;
  }
}

}  // namespace oakg
