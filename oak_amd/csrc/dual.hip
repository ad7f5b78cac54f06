// The analysis of a lightly observed zone (1 <= p < 64 local observations,
// N = 65..128 members) in observation space.  With S~ = diag(sqrt(r)) S the
// p x N weighted anomalies (r = w^2 / sigma^2, rrsqrt.F90:135) and
// d~ = sqrt(r) (yo - Hxf), the ensemble-space matrix is M = I + S~^T S~ and
// b = S~^T d~.  Its N - p eigenvalues beyond the rank of S~ are exactly 1,
// and every output of analysisIncrement is a matrix function of M, so it
// follows from the p x p matrix G = S~ S~^T = U diag(lambda) U^T
// (mu = max(1 + lambda, 1), as rrsqrt.F90:137 clamps):
//
//   a        = M^-1 b        = S~^T U diag(1/mu) U^T d~
//   M^-1/2                   = I - W W^T,  W = S~^T U diag(1/sqrt(mu^1/2 (1 + mu^1/2)))
//   M^1/2 1_N                = 1_N + S~^T U diag(1/(1 + mu^1/2)) U^T S~ 1_N
//
// (M^-1/2 - I and M^1/2 - I restricted to the range of S~^T; the weights are
// (mu^-1/2 - 1)/lambda and (mu^1/2 - 1)/lambda written without the
// cancellation at lambda -> 0).  The p x p eigenproblem runs through the
// NP = 64 kernels of tridiag.hip (Householder, QL, the eigenvector epilogue
// with its refinement of (I + G)^-1 d~); k_dual_lift then writes a, the
// normalised M^1/2 1_N and W into the zone's NP = 128 slots, with a flag
// (mu[0] = -1) that tells the transform kernels to form T = (I - W W^T) Omega
// instead of (Y Y^T) Omega.  The work per zone drops from that of a 128 x 128
// eigensolve to a 64 x 64 one plus two 128 x 64 x 64 products.
//
//   k_dual_gram   M_d = I + G (64 x 64), b_d = d~ (p rows of HS gathered and
//                 scaled in LDS, 4 x 4 register blocks of G per thread)
//   k_dual_lift   a, M^1/2 1_N, W from U diag(mu^-1/4), mu, a_d
#include "analysis.hpp"
#include "common.hpp"

namespace oakg {

namespace {

constexpr int PD = 64;        // observation-space size of the dual zones
constexpr int NPP = 128;      // ensemble-space size (NP of the primal slots)
constexpr int LDS_S = NPP + 1;  // padded row of S~ in LDS (rows 2 dwords apart in banks)

// Row q < p of S~ (natural column order) into s[q][*]; rows p..PD-1 zero.
// Returns the scaled innovation d~_q through dq (thread q < PD).
template <typename TS>
__device__ __forceinline__ void load_rows(double (*s)[LDS_S], double* dq, int64_t p0, int64_t p,
                                          const int32_t* __restrict__ sel_idx,
                                          const double* __restrict__ sel_w, int64_t m,
                                          const double* __restrict__ rho,
                                          const double* __restrict__ innov,
                                          const TS* __restrict__ HS, unsigned* __restrict__ err,
                                          double* sr) {
  const int tid = threadIdx.x;
  if (tid < PD) {
    double srq = 0.0, d = 0.0;
    if (tid < p) {
      const int64_t l = sel_idx[p0 + tid];
      if (l < 0 || l >= m) {
        atomicOr(err, OAKG_DERR_IDX);
      } else {
        const double w = sel_w[p0 + tid];
        srq = sqrt((w * w) * rho[l]);
        d = srq * innov[l];
      }
    }
    sr[tid] = srq;
    dq[tid] = d;
  }
  lds_barrier();
  // 128 consecutive stored values per row: position k holds column
  // 16 (k & 7) + (k >> 3) (hs_col<128>)
  for (int e = tid; e < PD * NPP; e += 256) {
    const int q = e >> 7, k = e & 127;
    double v = 0.0;
    if (q < p && sr[q] != 0.0) {
      const int64_t l = sel_idx[p0 + q];
      v = sr[q] * (double)HS[(size_t)l * NPP + k];
    }
    s[q][16 * (k & 7) + (k >> 3)] = v;
  }
  lds_barrier();
}

}  // namespace

template <typename TS>
__global__ __launch_bounds__(256) void k_dual_gram(int64_t z0, const int32_t* __restrict__ zorder,
                                                   const int64_t* __restrict__ sel_ptr,
                                                   const int32_t* __restrict__ sel_idx,
                                                   const double* __restrict__ sel_w, int64_t m,
                                                   const double* __restrict__ rho,
                                                   const double* __restrict__ innov,
                                                   const TS* __restrict__ HS,
                                                   double* __restrict__ mat_d,
                                                   double* __restrict__ vec_d,
                                                   unsigned* __restrict__ err) {
  __shared__ double s[PD][LDS_S];
  __shared__ double dq[PD], sr[PD];
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;  // not a dual zone
  const int64_t p0 = sel_ptr[z];
  const int64_t p = sel_ptr[z + 1] - p0;
  load_rows<TS>(s, dq, p0, p, sel_idx, sel_w, m, rho, innov, HS, err, sr);
  const int tid = threadIdx.x;
  // G[q][k] for q in 4 qb .. 4 qb + 3, k in 4 kb .. 4 kb + 3; the sums run in
  // column order, so G[q][k] and G[k][q] are the same sum (symmetric bits)
  const int qb = tid >> 4, kb = tid & 15;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int c = 0; c < NPP; ++c) {
    double a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = s[4 * qb + i][c];
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = s[4 * kb + j][c];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
  }
  double* Md = mat_d + (size_t)zl * PD * PD;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int q = 4 * qb + i, k = 4 * kb + j;
      Md[q * PD + k] = acc[i][j] + (q == k ? 1.0 : 0.0);
    }
  if (tid < PD) vec_d[(size_t)zl * 4 * PD + tid] = dq[tid];
}

template <typename TS>
__global__ __launch_bounds__(256) void k_dual_lift(int64_t z0, const int32_t* __restrict__ zorder,
                                                   int N, const int64_t* __restrict__ sel_ptr,
                                                   const int32_t* __restrict__ sel_idx,
                                                   const double* __restrict__ sel_w, int64_t m,
                                                   const double* __restrict__ rho,
                                                   const double* __restrict__ innov,
                                                   const TS* __restrict__ HS,
                                                   const double* __restrict__ ymat_d,
                                                   const double* __restrict__ vec_d,
                                                   double* __restrict__ ymat,
                                                   double* __restrict__ vec,
                                                   double* __restrict__ ampl,
                                                   unsigned* __restrict__ err) {
  __shared__ double s[PD][LDS_S];
  __shared__ double dq[PD], sr[PD], ad[PD], q2[PD], t1[PD], t4[PD], red[4];
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;
  const int64_t p0 = sel_ptr[z];
  const int64_t p = sel_ptr[z + 1] - p0;
  const int tid = threadIdx.x;
  const double* Yd = ymat_d + (size_t)zl * PD * PD;  // U diag(mu^-1/4)
  const double* Vd = vec_d + (size_t)zl * 4 * PD;     // [PD..2PD) a_d, [3PD..4PD) mu
  if (tid < PD) {
    q2[tid] = sqrt(Vd[3 * PD + tid]);  // mu^1/2
    ad[tid] = Vd[PD + tid];
  }
  load_rows<TS>(s, dq, p0, p, sel_idx, sel_w, m, rho, innov, HS, err, sr);
  // t1 = S~ 1_N
  if (tid < PD) {
    double a = 0.0;
    for (int c = 0; c < N; ++c) a += s[tid][c];
    t1[tid] = a;
  }
  lds_barrier();
  // t4 = U diag(1/(1 + mu^1/2)) U^T t1 = Yd diag(mu^1/2 / (1 + mu^1/2)) Yd^T t1
  if (tid < PD) {
    double a = 0.0;
    for (int q = 0; q < PD; ++q) a = fma(Yd[q * PD + tid], t1[q], a);
    dq[tid] = a * (q2[tid] / (1.0 + q2[tid]));  // dq is free after load_rows
  }
  lds_barrier();
  if (tid < PD) {
    double a = 0.0;
    for (int j = 0; j < PD; ++j) a = fma(Yd[tid * PD + j], dq[j], a);
    t4[tid] = a;
  }
  // W = S~^T Yd diag(c), c_j = 1/sqrt(1 + mu_j^1/2): thread (column j, block
  // of 32 rows); the rows of S~ are read as LDS broadcasts, Yd as coalesced
  // rows from global memory
  {
    const int j = tid & 63, ib = tid >> 6;
    const double cj = 1.0 / sqrt(1.0 + q2[j]);
    double acc[32];
#pragma unroll
    for (int ii = 0; ii < 32; ++ii) acc[ii] = 0.0;
    for (int q = 0; q < p; ++q) {
      const double wq = Yd[q * PD + j] * cj;
#pragma unroll
      for (int ii = 0; ii < 32; ++ii) acc[ii] = fma(s[q][32 * ib + ii], wq, acc[ii]);
    }
    double* Y = ymat + (size_t)zl * NPP * NPP;
#pragma unroll
    for (int ii = 0; ii < 32; ++ii) {
      Y[(size_t)(32 * ib + ii) * NPP + j] = acc[ii];
      Y[(size_t)(32 * ib + ii) * NPP + PD + j] = 0.0;
    }
  }
  lds_barrier();  // t4
  double* V = vec + (size_t)zl * 4 * NPP;
  // a, M^1/2 1_N (rows i < 128, threads 0..127; the other waves add zeros)
  const int i = tid;
  double vr = 0.0, ar = 0.0;
  if (tid < NPP) {
    for (int q = 0; q < p; ++q) {
      ar = fma(s[q][i], ad[q], ar);
      vr = fma(s[q][i], t4[q], vr);
    }
    vr += i < N ? 1.0 : 0.0;
  }
  const double ss = wave_sum_fast(vr * vr);
  if ((tid & 63) == 0) red[tid >> 6] = ss;
  lds_barrier();
  if (tid < NPP) {
    const double nv = (red[0] + red[1]) + (red[2] + red[3]);
    V[i] = 0.0;
    V[NPP + i] = ar;
    V[2 * NPP + i] = vr / sqrt(nv);
    V[3 * NPP + i] = i == 0 ? -1.0 : 0.0;  // dual flag for the transform kernels
    if (ampl && i < N) ampl[(size_t)z * N + i] = ar;
    if (ar != ar) atomicOr(err, OAKG_DERR_NAN);
  }
}

void launch_dual128(const ChunkArgs& A, int N, double* ymat128, double* vec128, double* ampl,
                    hipStream_t st, PhaseTimer* tm) {
  if (tm) tm->mark(PH_GRAM);
  if (A.prec == 1)
    hipLaunchKernelGGL(k_dual_gram<float>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                       A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS32, A.mat, A.vec, A.err);
  else
    hipLaunchKernelGGL(k_dual_gram<double>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                       A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS, A.mat, A.vec, A.err);
  if (tm) tm->mark(PH_EIG);
  launch_eig_small(PD, A, st);
  if (A.prec == 1)
    hipLaunchKernelGGL(k_dual_lift<float>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, N,
                       A.sel_ptr, A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS32, A.ymat, A.vec,
                       ymat128, vec128, ampl, A.err);
  else
    hipLaunchKernelGGL(k_dual_lift<double>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, N,
                       A.sel_ptr, A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS, A.ymat, A.vec,
                       ymat128, vec128, ampl, A.err);
}

}  // namespace oakg
