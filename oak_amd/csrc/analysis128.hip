// FP64 per-zone analysis for ensembles of 65..128 members (NP = 128):
// the same steps as analysis.hip (rrsqrt.F90:100-190), with kernels shaped
// for a 128 x 128 ensemble-space matrix.
//
//   k_gram128  M = I + S^T diag(r) S on v_mfma_f64_16x16x4f64 (S stored FP64, or
//              FP32 for the mixed path): 36 lower tiles,
//              wave w owns tile rows {w, 7-w} (9 tiles each, no cross-wave sum)
//   k_eig128   block one-sided Jacobi: each wave holds two 16-column blocks of
//              all 128 rows (2 rows per lane, 64 doubles in registers); a sweep
//              runs the 120 intra-block pairs of its blocks, then 7 block
//              rounds (round-robin over the 8 blocks) of 16 cross-block rounds;
//              blocks move between waves through the zone's scratch matrix
//   k_xform128 T = (Y Y^T) Omega formed in LDS and the update Ea = xa + Sf T
//              run from it (FP64 MFMA; FP32 MFMA for FP32 members)
#include <type_traits>

#include <algorithm>

#include "analysis.hpp"
#include "common.hpp"

namespace oakg {

constexpr int P128 = 128;

// ---------------------------------------------------------------------------
// Gram matrix, NP = 128
// ---------------------------------------------------------------------------
template <int W, typename TS>
__device__ __forceinline__ void gram128_wave(int64_t p0, int64_t p, const int32_t* __restrict__ sel_idx,
                                             const double* __restrict__ sel_w, int64_t m,
                                             const double* __restrict__ rho,
                                             const double* __restrict__ innov,
                                             const TS* __restrict__ HS, double* __restrict__ M,
                                             double* __restrict__ bout, unsigned* __restrict__ err,
                                             int lane) {
  constexpr int IA = W, IB = 7 - W;
  const int kq = lane >> 4, jl = lane & 15;
  d4 accA[IA + 1], accB[IB + 1];
#pragma unroll
  for (int J = 0; J <= IA; ++J) accA[J] = d4{0, 0, 0, 0};
#pragma unroll
  for (int J = 0; J <= IB; ++J) accB[J] = d4{0, 0, 0, 0};
  double bacc[8];
#pragma unroll
  for (int J = 0; J < 8; ++J) bacc[J] = 0;
  const int64_t nsteps = (p + 3) >> 2;
  // Each step's gathers hang on a load (index -> rho, innov and the row of
  // S), so the steps are software-pipelined: the index and weight of step
  // s + 2 and the gathers of step s + 1 are in flight while the MFMAs of step
  // s run (same arithmetic as one step at a time).  The step is branch-free
  // (indices past the zone's rows and invalid ones are clamped to valid
  // entries and their values dropped), so that the compiler keeps the loads
  // where they are issued instead of sinking them to their use.
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
  // columns 16 J + jl, J = 0..7, stored contiguously (hs_col<128>); row 0
  // stands in for a step past the zone's rows (its values are dropped)
  auto fetch_row = [&](Idx& x, double& rh, double& inn, TS* rv) {
    const bool inr = x.l >= 0 && x.l < m;
    bad = bad | (x.ok & !inr);
    x.ok = x.ok & inr;
    x.l = x.ok ? x.l : 0;
    rh = rho[x.l];
    inn = innov[x.l];
    const TS* row = HS + (size_t)x.l * P128 + 8 * jl;
    if constexpr (sizeof(TS) == 4) {
      const float4 u0 = reinterpret_cast<const float4*>(row)[0];
      const float4 u1 = reinterpret_cast<const float4*>(row)[1];
      rv[0] = u0.x; rv[1] = u0.y; rv[2] = u0.z; rv[3] = u0.w;
      rv[4] = u1.x; rv[5] = u1.y; rv[6] = u1.z; rv[7] = u1.w;
    } else {
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const double2 u = reinterpret_cast<const double2*>(row)[h];
        rv[2 * h] = u.x;
        rv[2 * h + 1] = u.y;
      }
    }
  };
  Idx i0 = fetch_idx(0), i1 = fetch_idx(1);
  double rh0, inn0;
  TS rv0[8];
  fetch_row(i0, rh0, inn0, rv0);
  for (int64_t s = 0; s < nsteps; ++s) {
    const Idx i2 = fetch_idx(s + 2);
    double rh1, inn1;
    TS rv1[8];
    fetch_row(i1, rh1, inn1, rv1);
    // a compiler barrier (no instruction): without it the loads above, whose
    // only use is the next iteration, are folded into one load at the top
    // of the loop and the pipeline is gone
    asm volatile("" ::: "memory");
    // step s
    const bool ok = i0.ok;
    const double r = ok ? (i0.w * i0.w) * rh0 : 0.0;
    const double rd = ok ? r * inn0 : 0.0;
    double bv[8];
#pragma unroll
    for (int J = 0; J < 8; ++J) bv[J] = ok ? (double)rv0[J] : 0.0;
    const double aA = r * bv[IA], aB = r * bv[IB];
#pragma unroll
    for (int J = 0; J <= IA; ++J) accA[J] = mfma_f64(aA, bv[J], accA[J]);
#pragma unroll
    for (int J = 0; J <= IB; ++J) accB[J] = mfma_f64(aB, bv[J], accB[J]);
    if (W == 0) {
#pragma unroll
      for (int J = 0; J < 8; ++J) bacc[J] = fma(rd, bv[J], bacc[J]);
    }
    i0 = i1;
    i1 = i2;
    rh0 = rh1;
    inn0 = inn1;
#pragma unroll
    for (int J = 0; J < 8; ++J) rv0[J] = rv1[J];
  }
  if (bad) atomicOr(err, OAKG_DERR_IDX);
  auto store = [&](int I, int J, const d4& acc) {
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int i = (lane >> 4) + 4 * rr, j = lane & 15;
      const int gi = 16 * I + i, gj = 16 * J + j;
      if (I != J || i >= j) {
        const double mv = acc[rr] + (gi == gj ? 1.0 : 0.0);
        M[gi * P128 + gj] = mv;
        M[gj * P128 + gi] = mv;
      }
    }
  };
#pragma unroll
  for (int J = 0; J <= IA; ++J) store(IA, J, accA[J]);
#pragma unroll
  for (int J = 0; J <= IB; ++J) store(IB, J, accB[J]);
  if (W == 0) {
#pragma unroll
    for (int J = 0; J < 8; ++J) {
      bacc[J] += shfl_xor_d(bacc[J], 16);
      bacc[J] += shfl_xor_d(bacc[J], 32);
    }
    if (lane < 16) {
#pragma unroll
      for (int J = 0; J < 8; ++J) bout[16 * J + lane] = bacc[J];
    }
  }
}

template <typename TS>
__global__ __launch_bounds__(256) void k_gram128(int64_t z0, const int32_t* __restrict__ zorder, const int64_t* __restrict__ sel_ptr,
                                                 const int32_t* __restrict__ sel_idx,
                                                 const double* __restrict__ sel_w, int64_t m,
                                                 const double* __restrict__ rho,
                                                 const double* __restrict__ innov,
                                                 const TS* __restrict__ HS,
                                                 double* __restrict__ mat, double* __restrict__ vec,
                                                 unsigned* __restrict__ err) {
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;  // the zone is solved by the other pipeline
  const int64_t p0 = sel_ptr[z];
  const int64_t p = sel_ptr[z + 1] - p0;
  double* M = mat + (size_t)zl * P128 * P128;
  double* bout = vec + (size_t)zl * 4 * P128;
  if (p == 0) {
    for (int e = threadIdx.x; e < P128 * P128; e += 256) M[e] = (e / P128 == e % P128) ? 1.0 : 0.0;
    for (int e = threadIdx.x; e < P128; e += 256) bout[e] = 0.0;
    return;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  switch (wave) {
    case 0: gram128_wave<0, TS>(p0, p, sel_idx, sel_w, m, rho, innov, HS, M, bout, err, lane); break;
    case 1: gram128_wave<1, TS>(p0, p, sel_idx, sel_w, m, rho, innov, HS, M, bout, err, lane); break;
    case 2: gram128_wave<2, TS>(p0, p, sel_idx, sel_w, m, rho, innov, HS, M, bout, err, lane); break;
    default: gram128_wave<3, TS>(p0, p, sel_idx, sel_w, m, rho, innov, HS, M, bout, err, lane); break;
  }
}

// The same Gram as a persistent kernel: a fixed grid takes zones from work
// counters in the order of the zone loop until the chunk is done.  Wave w of
// a workgroup always computes quarter w of a zone's tiles (the quarters need
// no common barrier), so the quarter is fixed per wave and its tile loop is
// compiled alone.  WG: the workgroup takes one zone for its four waves from
// counter 0 (the zone broadcast through LDS); else each wave takes the next
// zone of its own quarter from counter w.  The counters are cleared by the
// caller (a fill kernel) before each launch; every wave leaves its loop once
// the zone it reads passes the chunk (with WG all four waves read the same
// value, so they pass the same number of barriers).
constexpr int kPersistBlocksPerCU = 2;  // the register bound the kernel is compiled for

__global__ void k_clear_wq(unsigned* wq) {
  if (threadIdx.x < 4) wq[threadIdx.x] = 0u;
}

template <int W, typename TS, bool WG>
__device__ __forceinline__ void gram128_persist_loop(int64_t z0, const int32_t* __restrict__ zorder,
                                                     int nzc, unsigned* __restrict__ counter,
                                                     unsigned* sh_item,
                                                     const int64_t* __restrict__ sel_ptr,
                                                     const int32_t* __restrict__ sel_idx,
                                                     const double* __restrict__ sel_w, int64_t m,
                                                     const double* __restrict__ rho,
                                                     const double* __restrict__ innov,
                                                     const TS* __restrict__ HS,
                                                     double* __restrict__ mat,
                                                     double* __restrict__ vec,
                                                     unsigned* __restrict__ err, int lane) {
  auto zone = [&](unsigned item) {
    const int64_t z = zone_at(zorder, z0 + (int64_t)item);
    if (z < 0) return;  // the zone is solved by the other pipeline
    const int64_t p0 = sel_ptr[z];
    const int64_t p = sel_ptr[z + 1] - p0;
    double* M = mat + (size_t)item * P128 * P128;
    double* bout = vec + (size_t)item * 4 * P128;
    if (p == 0) {  // identity: rows 32 W .. 32 W + 31; b from quarter 0
      for (int e = lane; e < 32 * P128; e += 64) {
        const int r = 32 * W + e / P128, c = e % P128;
        M[r * P128 + c] = r == c ? 1.0 : 0.0;
      }
      if (W == 0)
        for (int e = lane; e < P128; e += 64) bout[e] = 0.0;
      return;
    }
    gram128_wave<W, TS>(p0, p, sel_idx, sel_w, m, rho, innov, HS, M, bout, err, lane);
  };
  if constexpr (WG) {
    // wave 0 takes the zone; the first barrier publishes it (the store is
    // waited for before the barrier), the second keeps the slot until every
    // wave has read it
    for (;;) {
      if (W == 0 && lane == 0) *sh_item = atomicAdd(counter, 1u);
      __syncthreads();
      const unsigned item = *sh_item;
      __syncthreads();
      if (item >= (unsigned)nzc) break;  // the same value in every wave
      zone(item);
    }
  } else {
    for (;;) {
      unsigned item = 0;
      if (lane == 0) item = atomicAdd(counter + W, 1u);
      item = __builtin_amdgcn_readfirstlane(item);
      if (item >= (unsigned)nzc) break;
      zone(item);
    }
  }
}

template <typename TS, bool WG>
__global__ __launch_bounds__(256, kPersistBlocksPerCU) void k_gram128_persist(
    int64_t z0, const int32_t* __restrict__ zorder, int nzc, unsigned* __restrict__ counter,
    const int64_t* __restrict__ sel_ptr, const int32_t* __restrict__ sel_idx,
    const double* __restrict__ sel_w, int64_t m, const double* __restrict__ rho,
    const double* __restrict__ innov, const TS* __restrict__ HS, double* __restrict__ mat,
    double* __restrict__ vec, unsigned* __restrict__ err) {
  __shared__ unsigned sh_item;
  // the wave index as a scalar: each wave runs one branch, with no exec mask
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  switch (wave) {
    case 0: gram128_persist_loop<0, TS, WG>(z0, zorder, nzc, counter, &sh_item, sel_ptr, sel_idx, sel_w, m, rho, innov, HS, mat, vec, err, lane); break;
    case 1: gram128_persist_loop<1, TS, WG>(z0, zorder, nzc, counter, &sh_item, sel_ptr, sel_idx, sel_w, m, rho, innov, HS, mat, vec, err, lane); break;
    case 2: gram128_persist_loop<2, TS, WG>(z0, zorder, nzc, counter, &sh_item, sel_ptr, sel_idx, sel_w, m, rho, innov, HS, mat, vec, err, lane); break;
    default: gram128_persist_loop<3, TS, WG>(z0, zorder, nzc, counter, &sh_item, sel_ptr, sel_idx, sel_w, m, rho, innov, HS, mat, vec, err, lane); break;
  }
}

// Block one-sided Jacobi, NP = 128
// ---------------------------------------------------------------------------
struct Eig128Lds {
  double red[4][64][17];  // per-wave products, one row per lane
  double cs[4][16][2];    // rotations of the wave's 16 pairs
  double nt[4][16];       // squared norms of the top-block slots
  double nb[4][16];       // squared norms of the bottom-block slots
  double col[8][P128];    // epilogue: per-column / per-row values
  double part[2][P128];   // epilogue: two partial sums
  int flag;
};

// block held at circle position pos in block round t (8 blocks, position 0 fixed)
__device__ __forceinline__ int blk_at(int pos, int t) { return pos == 0 ? 0 : ((pos - 1 + t) % 7) + 1; }

// Sum over the 128 rows of 16 values per lane (the lane's two rows already
// added): lanes with equal (lane & 15) receive the total of column lane & 15.
__device__ __forceinline__ double colsum16(double (&red)[64][17], int lane, const double (&g)[16]) {
#pragma unroll
  for (int k = 0; k < 16; ++k) red[lane][k] = g[k];
  __builtin_amdgcn_wave_barrier();
  const int k = lane & 15, q = lane >> 4;
  double s = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) s += red[q * 16 + i][k];
  __builtin_amdgcn_wave_barrier();
  s += shfl_xor_d(s, 16);
  s += shfl_xor_d(s, 32);
  return s;
}

__device__ __forceinline__ void jacobi_cs(double gam, double alpha, double beta, double tol, bool& rot,
                                          double& c, double& s, double& t) {
  c = 1.0;
  s = 0.0;
  t = 0.0;
  if (fabs(gam) > tol * sqrt(alpha * beta)) {
    const double zeta = (beta - alpha) / (2.0 * gam);
    t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
    c = 1.0 / sqrt(1.0 + t * t);
    s = c * t;
    rot = true;
  }
}

__global__ __launch_bounds__(256) void k_eig128(int64_t z0, const int32_t* __restrict__ zorder, int nzc, int N,
                                                const int64_t* __restrict__ sel_ptr,
                                                const double* __restrict__ mat,
                                                double* __restrict__ ymat, double* __restrict__ vec,
                                                double* __restrict__ ampl,
                                                int32_t* __restrict__ sweeps_out,
                                                unsigned* __restrict__ err, double tol,
                                                int max_sweeps,
                                                const int32_t* __restrict__ only) {
  __shared__ Eig128Lds L;
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;  // the zone is solved by the other pipeline
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  if (only && !(only[zl] & 3)) return;
  double* V = vec + (size_t)zl * 4 * P128;
  if (sel_ptr[z + 1] == sel_ptr[z]) {
    if (tid < P128) V[P128 + tid] = 0.0;
    if (tid == 0) sweeps_out[zl] = 0;
    return;
  }
  const double* M = mat + (size_t)zl * P128 * P128;
  double* X = ymat + (size_t)zl * P128 * P128;  // scratch for the block exchanges, then Y
  const int r0 = lane, r1 = lane + 64;
  double xt[2][16], xb[2][16];
  int t_blk = 0;  // block-round counter (positions rotate with period 7)
  int bt = blk_at(wave, 0), bb = blk_at(7 - wave, 0);
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    xt[0][j] = M[r0 * P128 + 16 * bt + j];
    xt[1][j] = M[r1 * P128 + 16 * bt + j];
    xb[0][j] = M[r0 * P128 + 16 * bb + j];
    xb[1][j] = M[r1 * P128 + 16 * bb + j];
  }
  auto norms = [&]() {
    double g[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = xt[0][k] * xt[0][k] + xt[1][k] * xt[1][k];
    const double a = colsum16(L.red[wave], lane, g);
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = xb[0][k] * xb[0][k] + xb[1][k] * xb[1][k];
    const double b = colsum16(L.red[wave], lane, g);
    if (lane < 16) {
      L.nt[wave][lane] = a;
      L.nb[wave][lane] = b;
    }
    __builtin_amdgcn_wave_barrier();
  };
  int sweeps = 0;
  for (int sw = 0; sw < max_sweeps; ++sw) {
    bool rotated = false;
    // ---- intra-block pairs: circle ordering on each 16-column block ----
    norms();
    for (int rd = 0; rd < 15; ++rd) {
      double g[16];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        g[i] = xt[0][i] * xt[0][15 - i] + xt[1][i] * xt[1][15 - i];
        g[8 + i] = xb[0][i] * xb[0][15 - i] + xb[1][i] * xb[1][15 - i];
      }
      const double gam = colsum16(L.red[wave], lane, g);
      const int k = lane & 15, i = k & 7;
      const bool top = k < 8;
      const double alpha = top ? L.nt[wave][i] : L.nb[wave][i];
      const double beta = top ? L.nt[wave][15 - i] : L.nb[wave][15 - i];
      double c, s, t;
      jacobi_cs(gam, alpha, beta, tol, rotated, c, s, t);
      __builtin_amdgcn_wave_barrier();
      if (lane < 16) {
        L.cs[wave][k][0] = c;
        L.cs[wave][k][1] = s;
        const int sp = (i == 0) ? 0 : i + 1, sq = (i == 0) ? 1 : 16 - i;
        if (top) {
          L.nt[wave][sp] = alpha - t * gam;
          L.nt[wave][sq] = beta + t * gam;
        } else {
          L.nb[wave][sp] = alpha - t * gam;
          L.nb[wave][sq] = beta + t * gam;
        }
      }
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int ii = 0; ii < 8; ++ii) {
        const double ct = L.cs[wave][ii][0], st = L.cs[wave][ii][1];
        const double cb = L.cs[wave][8 + ii][0], sb = L.cs[wave][8 + ii][1];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const double p1 = xt[h][ii], q1 = xt[h][15 - ii];
          xt[h][ii] = ct * p1 - st * q1;
          xt[h][15 - ii] = st * p1 + ct * q1;
          const double p2 = xb[h][ii], q2 = xb[h][15 - ii];
          xb[h][ii] = cb * p2 - sb * q2;
          xb[h][15 - ii] = sb * p2 + cb * q2;
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double tt = xt[h][15], tb = xb[h][15];
#pragma unroll
        for (int j = 15; j >= 2; --j) {
          xt[h][j] = xt[h][j - 1];
          xb[h][j] = xb[h][j - 1];
        }
        xt[h][1] = tt;
        xb[h][1] = tb;
      }
    }
    // ---- cross-block pairs: 7 block rounds of 16 rounds ----
    for (int br = 0; br < 7; ++br) {
      if (br > 0) {
        // move blocks between waves through the scratch matrix
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          X[r0 * P128 + 16 * bt + j] = xt[0][j];
          X[r1 * P128 + 16 * bt + j] = xt[1][j];
          X[r0 * P128 + 16 * bb + j] = xb[0][j];
          X[r1 * P128 + 16 * bb + j] = xb[1][j];
        }
        __syncthreads();
        ++t_blk;
        bt = blk_at(wave, t_blk % 7);
        bb = blk_at(7 - wave, t_blk % 7);
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          xt[0][j] = X[r0 * P128 + 16 * bt + j];
          xt[1][j] = X[r1 * P128 + 16 * bt + j];
          xb[0][j] = X[r0 * P128 + 16 * bb + j];
          xb[1][j] = X[r1 * P128 + 16 * bb + j];
        }
        __syncthreads();
      }
      norms();
      for (int rd = 0; rd < 16; ++rd) {
        double g[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) g[i] = xt[0][i] * xb[0][i] + xt[1][i] * xb[1][i];
        const double gam = colsum16(L.red[wave], lane, g);
        const int k = lane & 15;
        const double alpha = L.nt[wave][k], beta = L.nb[wave][k];
        double c, s, t;
        jacobi_cs(gam, alpha, beta, tol, rotated, c, s, t);
        __builtin_amdgcn_wave_barrier();
        if (lane < 16) {
          L.cs[wave][k][0] = c;
          L.cs[wave][k][1] = s;
          L.nt[wave][k] = alpha - t * gam;
          L.nb[wave][(k + 15) & 15] = beta + t * gam;
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const double ci = L.cs[wave][i][0], si = L.cs[wave][i][1];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const double p1 = xt[h][i], q1 = xb[h][i];
            xt[h][i] = ci * p1 - si * q1;
            xb[h][i] = si * p1 + ci * q1;
          }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const double b0 = xb[h][0];
#pragma unroll
          for (int j = 0; j < 15; ++j) xb[h][j] = xb[h][j + 1];
          xb[h][15] = b0;
        }
      }
    }
    ++sweeps;
    // convergence: no rotation in any wave during the sweep
    if (tid == 0) L.flag = 0;
    __syncthreads();
    if (__ballot(rotated) != 0 && lane == 0) atomicOr(&L.flag, 1);
    __syncthreads();
    if (L.flag == 0) break;
  }
  // final matrix X into the scratch
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    X[r0 * P128 + 16 * bt + j] = xt[0][j];
    X[r1 * P128 + 16 * bt + j] = xt[1][j];
    X[r0 * P128 + 16 * bb + j] = xb[0][j];
    X[r1 * P128 + 16 * bb + j] = xb[1][j];
  }
  __syncthreads();

  // ---- epilogue over X (all 256 threads; column j = tid & 127, half h = tid >> 7) ----
  const int jc = tid & 127, hh = tid >> 7;
  // per-column sums over a half of the rows: sum_r X[r][j] * f(r)
  auto colsum = [&](auto f) -> double {
    double s = 0;
    for (int r = 64 * hh; r < 64 * hh + 64; ++r) s = fma(X[r * P128 + jc], f(r), s);
    L.part[hh][jc] = s;
    __syncthreads();
    const double tot = L.part[0][jc] + L.part[1][jc];
    __syncthreads();
    return tot;
  };
  // per-row sums over a half of the columns: sum_j X[r][j] * L.col[ci][j]
  auto rowsum = [&](int ci) -> double {
    const int r = jc;
    double s = 0;
    for (int j = 64 * hh; j < 64 * hh + 64; ++j) s = fma(X[r * P128 + j], L.col[ci][j], s);
    L.part[hh][r] = s;
    __syncthreads();
    const double tot = L.part[0][r] + L.part[1][r];
    __syncthreads();
    return tot;
  };
  // column norms (eigenvalues of I + C), clamp at 1 (rrsqrt.F90:137)
  const double nrm2 = colsum([&](int r) { return X[r * P128 + jc]; });
  const double mun = sqrt(nrm2), muc = fmax(mun, 1.0);
  if (hh == 0) {
    L.col[0][jc] = V[jc];                       // b
    L.col[1][jc] = jc < N ? 1.0 : 0.0;          // 1_N
  }
  __syncthreads();
  const double c1 = colsum([&](int r) { return L.col[0][r]; });
  const double c2 = colsum([&](int r) { return L.col[1][r]; });
  const double sa = 1.0 / (nrm2 * muc);
  if (hh == 0) {
    L.col[2][jc] = c1 * sa;
    L.col[3][jc] = c2 * sqrt(muc) / nrm2;
  }
  __syncthreads();
  const double ar0 = rowsum(2);   // a (row index jc)
  const double vr = rowsum(3);    // v before normalisation
  // refinement a += M^-1 (b - M a)
  if (hh == 0) L.col[4][jc] = ar0;
  __syncthreads();
  double ma = 0;
  for (int j = 64 * hh; j < 64 * hh + 64; ++j) ma = fma(M[jc * P128 + j], L.col[4][j], ma);
  L.part[hh][jc] = ma;
  __syncthreads();
  const double mat_a = L.part[0][jc] + L.part[1][jc];
  __syncthreads();
  if (hh == 0) L.col[5][jc] = L.col[0][jc] - mat_a;  // residual b - M a
  __syncthreads();
  const double c3 = colsum([&](int r) { return L.col[5][r]; });
  if (hh == 0) L.col[6][jc] = c3 * sa;
  __syncthreads();
  const double da = rowsum(6);
  const double ar = ar0 + da;
  // |v|
  if (hh == 0) L.col[7][jc] = vr * vr;
  __syncthreads();
  double nv = 0;
  for (int j = 0; j < P128; ++j) nv += L.col[7][j];
  const double vn = vr / sqrt(nv);
  // Y = X diag(muc^-1/4 / mun)
  if (hh == 0) L.col[6][jc] = 1.0 / (sqrt(sqrt(muc)) * mun);
  __syncthreads();
  for (int e = tid; e < P128 * P128; e += 256) X[e] = X[e] * L.col[6][e & 127];
  if (hh == 0) {
    V[P128 + jc] = ar;
    V[2 * P128 + jc] = vn;
    V[3 * P128 + jc] = muc;
    if (ampl && jc < N) ampl[(size_t)z * N + jc] = ar;
    if (ar != ar) atomicOr(err, OAKG_DERR_NAN);
  }
  if (tid == 0) sweeps_out[zl] = sweeps;
}

// ---------------------------------------------------------------------------
// Transform and update fused, NP = 128, FP64: T = (Y Y^T) Omega is formed in
// LDS and out = Sf_z T, xa = x0 + Sf_z a run from it, so T never goes through
// HBM.  One workgroup per zone (the 128 x 128 block fills the LDS):
//   1. Y -> LDS (row stride LDT: the MFMA operand reads, 16 rows x 4
//      consecutive doubles per instruction, hit distinct banks)
//   2. Y Y^T on v_mfma_f64_16x16x4f64, the 36 lower tiles (wave w owns tile
//      rows w and 7 - w, 9 tiles each), stored mirrored over Y
//   3. the Omega correction (the rank-structured formula of k_tmat in
//      analysis.hip), applied in place; the buffer then holds X[s][r] = T[r][s]
//   4. the update (Out^T strip = X Sf_strip^T) with the A operand X read
//      from LDS and each wave's first strip of Sf loaded during steps 1-3
// ---------------------------------------------------------------------------
constexpr int LDT = P128 + 2;

struct XformLds {
  double t[P128 * LDT];
  double u[4][P128];   // quarter sums of T v
  double w[4][P128];   // quarter sums of T over the first N-1 entries
  double v[P128];
  double a[P128 + 2];  // a, then v.w~
};

template <int MODE, typename TE>
__global__ __launch_bounds__(512) void k_xform128(int64_t z0, const int32_t* __restrict__ zorder, int N, const int64_t* __restrict__ sel_ptr,
                                                  const int64_t* __restrict__ zone_ptr,
                                                  const double* __restrict__ ymat,
                                                  const double* __restrict__ vec,
                                                  const double* __restrict__ X0,
                                                  const TE* __restrict__ S, int64_t ldS,
                                                  double* __restrict__ xa, TE* __restrict__ Out,
                                                  int64_t ldO, double scaling, double infl,
                                                  double maxcorr,
                                                  const double* __restrict__ maxcorr_v) {
#pragma clang fp contract(off)
  // FP64 state, update on FP64 MFMA (MODE 0: locAnalysis with Sf given,
  // MODE 1: ensemble); FP32 members run in k_xform128f.
  constexpr bool F32 = std::is_same<TE, float>::value;
  static_assert(!F32, "FP32 state runs in k_xform128f");
  __shared__ XformLds L;
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, kk = lane >> 4;
  const int64_t i0 = zone_ptr[z];
  const int64_t sz = zone_ptr[z + 1] - i0;
  // The k index of the MFMA steps is permuted within pairs of steps: in
  // steps 2p and 2p+1 lane group kk covers k = 8p + 2kk and 8p + 2kk + 1, so
  // one 16-byte LDS read gives a lane both A operands.  B uses the same map.
  // Sf values of strip rb: b[q] = Sf[rb + li][k(q)], k(q) = 8 (q/2) + 2 kk + q%2,
  // loaded raw from clamped addresses; the selection and centring are
  // applied where the values are first used, so the loads stay in flight
  TE b[P128 / 4];
  double x0 = 0.0;
  auto load_strip = [&](int64_t rb) {
    const int64_t row = rb + li;
    const int64_t rr = i0 + (row < sz ? row : 0);
    x0 = X0[rr];
#pragma unroll
    for (int q = 0; q < P128 / 4; ++q) {
      const int k = 8 * (q >> 1) + 2 * kk + (q & 1);
      b[q] = S[rr + (int64_t)(k < N ? k : 0) * ldS];
    }
  };
  const double* Vz = vec + (size_t)zl * 4 * P128;
  // a zone solved in observation space (dual.hip) has -1 in its mu slot and
  // W in place of Y: M^-1/2 = I - W W^T instead of Y Y^T
  const bool dual = Vz[3 * P128] < 0.0;
  if (tid < P128) L.a[tid] = Vz[P128 + tid];
  else if (tid < 2 * P128) L.v[tid - P128] = tid - P128 < N ? Vz[2 * P128 + tid - P128] : 0.0;
  if (sel_ptr[z + 1] == sel_ptr[z]) {
    // no observations: T = I (a = 0), the update copies the forecast anomalies
    for (int e = tid; e < P128 * P128; e += 512) L.t[(e >> 7) * LDT + (e & 127)] = (e >> 7) == (e & 127);
    if (sz > 0) load_strip(16 * wave);
    lds_barrier();
  } else {
    // Y is stored to LDS first, then the wave's first strip of Sf is
    // loaded; the barriers of this kernel order LDS only, so the Sf loads
    // stay in flight while the transform is formed
    const double2* Y2 = reinterpret_cast<const double2*>(ymat + (size_t)zl * P128 * P128);
#pragma unroll
    for (int it = 0; it < P128 * P128 / 1024; ++it) {
      const int e = tid + 512 * it;
      *reinterpret_cast<double2*>(L.t + (e >> 6) * LDT + 2 * (e & 63)) = Y2[e];
    }
    __builtin_amdgcn_sched_barrier(0);
    if (sz > 0) load_strip(16 * wave);
    lds_barrier();
    {
      // 36 lower tiles over 8 waves: the pair of tile rows R0 = w % 4 and
      // R1 = 7 - w % 4 has 9 tiles; wave w < 4 takes (R0, 0..R0) and
      // (R1, 0..3-R0), wave w + 4 takes (R1, 4-R0..7-R0).
      const int R0 = wave & 3, R1 = 7 - R0;
      const bool hi = wave >= 4;
      const int jlo = hi ? 4 - R0 : 0, jhi = hi ? 7 - R0 : 3 - R0;  // row R1 range
      d4 acc0[4], acc1[4];
#pragma unroll
      for (int J = 0; J < 4; ++J) acc0[J] = acc1[J] = d4{0, 0, 0, 0};
      const double* yr = L.t + li * LDT + 2 * kk;
      for (int p = 0; p < P128 / 8; ++p) {
        asm volatile("" ::: "memory");
        const double2 y0 = *reinterpret_cast<const double2*>(yr + 16 * R0 * LDT + 8 * p);
        const double2 y1 = *reinterpret_cast<const double2*>(yr + 16 * R1 * LDT + 8 * p);
        // tile columns jlo + J (for the waves w < 4, jlo = 0: also the columns of row R0)
        double2 yb[4];
#pragma unroll
        for (int J = 0; J < 4; ++J)
          yb[J] = *reinterpret_cast<const double2*>(yr + 16 * (jlo + J) * LDT + 8 * p);
        if (!hi) {
#pragma unroll
          for (int J = 0; J < 4; ++J)
            if (J <= R0) {
              acc0[J] = mfma_f64(y0.x, yb[J].x, acc0[J]);
              acc0[J] = mfma_f64(y0.y, yb[J].y, acc0[J]);
            }
        }
#pragma unroll
        for (int J = 0; J < 4; ++J)
          if (jlo + J <= jhi) {
            acc1[J] = mfma_f64(y1.x, yb[J].x, acc1[J]);
            acc1[J] = mfma_f64(y1.y, yb[J].y, acc1[J]);
          }
      }
      lds_barrier();  // every wave has read Y
      // D[16 I + lane / 16 + 4 r][16 J + lane % 16]; mirrored entries keep T symmetric
      auto put = [&](int I, int J, const d4& acc) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = 16 * I + kk + 4 * r, n = 16 * J + li;
          const double tv = dual ? (m == n ? 1.0 : 0.0) - acc[r] : acc[r];
          L.t[m * LDT + n] = tv;
          L.t[n * LDT + m] = tv;
        }
      };
      if (!hi) {
#pragma unroll
        for (int J = 0; J < 4; ++J)
          if (J <= R0) put(R0, J, acc0[J]);
      }
#pragma unroll
      for (int J = 0; J < 4; ++J)
        if (jlo + J <= jhi) put(R1, jlo + J, acc1[J]);
    }
    lds_barrier();
    // Omega = v w^T + P(v) P(w)^T (rrsqrt.F90:737-756, matoper.F90:509-535)
    // applied on the right of the symmetric T (as k_tmat in analysis.hip).
    // With w = 1/sqrt(N) 1,
    // alpha_v = -1/(|v_N| + 1), alpha_w = -1/(1/sqrt(N) + 1), row r of T Omega is
    //   x[s] = T[r][s] - [s = N-1] T[r][N-1] + wN (T v)_r
    //          + (alpha_w U2 + alpha_v alpha_w (v.w~) U3) zw[s] + alpha_v U3 vt[s]
    // with U2 = wN sum_{q<N-1} T[r][q], U3 = (T v)_r + sign(v_N) T[r][N-1],
    // zw[s] = wN + [s = N-1], vt[s] = v[s] for s < N-1 (0 at N-1); the row
    // sums are read as column sums of the symmetric T.
    const double wN = 1.0 / sqrt((double)N);
    const double vN = L.v[N - 1];
    const double sgv = vN >= 0 ? 1.0 : -1.0;
    const double alv = -1.0 / (fabs(vN) + 1.0);
    const double alw = -1.0 / (wN + 1.0);
    {
      const int r = tid & 127, h = tid >> 7;  // quarter h of the column
      double su = 0, sw = 0;
#pragma unroll 8
      for (int q = 32 * h; q < 32 * h + 32; ++q) {
        const double tq = L.t[q * LDT + r];
        su = fma(tq, L.v[q], su);
        sw += q < N - 1 ? tq : 0.0;
      }
      L.u[h][r] = su;
      L.w[h][r] = sw;
    }
    if (wave == 0) {
      double s = 0;
      for (int q = lane; q < N - 1; q += 64) s = fma(L.v[q], wN, s);
      s = wave_sum_fast(s);
      if (lane == 0) L.a[P128] = s;
    }
    lds_barrier();
    const int r = tid & 127;
    const double u0 = L.t[(N - 1) * LDT + r];
    const double u1 = (L.u[0][r] + L.u[1][r]) + (L.u[2][r] + L.u[3][r]);
    const double u2 = wN * ((L.w[0][r] + L.w[1][r]) + (L.w[2][r] + L.w[3][r]));
    const double u3 = u1 + sgv * u0;
    const double vw = L.a[P128];
    const double c1 = u1 * wN;
    const double c2 = alw * u2 + alv * alw * vw * u3;
    const double c3 = alv * u3;
    // X[s][r] = (T Omega)[r][s], in place (T symmetric: T[r][s] = t[s][r])
#pragma unroll 4
    for (int s = tid >> 7; s < N; s += 4) {
      const double zw = s == N - 1 ? wN + 1.0 : wN;
      const double vt = s < N - 1 ? L.v[s] : 0.0;
      double x = L.t[s * LDT + r];
      if (s == N - 1) x -= u0;
      x += c1 + c2 * zw + c3 * vt;
      L.t[s * LDT + r] = x;
    }
    lds_barrier();
  }
  // out = Sf_z T: Out^T strip = T^T Sf_strip^T, A operand X[16 J + li][k]
  const double* xr = L.t + li * LDT + 2 * kk;
  for (int64_t rb = 16 * wave; rb < sz; rb += 128) {
    if (rb != 16 * wave) load_strip(rb);
    const int64_t row = rb + li;
    const bool rok = row < sz;
    double dxp = 0.0;
#pragma unroll
    for (int q = 0; q < P128 / 4; ++q) {
      const bool ok = rok && (8 * (q >> 1) + 2 * kk + (q & 1)) < N;
      b[q] = ok ? ((MODE == 1) ? (b[q] - x0) / scaling : b[q]) : 0.0;
    }
#pragma unroll
    for (int p = 0; p < P128 / 8; ++p) {
      const double2 ak = *reinterpret_cast<const double2*>(L.a + 8 * p + 2 * kk);
      dxp = fma(b[2 * p], ak.x, dxp);
      dxp = fma(b[2 * p + 1], ak.y, dxp);
    }
    dxp += shfl_xor_d(dxp, 16);
    dxp += shfl_xor_d(dxp, 32);
    double xar = x0 + dxp;
    if (MODE == 1) {
      const double mc = (maxcorr_v && rok) ? maxcorr_v[i0 + row] : maxcorr;
      if (xar - mc > x0) xar = x0 + mc;
      if (xar < x0 - mc) xar = x0 - mc;
    }
    if (MODE == 0 && rok && kk == 0) xa[i0 + row] = xar;
    // two passes over the output tiles (4 each) keep the accumulators small
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      const double* xh = xr + 64 * h * LDT;
      d4 acc[4];
#pragma unroll
      for (int J = 0; J < 4; ++J) acc[J] = d4{0, 0, 0, 0};
#pragma unroll
      for (int p = 0; p < P128 / 8; ++p) {
        // keeps the compiler from hoisting the reads of later pairs; the
        // other wave on the SIMD covers the read latency
        asm volatile("" ::: "memory");
        double2 A[4];
#pragma unroll
        for (int J = 0; J < 4; ++J) A[J] = *reinterpret_cast<const double2*>(xh + 16 * J * LDT + 8 * p);
#pragma unroll
        for (int J = 0; J < 4; ++J) acc[J] = mfma_f64(A[J].x, b[2 * p], acc[J]);
#pragma unroll
        for (int J = 0; J < 4; ++J) acc[J] = mfma_f64(A[J].y, b[2 * p + 1], acc[J]);
      }
      if (rok) {
#pragma unroll
        for (int J = 0; J < 4; ++J) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int col = 64 * h + 16 * J + kk + 4 * r;
            if (col < N) {
              double o = acc[J][r];
              if (MODE == 1) {
                const double sa = (infl != 1.0) ? infl * o : o;
                o = xar + sa * scaling;
              }
              Out[(i0 + row) + (int64_t)col * ldO] = o;
            }
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// The same for FP32 members (mixed precision): Y is staged in LDS rounded to
// FP32, Y Y^T runs on v_mfma_f32_16x16x4f32 and T Omega is kept in FP32, the
// operand precision of the FP32 update that consumes it (the Omega row sums
// and xa = x0 + Sf a stay in FP64).  The buffers take 78 KB, so two zones
// share a CU; four waves per zone: wave w owns tile rows w and 7 - w.
//   FP32 MFMA layout: A (16x4) lane l holds A[l % 16][l / 16], B (4x16)
//   B[l / 16][l % 16], D (16x16) register r holds D[4 (l / 16) + r][l % 16].
// ---------------------------------------------------------------------------
constexpr int LDF = P128 + 4;

struct XformLdsF {
  float t[P128 * LDF];
  double u[2][P128];   // half sums of T v
  double w[2][P128];   // half sums of T over the first N-1 entries
  double v[P128];
  double a[P128 + 2];  // a / scaling, then v.w~
};

__global__ __launch_bounds__(256) void k_xform128f(int64_t z0, const int32_t* __restrict__ zorder, int N, const int64_t* __restrict__ sel_ptr,
                                                   const int64_t* __restrict__ zone_ptr,
                                                   const double* __restrict__ ymat,
                                                   const double* __restrict__ vec,
                                                   const double* __restrict__ X0,
                                                   const float* __restrict__ S, int64_t ldS,
                                                   float* __restrict__ Out, int64_t ldO,
                                                   double scaling, double infl, double maxcorr,
                                                   const double* __restrict__ maxcorr_v) {
#pragma clang fp contract(off)
  __shared__ XformLdsF L;
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, kk = lane >> 4;
  const int64_t i0 = zone_ptr[z];
  const int64_t sz = zone_ptr[z + 1] - i0;
  // k permuted within pairs of MFMA steps (as in k_xform128): in steps 2p and
  // 2p+1 lane group kk covers k = 8p + 2kk and 8p + 2kk + 1
  float b[P128 / 4];
  double x0 = 0.0;
  auto load_strip = [&](int64_t rb) {
    const int64_t row = rb + li;
    const int64_t rr = i0 + (row < sz ? row : 0);
    x0 = X0[rr];
#pragma unroll
    for (int q = 0; q < P128 / 4; ++q) {
      const int k = 8 * (q >> 1) + 2 * kk + (q & 1);
      b[q] = S[rr + (int64_t)(k < N ? k : 0) * ldS];
    }
  };
  const double* Vz = vec + (size_t)zl * 4 * P128;
  const bool dual = Vz[3 * P128] < 0.0;  // T = (I - W W^T) Omega (dual.hip)
  if (tid < P128) L.a[tid] = Vz[P128 + tid] / scaling;
  else L.v[tid - P128] = tid - P128 < N ? Vz[2 * P128 + tid - P128] : 0.0;
  if (sel_ptr[z + 1] == sel_ptr[z]) {
    // no observations: T = I (a = 0), the update copies the forecast anomalies
    for (int e = tid; e < P128 * P128; e += 256)
      L.t[(e >> 7) * LDF + (e & 127)] = (e >> 7) == (e & 127) ? 1.0f : 0.0f;
    if (sz > 0) load_strip(16 * wave);
    lds_barrier();
  } else {
    const double2* Y2 = reinterpret_cast<const double2*>(ymat + (size_t)zl * P128 * P128);
#pragma unroll
    for (int it0 = 0; it0 < P128 * P128 / 512; it0 += 16) {
      double2 yv[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) yv[u] = Y2[tid + 256 * (it0 + u)];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int e = tid + 256 * (it0 + u);
        *reinterpret_cast<float2*>(L.t + (e >> 6) * LDF + 2 * (e & 63)) =
            make_float2((float)yv[u].x, (float)yv[u].y);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    if (sz > 0) load_strip(16 * wave);
    lds_barrier();
    {
      const int R0 = wave, R1 = 7 - wave;  // tiles (R0, 0..R0) and (R1, 0..R1)
      f4 acc0[4], acc1[8];
#pragma unroll
      for (int J = 0; J < 4; ++J) acc0[J] = f4{0, 0, 0, 0};
#pragma unroll
      for (int J = 0; J < 8; ++J) acc1[J] = f4{0, 0, 0, 0};
      const float* yr = L.t + li * LDF + 2 * kk;
#pragma unroll 1
      for (int p = 0; p < P128 / 8; ++p) {
        const float2 y0 = *reinterpret_cast<const float2*>(yr + 16 * R0 * LDF + 8 * p);
        const float2 y1 = *reinterpret_cast<const float2*>(yr + 16 * R1 * LDF + 8 * p);
        float2 yb[8];
#pragma unroll
        for (int J = 0; J < 8; ++J) yb[J] = *reinterpret_cast<const float2*>(yr + 16 * J * LDF + 8 * p);
#pragma unroll
        for (int J = 0; J < 4; ++J)
          if (J <= R0) {
            acc0[J] = __builtin_amdgcn_mfma_f32_16x16x4f32(y0.x, yb[J].x, acc0[J], 0, 0, 0);
            acc0[J] = __builtin_amdgcn_mfma_f32_16x16x4f32(y0.y, yb[J].y, acc0[J], 0, 0, 0);
          }
#pragma unroll
        for (int J = 0; J < 8; ++J)
          if (J <= R1) {
            acc1[J] = __builtin_amdgcn_mfma_f32_16x16x4f32(y1.x, yb[J].x, acc1[J], 0, 0, 0);
            acc1[J] = __builtin_amdgcn_mfma_f32_16x16x4f32(y1.y, yb[J].y, acc1[J], 0, 0, 0);
          }
      }
      lds_barrier();  // every wave has read Y
      // D[16 I + 4 (lane / 16) + r][16 J + lane % 16]; mirrored entries keep T symmetric
      auto put = [&](int I, int J, const f4& acc) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = 16 * I + 4 * kk + r, n = 16 * J + li;
          const float tv = dual ? (m == n ? 1.0f : 0.0f) - acc[r] : acc[r];
          L.t[m * LDF + n] = tv;
          L.t[n * LDF + m] = tv;
        }
      };
#pragma unroll
      for (int J = 0; J < 4; ++J)
        if (J <= R0) put(R0, J, acc0[J]);
#pragma unroll
      for (int J = 0; J < 8; ++J)
        if (J <= R1) put(R1, J, acc1[J]);
    }
    lds_barrier();
    // Omega on the right of the symmetric T, as in k_xform128 (row sums read
    // as column sums, accumulated in FP64)
    const double wN = 1.0 / sqrt((double)N);
    const double vN = L.v[N - 1];
    const double sgv = vN >= 0 ? 1.0 : -1.0;
    const double alv = -1.0 / (fabs(vN) + 1.0);
    const double alw = -1.0 / (wN + 1.0);
    {
      const int r = tid & 127, h = tid >> 7;  // half h of the column
      double su = 0, sw = 0;
#pragma unroll 8
      for (int q = 64 * h; q < 64 * h + 64; ++q) {
        const double tq = L.t[q * LDF + r];
        su = fma(tq, L.v[q], su);
        sw += q < N - 1 ? tq : 0.0;
      }
      L.u[h][r] = su;
      L.w[h][r] = sw;
    }
    if (wave == 0) {
      double s = 0;
      for (int q = lane; q < N - 1; q += 64) s = fma(L.v[q], wN, s);
      s = wave_sum_fast(s);
      if (lane == 0) L.a[P128] = s;
    }
    lds_barrier();
    const int r = tid & 127;
    const double u0 = L.t[(N - 1) * LDF + r];
    const double u1 = L.u[0][r] + L.u[1][r];
    const double u2 = wN * (L.w[0][r] + L.w[1][r]);
    const double u3 = u1 + sgv * u0;
    const double vw = L.a[P128];
    const double c1 = u1 * wN;
    const double c2 = alw * u2 + alv * alw * vw * u3;
    const double c3 = alv * u3;
    // X[s][r] = (T Omega)[r][s], in place (T symmetric: T[r][s] = t[s][r])
#pragma unroll 4
    for (int s = tid >> 7; s < N; s += 2) {
      const double zw = s == N - 1 ? wN + 1.0 : wN;
      const double vt = s < N - 1 ? L.v[s] : 0.0;
      double x = L.t[s * LDF + r];
      if (s == N - 1) x -= u0;
      x += c1 + c2 * zw + c3 * vt;
      L.t[s * LDF + r] = (float)x;
    }
    lds_barrier();
  }
  // out = Sf_z T: Out^T strip = T^T Sf_strip^T (as the FP32 branch of k_xform128)
  const float* xr = L.t + li * LDF + 2 * kk;
  const double sc_infl = scaling * infl;
  const float inv = (float)(1.0 / scaling);
  for (int64_t rb = 16 * wave; rb < sz; rb += 64) {
    if (rb != 16 * wave) load_strip(rb);
    const int64_t row = rb + li;
    const bool rok = row < sz;
    const float xb = (float)x0;
    double dxp = 0.0;
#pragma unroll
    for (int q = 0; q < P128 / 4; ++q) {
      const int kc = 8 * (q >> 1) + 2 * kk + (q & 1);
      const bool ok = rok && kc < N;
      const float ev = b[q];
      if (ok) dxp = fma((double)ev - x0, L.a[kc], dxp);
      b[q] = ok ? (ev - xb) * inv : 0.0f;
    }
    dxp += shfl_xor_d(dxp, 16);
    dxp += shfl_xor_d(dxp, 32);
    double xar = x0 + dxp;
    const double mc = (maxcorr_v && rok) ? maxcorr_v[i0 + row] : maxcorr;
    if (xar - mc > x0) xar = x0 + mc;
    if (xar < x0 - mc) xar = x0 - mc;
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      const float* xh = xr + 64 * h * LDF;
      f4 acc[4];
#pragma unroll
      for (int J = 0; J < 4; ++J) acc[J] = f4{0, 0, 0, 0};
#pragma unroll
      for (int p = 0; p < P128 / 8; ++p) {
        asm volatile("" ::: "memory");
        float2 A[4];
#pragma unroll
        for (int J = 0; J < 4; ++J) A[J] = *reinterpret_cast<const float2*>(xh + 16 * J * LDF + 8 * p);
#pragma unroll
        for (int J = 0; J < 4; ++J)
          acc[J] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[J].x, b[2 * p], acc[J], 0, 0, 0);
#pragma unroll
        for (int J = 0; J < 4; ++J)
          acc[J] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[J].y, b[2 * p + 1], acc[J], 0, 0, 0);
      }
      if (rok) {
#pragma unroll
        for (int J = 0; J < 4; ++J) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int col = 64 * h + 16 * J + 4 * kk + r;
            if (col < N)
              Out[(i0 + row) + (int64_t)col * ldO] = (float)(xar + sc_infl * (double)acc[J][r]);
          }
        }
      }
    }
  }
}

int persist_grid128() {
  // the occupancy the launch bounds guarantee (the runtime's occupancy
  // query reports one workgroup per CU for these 256-VGPR waves)
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  return kPersistBlocksPerCU * ncu;
}

void launch_solve128(const ChunkArgs& A, hipStream_t st, PhaseTimer* tm) {
  if (tm) tm->mark(PH_GRAM);
  if (A.persist_wq) {
    hipLaunchKernelGGL(k_clear_wq, dim3(1), dim3(64), 0, st, A.persist_wq);
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(A.persist_grid, A.nzc));
    const dim3 g(grid), b(256);
    if (A.prec == 1 && A.persist_item == 1)
      hipLaunchKernelGGL((k_gram128_persist<float, true>), g, b, 0, st, A.z0, A.zorder, A.nzc,
                         A.persist_wq, A.sel_ptr, A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS32,
                         A.mat, A.vec, A.err);
    else if (A.prec == 1)
      hipLaunchKernelGGL((k_gram128_persist<float, false>), g, b, 0, st, A.z0, A.zorder, A.nzc,
                         A.persist_wq, A.sel_ptr, A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS32,
                         A.mat, A.vec, A.err);
    else if (A.persist_item == 1)
      hipLaunchKernelGGL((k_gram128_persist<double, true>), g, b, 0, st, A.z0, A.zorder, A.nzc,
                         A.persist_wq, A.sel_ptr, A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS,
                         A.mat, A.vec, A.err);
    else
      hipLaunchKernelGGL((k_gram128_persist<double, false>), g, b, 0, st, A.z0, A.zorder, A.nzc,
                         A.persist_wq, A.sel_ptr, A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS,
                         A.mat, A.vec, A.err);
  } else if (A.prec == 1)
    hipLaunchKernelGGL(k_gram128<float>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                       A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS32, A.mat, A.vec, A.err);
  else
    hipLaunchKernelGGL(k_gram128<double>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                       A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS, A.mat, A.vec, A.err);
  if (tm) tm->mark(PH_EIG);
  if (A.eig_method == 1) {
    launch_tql_pipeline(128, A, st);
    hipLaunchKernelGGL(k_eig128, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.nzc, A.N, A.sel_ptr,
                       A.mat, A.ymat, A.vec, A.ampl, A.sweeps, A.err, A.tol, A.max_sweeps,
                       (const int32_t*)A.tql_status);
  } else {
    hipLaunchKernelGGL(k_eig128, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.nzc, A.N, A.sel_ptr,
                       A.mat, A.ymat, A.vec, A.ampl, A.sweeps, A.err, A.tol, A.max_sweeps,
                       (const int32_t*)nullptr);
  }
}

void launch_apply128(const ChunkArgs& A, hipStream_t st, PhaseTimer* tm) {
  if (A.mode == 2) {  // innovation likelihood: adjoint instead of transform + update
    if (tm) tm->mark(PH_UPDATE);
    launch_nll_adjoint(128, A, st);
    if (tm) tm->mark(PH_END);
    return;
  }
  // transform and update in one kernel (the timer books it as the update)
  if (tm) tm->mark(PH_UPDATE);
  if (A.prec == 1)
    hipLaunchKernelGGL(k_xform128f, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.N, A.sel_ptr,
                       A.zone_ptr, A.ymat, A.vec, A.X0, A.S32, A.ldS, A.Out32, A.ldO, A.scaling,
                       A.infl, A.maxcorr, A.maxcorr_v);
  else if (A.mode == 0)
    hipLaunchKernelGGL((k_xform128<0, double>), dim3(A.nzc), dim3(512), 0, st, A.z0, A.zorder, A.N, A.sel_ptr,
                       A.zone_ptr, A.ymat, A.vec, A.X0, A.S, A.ldS, A.xa, A.Out, A.ldO, A.scaling,
                       A.infl, A.maxcorr, A.maxcorr_v);
  else
    hipLaunchKernelGGL((k_xform128<1, double>), dim3(A.nzc), dim3(512), 0, st, A.z0, A.zorder, A.N, A.sel_ptr,
                       A.zone_ptr, A.ymat, A.vec, A.X0, A.S, A.ldS, A.xa, A.Out, A.ldO, A.scaling,
                       A.infl, A.maxcorr, A.maxcorr_v);
  if (tm) tm->mark(PH_END);
}

void launch_chunk128(const ChunkArgs& A, hipStream_t st, PhaseTimer* tm) {
  launch_solve128(A, st, tm);
  launch_apply128(A, st, tm);
}

}  // namespace oakg
