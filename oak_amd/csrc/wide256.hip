// Analysis kernels for ensembles of 129..256 members (NP = 256).
//
// A zone's 256 x 256 matrix is 512 KB in FP64: it fits neither the
// registers of one workgroup nor its LDS, where the N <= 128 kernels keep
// a zone whole, so each phase splits it its own way:
//   k_gram256      M = I + S~^T S~ and b = S~^T R^-1 d: 16 waves per zone
//                  (four workgroups), each over all local observations for
//                  two or three 32 x 32 blocks of the lower triangle
//                  (v_mfma_f64_16x16x4f64).
//   k_tridiag_lower<256>  Householder reduction to tridiagonal form (lower, as
//                  dsytd2), steps 0..127, with the lower triangle in 8 x 8
//                  register tiles over 512 threads, one workgroup per zone:
//                  496 tiles hold columns 8..255, and columns 0..7 sit in
//                  LDS for the seven steps that change them; steps 128..253
//                  on the trailing 128 x 128 block by k_tridiag<128>.
//   k_formq256     Q = H_0 ... H_253 in blocked compact-WY form on MFMA (the
//                  scheme of k_formq128 on a 256 x 64 slice), four workgroups
//                  per zone, one 64-column block of Q each.
//   tridiag.hip    k_tql (QL on (d, e), rotations logged), k_tql_apply256
//                  (the rotations on the rows of Q in registers, two threads
//                  a row) and k_eigvec_epi: eigenvalues, U, a, v and Y.
//   k_tmat256      T = (Y Y^T) Omega on MFMA, four workgroups per zone, one
//                  64-row block of T each.
//   k_update256    out = S [T | a] for the zone's rows, eight workgroups per
//                  zone (one 32-column block of T each) on one XCD.
// The operations are those of the N <= 128 path (rrsqrt.F90:100-190 and the
// post-processing of assimilation.F90:3299-3363); only the split of the work
// and the order of the partial sums differ.  k_tridiag_lower<128> (the same
// lower-triangle reduction at NP = 128, steps 0..63) also serves the
// N = 128 path as its reduction options 3 and 4 (launch_tridiag_lower128),
// k_tridiag_lower<64, true> the trailing block of option 4 and the whole
// reduction at NP = 64 (with k_formq64 for Q).
#include "analysis.hpp"
#include "common.hpp"
#include "eigen_common.hpp"

namespace oakg {

namespace {
constexpr int NPW = 256;
}

// ---------------------------------------------------------------------------
// Gram.  The lower triangle of M in 32 x 32 blocks (2 x 2 tiles of 16 x 16):
// 28 off-diagonal blocks (4 tiles) and 8 diagonal ones (3 tiles), dealt to
// 16 wave slots per zone (four workgroups): slots 0..11 two off-diagonal
// blocks each (8 tiles), slots 12..15 one off-diagonal and two diagonal
// blocks (10 tiles), which also form b for the columns of their diagonal
// blocks (all 256 together).  Every wave runs over all the zone's
// observations, four per MFMA step, with the next step's gathers in flight
// (the scheme of k_gram); at about 200 registers two waves share a SIMD, so
// one wave's gathers and VALU work overlap another's MFMAs.
// ---------------------------------------------------------------------------
// off-diagonal block k (0..27) of the 8 x 8 block grid: (bi, bj), bi > bj
__device__ __forceinline__ void gram256_offdiag(int k, int& bi, int& bj) {
  int i = 1;
  while (k >= i) {
    k -= i;
    ++i;
  }
  bi = i;
  bj = k;
}

// MIX = false: off-diagonal blocks (X0, Y0) and (X1, Y1); MIX = true:
// off-diagonal block (X0, Y0) and diagonal blocks X1 and Y1.  The lane's
// values u[4 h + 2 s + k] = row[32 B + 16 k + jl] with B = (h ? (s ? Y1 : X1)
// : (s ? Y0 : X0)).
template <typename TS, bool MIX>
__device__ __forceinline__ void gram256_wave(int X0, int Y0, int X1, int Y1, int64_t p0, int64_t p,
                                             const int32_t* __restrict__ sel_idx,
                                             const double* __restrict__ sel_w, int64_t m,
                                             const double* __restrict__ rho,
                                             const double* __restrict__ innov,
                                             const TS* __restrict__ HS, double* __restrict__ M,
                                             double* __restrict__ bv_out,
                                             unsigned* __restrict__ err) {
  constexpr int NP = NPW;
  constexpr int NT = MIX ? 10 : 8;
  const int lane = threadIdx.x & 63, kq = lane >> 4, jl = lane & 15;
  d4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = d4{0, 0, 0, 0};
  double bacc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) bacc[k] = 0.0;
  const int64_t nsteps = (p + 3) >> 2;
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
  const int c0 = 32 * X0 + jl, c1 = 32 * Y0 + jl, c2 = 32 * X1 + jl, c3 = 32 * Y1 + jl;
  auto fetch_row = [&](Idx& x, double& rh, double& inn, TS* rv) {
    const bool inr = x.l >= 0 && x.l < m;
    bad = bad | (x.ok & !inr);
    x.ok = x.ok & inr;
    x.l = x.ok ? x.l : 0;
    rh = rho[x.l];
    inn = innov[x.l];
    const TS* row = HS + (size_t)x.l * NP;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      rv[k] = row[c0 + 16 * k];
      rv[2 + k] = row[c1 + 16 * k];
      rv[4 + k] = row[c2 + 16 * k];
      rv[6 + k] = row[c3 + 16 * k];
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
    asm volatile("" ::: "memory");  // compiler barrier: keeps the loads above here
    const bool ok = i0.ok;
    const double r = ok ? (i0.w * i0.w) * rh0 : 0.0;
    const double rd = ok ? r * inn0 : 0.0;
    double bv[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) bv[k] = ok ? (double)rv0[k] : 0.0;
    int t = 0;
#pragma unroll
    for (int I = 0; I < 2; ++I) {  // off-diagonal block (X0, Y0)
      const double a = r * bv[I];
#pragma unroll
      for (int J = 0; J < 2; ++J) {
        acc[t] = mfma_f64(a, bv[2 + J], acc[t]);
        ++t;
      }
    }
    if constexpr (MIX) {  // diagonal blocks X1 and Y1
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int I = 0; I < 2; ++I) {
          const double a = r * bv[4 + 2 * h + I];
#pragma unroll
          for (int J = 0; J <= I; ++J) {
            acc[t] = mfma_f64(a, bv[4 + 2 * h + J], acc[t]);
            ++t;
          }
        }
#pragma unroll
      for (int k = 0; k < 4; ++k) bacc[k] = fma(rd, bv[4 + k], bacc[k]);
    } else {  // off-diagonal block (X1, Y1)
#pragma unroll
      for (int I = 0; I < 2; ++I) {
        const double a = r * bv[4 + I];
#pragma unroll
        for (int J = 0; J < 2; ++J) {
          acc[t] = mfma_f64(a, bv[6 + J], acc[t]);
          ++t;
        }
      }
    }
    i0 = i1;
    i1 = i2;
    rh0 = rh1;
    inn0 = inn1;
#pragma unroll
    for (int k = 0; k < 8; ++k) rv0[k] = rv1[k];
  }
  if (bad) atomicOr(err, OAKG_DERR_IDX);
  if constexpr (MIX) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      bacc[k] += shfl_xor_d(bacc[k], 16);
      bacc[k] += shfl_xor_d(bacc[k], 32);
    }
    if (lane < 16) {
#pragma unroll
      for (int k = 0; k < 4; ++k) bv_out[32 * (k < 2 ? X1 : Y1) + 16 * (k & 1) + lane] = bacc[k];
    }
  }
  // tile (gi0, gj0): register q of lane l is C[l/16 + 4q][l%16]; a diagonal
  // tile writes its lower half into both triangles, so that M is exactly
  // symmetric
  auto put = [&](int gi0, int gj0, const d4& c, bool dtile) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = kq + 4 * q, j = jl;
      if (!dtile || i >= j) {
        const int gi = gi0 + i, gj = gj0 + j;
        const double v = c[q] + (gi == gj ? 1.0 : 0.0);
        M[(size_t)gi * NP + gj] = v;
        M[(size_t)gj * NP + gi] = v;
      }
    }
  };
  int t = 0;
#pragma unroll
  for (int I = 0; I < 2; ++I)
#pragma unroll
    for (int J = 0; J < 2; ++J) {
      put(32 * X0 + 16 * I, 32 * Y0 + 16 * J, acc[t], false);
      ++t;
    }
  if constexpr (MIX) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int I = 0; I < 2; ++I)
#pragma unroll
        for (int J = 0; J <= I; ++J) {
          const int B = h ? Y1 : X1;
          put(32 * B + 16 * I, 32 * B + 16 * J, acc[t], I == J);
          ++t;
        }
  } else {
#pragma unroll
    for (int I = 0; I < 2; ++I)
#pragma unroll
      for (int J = 0; J < 2; ++J) {
        put(32 * X1 + 16 * I, 32 * Y1 + 16 * J, acc[t], false);
        ++t;
      }
  }
}

// grid (zones, 4): workgroup y holds wave slots 4 y .. 4 y + 3.  (Giving each
// XCD a contiguous run of zones, so that neighbours share their rows of S in
// one L2, was slower: the densely observed zones are contiguous too, and
// their XCDs finish last.)
template <typename TS>
__global__ __launch_bounds__(256, 2) void k_gram256(int64_t z0, const int32_t* __restrict__ zorder,
                                                    const int64_t* __restrict__ sel_ptr,
                                                    const int32_t* __restrict__ sel_idx,
                                                    const double* __restrict__ sel_w, int64_t m,
                                                    const double* __restrict__ rho,
                                                    const double* __restrict__ innov,
                                                    const TS* __restrict__ HS,
                                                    double* __restrict__ mat,
                                                    double* __restrict__ vec,
                                                    unsigned* __restrict__ err) {
  constexpr int NP = NPW;
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;
  const int64_t p0 = sel_ptr[z];
  const int64_t p = sel_ptr[z + 1] - p0;
  double* M = mat + (size_t)zl * NP * NP;
  double* bv_out = vec + (size_t)zl * 4 * NP;
  if (p == 0) {  // M = I (rows 64 y .. 64 y + 63), b = 0
    const int y = blockIdx.y;
    for (int e = threadIdx.x; e < NP * NP / 4; e += 256) {
      const int g = y * (NP * NP / 4) + e;
      M[g] = (g / NP == g % NP) ? 1.0 : 0.0;
    }
    if (y == 0) bv_out[threadIdx.x] = 0.0;
    return;
  }
  const int ws = __builtin_amdgcn_readfirstlane((int)blockIdx.y * 4 + (int)(threadIdx.x >> 6));
  int X0, Y0, X1, Y1;
  if (ws < 12) {
    gram256_offdiag(2 * ws, X0, Y0);
    gram256_offdiag(2 * ws + 1, X1, Y1);
    gram256_wave<TS, false>(X0, Y0, X1, Y1, p0, p, sel_idx, sel_w, m, rho, innov, HS, M, bv_out,
                            err);
  } else {
    gram256_offdiag(24 + (ws - 12), X0, Y0);
    X1 = 2 * (ws - 12);
    Y1 = X1 + 1;
    gram256_wave<TS, true>(X0, Y0, X1, Y1, p0, p, sel_idx, sel_w, m, rho, innov, HS, M, bv_out,
                           err);
  }
}

// ---------------------------------------------------------------------------
// Householder reduction of M (lower, as dsytd2: H_i x = beta e_1 for the
// column below the diagonal) with the lower triangle in registers.  Thread t
// < 496 holds tile (tr, tc), 1 <= tc <= tr <= 31, of 8 x 8 (rows 8 tr.., columns
// 8 tc..), the tile columns in decreasing order so that the active tiles
// (those meeting the trailing block, 8 tc + 7 > i) are always threads
// 0 .. k - 1 and whole waves fall idle as the block shrinks.  A diagonal tile
// holds the full 8 x 8 block.  Columns 0..7 (every row) are in LDS: only
// steps 0..6 change them, and the row threads treat them.
//
// Step i: every wave forms the reflector from column i, which it carries in
// registers (x[h] = A[lane + 64 h][i]), and keeps its own copy of v in LDS.
// A v is summed per row from 32 partial sums: a row r of tile row R >= 1 gets
// the row sums of tiles (R, 1..R-1) and (R, R), the column sums of tiles
// (R+1..31, R) (the upper triangle's part, by symmetry) and the row sum of
// columns 0..7; each tile writes its sums into the slot of its tile column
// (row sums) or tile row (column sums), so every slot has one writer.  Rows
// 0..7 (steps 0..6) take the column sums of the LDS block instead.  Row
// thread r (waves 4..7) forms p_r = tau (A v)_r and its share of v^T p; the
// tiles then apply A -= v w^T + w v^T with w = p - (tau/2)(v^T p) v.  Column
// i+1 as it stands before the update is published by its tile column (or
// the LDS block), and every wave forms its updated value itself.  Two
// barriers per step.  Steps 0..127 run here; the trailing 128 x 128 block
// then goes to the NP = 128 kernels (fewer registers, two to four zones per
// CU through the late, latency-bound steps; launch_tridiag256_tail).
// Output as k_tridiag: reflector i in row i of qmat (tau_i in entry i), T
// reversed (d'_k = d_{255-k}).
// ---------------------------------------------------------------------------
template <int NP>
struct TriLowerLds {
  static constexpr int PR = (tri_rx(NP - 1) + 2) & ~1;  // padded length of a row vector
  static constexpr int NW = NP == 256 ? 8 : (NP == 128 ? 2 : 1);  // waves of the workgroup
  static constexpr int NRW = NP / 64;                  // row waves
  double colw[NW][PR];     // each wave's copy of v
  double pub[2][PR];       // column i+1 before step i's update (double-buffered)
  double pw[2][PR];        // p (double-buffered)
  double part[NP / 8][PR]; // slot s of row r: its partial sum of A v
  double l0[NP][9];        // columns 0..7 of every row (steps 0..6)
  double hp[NRW][8];       // the row waves' shares of (A v)_0..7 (steps 0..6)
  double red[2][NRW];      // the row waves' shares of v^T p (double-buffered)
};

// FULL = false: steps 0 .. NP/2 - 1, then the trailing block is handed on
// (below); FULL = true: every step, and the trailing 2 x 2 block gives the
// last two d and the last e (e_{NP-1} is left to the caller).  zstride, ldm
// and ldv: the distance between the zones' matrices and the leading
// dimensions of M and of the reflector rows (NP for a zone's own block; the
// N = 128 trailing block, reduced by k_tridiag_lower<64, true>, sits in rows
// 64.. of the zone's NP = 128 block).
template <int NP, bool FULL = false>
__global__ __launch_bounds__(NP == 256 ? 512 : (NP == 128 ? 128 : 64)) void k_tridiag_lower(
    int64_t z0, const int32_t* __restrict__ zorder, const int64_t* __restrict__ sel_ptr,
    const double* __restrict__ mat, double* __restrict__ qmat, double* __restrict__ dvec,
    double* __restrict__ evec, int nzc, int64_t zstride, int ldm, int ldv) {
  using L = TriLowerLds<NP>;
  constexpr int NH = NP / 64, PR = L::PR, H = NP / 2, NB = NP / 8;
  constexpr int NT = (NB - 1) * NB / 2;  // register tiles (tile columns 1..NB-1)
  constexpr int BS = 64 * L::NW;         // threads
  __shared__ L S;
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;
  if (sel_ptr[z + 1] == sel_ptr[z]) return;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const bool own = t < NT;
  int tr = NB - 1, tc = NB - 1;
  if (own) {  // ring s = NB - 1 - tc holds the s + 1 tiles of tile column tc
    int s = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while (s * (s + 1) / 2 > t) --s;
    while ((s + 1) * (s + 2) / 2 <= t) ++s;
    tc = NB - 1 - s;
    tr = tc + (t - s * (s + 1) / 2);
  }
  const int r0 = 8 * tr, c0 = 8 * tc;
  const int R0 = tri_rx(r0), C0 = tri_rx(c0);  // tri_rx(r0 + ii) = R0 + ii
  const double* M = mat + (size_t)zl * zstride;
  double* V = qmat + (size_t)zl * zstride;
  double a[8][8];
#pragma unroll
  for (int ii = 0; ii < 8; ++ii)
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) a[ii][jj] = own ? M[(size_t)(r0 + ii) * ldm + c0 + jj] : 0.0;
  const int row = t - (BS - NP);  // the last NP threads: row threads
  if (row >= 0) {
#pragma unroll
    for (int c = 0; c < 8; ++c) S.l0[row][c] = M[(size_t)row * ldm + c];
  }
  for (int e = t; e < NB * PR; e += BS) (&S.part[0][0])[e] = 0.0;
  double x[NH];
#pragma unroll
  for (int h = 0; h < NH; ++h) x[h] = M[(size_t)(lane + 64 * h) * ldm];  // column 0 (M symmetric)
  __syncthreads();
  auto pick = [&](int h) {
    double v = x[0];
#pragma unroll
    for (int k = 1; k < NH; ++k) v = h == k ? x[k] : v;
    return v;
  };
  double* col = S.colw[wave];
  for (int i = 0; i < (FULL ? NP - 2 : H); ++i) {
    const int b = i & 1;
    double* pw = S.pw[b];
    double* pub = S.pub[b];
    // reflector H = I - tau v v^T with H x = beta e_1, x = A[i+1.., i]
    double tau = 0.0, scal = 0.0, beta;
    double v[NH];
    {
      double s = 0.0;
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        const int r = lane + 64 * h;
        if (r >= i + 2) s = fma(x[h], x[h], s);
      }
      s = wave_sum_d(s);
      const int ia = i + 1;
      const double alpha = readlane_d(pick(ia >> 6), ia & 63);
      const double di = readlane_d(pick(i >> 6), i & 63);
      beta = alpha;
      if (s > 0.0) {
        // beta = -sign(alpha) nrm, tau = (beta - alpha) / beta = 1 + |alpha| / nrm,
        // 1 / (alpha - beta) = sign(alpha) / (|alpha| + nrm)
        const double hh = fma(alpha, alpha, s);
        const double aa = fabs(alpha);
        if (hh > 0x1p-900 && hh < 0x1p+900) {
          const double rn = rsqrt_acc(hh);
          const double nrm = hh * rn;
          beta = alpha >= 0.0 ? -nrm : nrm;
          tau = fma(aa, rn, 1.0);
          const double rd = rcp_acc(aa + nrm);
          scal = alpha >= 0.0 ? rd : -rd;
        } else {
          const double nrm = sqrt(hh);
          beta = alpha >= 0.0 ? -nrm : nrm;
          tau = (beta - alpha) / beta;
          scal = 1.0 / (alpha - beta);
        }
      }
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        const int r = lane + 64 * h;
        v[h] = r > i + 1 ? x[h] * scal : (r == i + 1 ? 1.0 : 0.0);
        col[tri_rx(r)] = v[h];
      }
      __builtin_amdgcn_wave_barrier();
      if (wave == L::NW - 1) {
#pragma unroll
        for (int h = 0; h < NH; ++h) {
          const int r = lane + 64 * h;
          V[(size_t)i * ldv + r] = r == i ? tau : v[h];
        }
        if (lane == 0) {
          dvec[(size_t)(NP - 1 - i) * nzc + zl] = di;
          evec[(size_t)(NP - 2 - i) * nzc + zl] = beta;
        }
      }
    }
    // partial sums of the tiles (zeros from a tile outside the trailing block)
    const bool act = own && (c0 + 7 > i);
    double vR[8], vC[8];
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) vC[jj] = col[C0 + jj];
#pragma unroll
    for (int ii = 0; ii < 8; ++ii) vR[ii] = col[R0 + ii];
    if (own) {
#pragma unroll
      for (int ii = 0; ii < 8; ++ii) {
        double ps = 0.0;
        if (act) {
#pragma unroll
          for (int jj = 0; jj < 8; ++jj) ps = fma(a[ii][jj], vC[jj], ps);
        }
        S.part[tc][R0 + ii] = ps;
      }
      if (tr != tc) {
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          double cs = 0.0;
          if (act) {
#pragma unroll
            for (int ii = 0; ii < 8; ++ii) cs = fma(a[ii][jj], vR[ii], cs);
          }
          S.part[tr][C0 + jj] = cs;
        }
      }
      // column i+1 before this step's update, from its tile column
      if (tc == ((i + 1) >> 3)) {
        const int j0 = (i + 1) & 7;
#pragma unroll
        for (int jj = 0; jj < 8; ++jj)
          if (jj == j0) {
#pragma unroll
            for (int ii = 0; ii < 8; ++ii) pub[R0 + ii] = a[ii][jj];
          }
      }
    }
    // columns 0..7 (steps 0..6): the row sums of the LDS block go to slot 0
    // of rows 8..255; its column sums are (A v)_0..7
    if (row >= 0) {
      if (i < 7) {
        const double vr = col[tri_rx(row)];
        double rs = 0.0, hs[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const double lv = S.l0[row][c];
          rs = fma(lv, col[c], rs);  // tri_rx(c) = c for c < 8
          hs[c] = lv * vr;
        }
        if (row >= 8) S.part[0][tri_rx(row)] = rs;
        if (i + 1 < 8) pub[tri_rx(row)] = S.l0[row][i + 1];
#pragma unroll
        for (int c = 0; c < 8; ++c) hs[c] = wave_sum_d(hs[c]);
        if (lane == 0) {
#pragma unroll
          for (int c = 0; c < 8; ++c) S.hp[row >> 6][c] = hs[c];
        }
      } else if (i == 7) {
        S.part[0][tri_rx(row)] = 0.0;  // the block has left the trailing matrix
      }
    }
    lds_barrier();
    // p = tau A v and the row waves' shares of v^T p
    if (row >= 0) {
      const int rowx = tri_rx(row);
      double s4[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) s4[q] = S.part[q][rowx];
#pragma unroll
      for (int q = 4; q < NB; ++q) s4[q & 3] += S.part[q][rowx];
      double s = (s4[0] + s4[1]) + (s4[2] + s4[3]);
      if (i < 7 && row < 8) {
        if constexpr (NP == 256)
          s = (S.hp[0][row] + S.hp[1][row]) + (S.hp[2][row] + S.hp[3][row]);
        else if constexpr (NP == 128)
          s = S.hp[0][row] + S.hp[1][row];
        else
          s = S.hp[0][row];
      }
      const double pr = row > i ? tau * s : 0.0;
      pw[rowx] = pr;
      double vp = col[rowx] * pr;
      vp = wave_sum_d(vp);
      if (lane == 0) S.red[b][row >> 6] = vp;
    }
    lds_barrier();
    // A -= v w^T + w v^T, w = p + K v, K = -tau/2 v^T p
    double vpt;
    if constexpr (NP == 256) {
      vpt = (S.red[b][0] + S.red[b][1]) + (S.red[b][2] + S.red[b][3]);
    } else if constexpr (NP == 128) {
      vpt = S.red[b][0] + S.red[b][1];
    } else {
      vpt = S.red[b][0];
    }
    const double K = -0.5 * tau * vpt;
    if (act) {
      double wC[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) wC[jj] = fma(K, vC[jj], pw[C0 + jj]);
#pragma unroll
      for (int ii = 0; ii < 8; ++ii) {
        const double wR = fma(K, vR[ii], pw[R0 + ii]);
#pragma unroll
        for (int jj = 0; jj < 8; ++jj)
          a[ii][jj] = fma(-vR[ii], wC[jj], fma(-wR, vC[jj], a[ii][jj]));
      }
    }
    if (i < 7 && row >= 0) {
      const int rowx = tri_rx(row);
      const double vr = col[rowx], wr = fma(K, vr, pw[rowx]);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const double vc = col[c], wc = fma(K, vc, pw[c]);
        S.l0[row][c] = fma(-vr, wc, fma(-wr, vc, S.l0[row][c]));
      }
    }
    // column i+1 after the update (v_{i+1} = 1, w_{i+1} = p_{i+1} + K)
    {
      const double w1 = fma(K, 1.0, pw[tri_rx(i + 1)]);
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        const int r = lane + 64 * h;
        const double w = fma(K, v[h], pw[tri_rx(r)]);
        const double nc = fma(-v[h], w1, fma(-w, 1.0, pub[tri_rx(r)]));
        x[h] = r > i ? nc : 0.0;
      }
    }
  }
  // the trailing block [H, NP)^2 after step H - 1 (the tiles of tile rows and
  // columns >= H / 8; both triangles) to the rows from H of the zone's
  // reflector block: NP = 256, as a dense H x H matrix that the NP = 128
  // kernels reduce (launch_tridiag256_tail); NP = 128, as row r, column
  // c - H of rows H.. (the layout k_tridiag<128, 1> leaves for
  // k_tridiag<128, 2>)
  if constexpr (FULL) {
    if (t == 0) {  // tile (NB-1, NB-1): the trailing 2 x 2 block, rows NP-2, NP-1
      dvec[(size_t)1 * nzc + zl] = a[6][6];
      evec[(size_t)0 * nzc + zl] = a[7][6];
      dvec[(size_t)0 * nzc + zl] = a[7][7];
    }
    return;
  }
  constexpr int LDO = NP == 256 ? H : NP;
  if (own && tc >= H / 8) {
    double* blk = V + (size_t)H * NP;
#pragma unroll
    for (int ii = 0; ii < 8; ++ii)
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        blk[(size_t)(r0 - H + ii) * LDO + (c0 - H + jj)] = a[ii][jj];
        if (tr != tc) blk[(size_t)(c0 - H + jj) * LDO + (r0 - H + ii)] = a[ii][jj];
      }
  }
  if (NP == 256 && t == 0) evec[(size_t)(NP - 1) * nzc + zl] = 0.0;
}

// e_{NP-1} = 0 of every zone of a chunk (NP = 128)
__global__ void k_fill_elast128(double* __restrict__ evec, int nzc) {
  const int zl = blockIdx.x * 256 + threadIdx.x;
  if (zl < nzc) evec[(size_t)127 * nzc + zl] = 0.0;
}

// NP = 128, tri_split 3: steps 0..63 on the lower-triangle tiles (two waves,
// four zones per CU), then k_tridiag<128, 2> from the trailing block
void launch_tridiag_lower128(const ChunkArgs& A, hipStream_t st) {
  hipLaunchKernelGGL((k_tridiag_lower<128>), dim3(A.nzc), dim3(128), 0, st, A.z0, A.zorder,
                     A.sel_ptr, A.mat, A.ymat, A.dvec, A.evec, A.nzc, (int64_t)128 * 128, 128, 128);
}

// tri_split 4: the trailing 64 x 64 block (rows 64.., columns 0..63 of the
// zone's reflector block) reduced in full by k_tridiag_lower<64> (one wave
// per zone, eight zones per CU): its reflectors go to columns 64..127 of the
// same rows (their columns 0..63 keep the block's values, which k_formq128
// does not read: entries below a reflector's own index are zero by
// definition), its d and e to entries 0..63 and 0..62 of the reversed
// tridiagonal, as steps 64..125 of the whole reduction would write them;
// e_127 = 0 is written here
void launch_tridiag_lower128_tail(const ChunkArgs& A, hipStream_t st) {
  hipLaunchKernelGGL((k_tridiag_lower<64, true>), dim3(A.nzc), dim3(64), 0, st, A.z0, A.zorder,
                     A.sel_ptr, A.ymat + 64 * 128, A.ymat + 64 * 128 + 64, A.dvec, A.evec, A.nzc,
                     (int64_t)128 * 128, 128, 128);
  hipLaunchKernelGGL(k_fill_elast128, dim3((A.nzc + 255) / 256), dim3(256), 0, st, A.evec, A.nzc);
}


// ---------------------------------------------------------------------------
// Q = H_0 H_1 ... H_253 from the reflectors in the rows of qmat (row k: v_k,
// tau_k in entry k), accumulated backwards in blocks of 16 in compact WY
// form (LAPACK dlarft / dlarfb): H_k0 ... H_k0+15 = I - Vb Tb Vb^T and
// Q <- Q - Vb (Tb (Vb^T Q)), on v_mfma_f64_16x16x4f64.  Workgroup y of a
// zone forms columns 64 y .. 64 y + 63 of Q (they depend on no other
// columns): wave w holds tile column J = 4 y + w in its accumulators, 16
// tiles of 16 x 16; block b touches tile rows and columns >= b only.  Q goes
// to umat with its columns reversed, as the QL kernels expect.
// ---------------------------------------------------------------------------
struct FormQ256Lds {
  static constexpr int LV = 17;  // padded row of Vb and Tb
  double vb[2][256 * LV];        // Vb[r][j]: row r of reflector k0 + j
  double tb[2][16 * LV];         // Tb
  double g[16 * LV];             // Vb^T Vb
  double tau[2][16];
};

__global__ __launch_bounds__(256) void k_formq256(int64_t z0, const int32_t* __restrict__ zorder,
                                                  const int64_t* __restrict__ sel_ptr,
                                                  const double* __restrict__ qmat,
                                                  double* __restrict__ umat) {
  constexpr int NP = NPW, NB = 16, LV = FormQ256Lds::LV;
  __shared__ FormQ256Lds S;
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;
  if (sel_ptr[z + 1] == sel_ptr[z]) return;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int lr = lane >> 4, lc = lane & 15;
  const double* V = qmat + (size_t)zl * NP * NP;
  double* U = umat + (size_t)zl * NP * NP;
  const int J = 4 * (int)blockIdx.y + wave;  // the wave's tile column
  d4 q[16];
#pragma unroll
  for (int I = 0; I < 16; ++I)
#pragma unroll
    for (int r = 0; r < 4; ++r) q[I][r] = (I == J && lr + 4 * r == lc) ? 1.0 : 0.0;
  // thread t moves reflector pj = t / 16 of a block, rows pr + 16 e (e < 16)
  const int pj = t >> 4, pr = t & 15;
  double pf[16];
  // reflectors 0..127 are rows of the zone's block; 128..253 (from the NP =
  // 128 kernels) are rows 0..125 of the 128 x 128 block at row 192, for
  // rows 128..255 of Q
  const double* V2 = V + (size_t)192 * NP;
  auto load_block = [&](int b) {
    const int k = 16 * b + pj;
    if (k < NP / 2) {
      const double* src = V + (size_t)k * NP + pr;
#pragma unroll
      for (int e = 0; e < 16; ++e) pf[e] = src[16 * e];
    } else if (k < NP - 2) {
      const double* src = V2 + (size_t)(k - NP / 2) * (NP / 2) + pr;
#pragma unroll
      for (int e = 0; e < 16; ++e) pf[e] = e < 8 ? 0.0 : src[16 * (e - 8)];
    } else {
#pragma unroll
      for (int e = 0; e < 16; ++e) pf[e] = 0.0;
    }
  };
  load_block(NB - 1);
  for (int b = NB - 1; b >= 0; --b) {
    const int buf = b & 1;
    double* vb = S.vb[buf];
    double* tb = S.tb[buf];
    {
      const int k = 16 * b + pj;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int r = pr + 16 * e;
        vb[r * LV + pj] = r == k ? 0.0 : pf[e];
        if (r == k) S.tau[buf][pj] = pf[e];
      }
    }
    if (b > 0) load_block(b - 1);
    lds_barrier();
    // wave 0: G = Vb^T Vb (rows >= 16 b), then Tb column by column:
    // Tb[j][j] = tau_j, Tb[0:j, j] = -tau_j Tb[0:j, 0:j] G[0:j, j]
    if (wave == 0) {
      d4 g = {0.0, 0.0, 0.0, 0.0};
      for (int kk = 4 * b; kk < 64; ++kk) {
        const double xv = vb[(4 * kk + lr) * LV + lc];
        g = mfma_f64(xv, xv, g);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) S.g[(lr + 4 * r) * LV + lc] = g[r];
      __builtin_amdgcn_wave_barrier();
      if (lane < 16) {
        double trr[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          asm volatile("" ::: "memory");  // G column by column, not all loads at once
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < j; ++k) s = fma(trr[k], S.g[k * LV + j], s);
          const double tj = S.tau[buf][j];
          trr[j] = lane < j ? -tj * s : (lane == j ? tj : 0.0);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) tb[lane * LV + j] = trr[j];
      }
    }
    // W = Vb^T Q[:, J]
    d4 w = {0.0, 0.0, 0.0, 0.0};
    if (J >= b) {
#pragma unroll
      for (int I = 0; I < 16; ++I)
        if (I >= b) {
#pragma unroll
          for (int kk = 0; kk < 4; ++kk)
            w = mfma_f64(vb[(16 * I + 4 * kk + lr) * LV + lc], q[I][kk], w);
        }
    }
    lds_barrier();  // Tb
    // W' = -Tb W;  Q[I][J] += Vb[I rows] W'
    if (J >= b) {
      d4 wp = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) wp = mfma_f64(-tb[lc * LV + 4 * kk + lr], w[kk], wp);
#pragma unroll
      for (int I = 0; I < 16; ++I)
        if (I >= b) {
#pragma unroll
          for (int kk = 0; kk < 4; ++kk)
            q[I] = mfma_f64(vb[(16 * I + lc) * LV + 4 * kk + lr], wp[kk], q[I]);
        }
    }
  }
#pragma unroll
  for (int I = 0; I < 16; ++I)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      U[(size_t)(16 * I + lr + 4 * r) * NP + (NP - 1 - (16 * J + lc))] = q[I][r];
}

// ---------------------------------------------------------------------------
// Q = H_0 ... H_61 for NP = 64 (the reflectors k_tridiag_lower<64, true>
// leaves in the rows of qmat), the scheme of k_formq256 with one workgroup
// per zone: wave w holds tile column w of Q (4 tiles), blocks of 16
// reflectors in compact WY form on MFMA.  Every reflector is read before Q
// is written over them (in place, columns reversed as the QL kernels expect).
// ---------------------------------------------------------------------------
struct FormQ64Lds {
  static constexpr int LV = 17;
  double vb[2][64 * LV];
  double tb[2][16 * LV];
  double g[16 * LV];
  double tau[2][16];
};

__global__ __launch_bounds__(256) void k_formq64(int64_t z0, const int32_t* __restrict__ zorder,
                                                 const int64_t* __restrict__ sel_ptr,
                                                 double* __restrict__ qmat) {
  constexpr int NP = 64, NB = 4, LV = FormQ64Lds::LV;
  __shared__ FormQ64Lds S;
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;
  if (sel_ptr[z + 1] == sel_ptr[z]) return;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int lr = lane >> 4, lc = lane & 15;
  double* V = qmat + (size_t)zl * NP * NP;
  const int J = wave;
  d4 q[NB];
#pragma unroll
  for (int I = 0; I < NB; ++I)
#pragma unroll
    for (int r = 0; r < 4; ++r) q[I][r] = (I == J && lr + 4 * r == lc) ? 1.0 : 0.0;
  // thread t moves reflector pj = t / 16 of a block, rows pr + 16 e (e < 4)
  const int pj = t >> 4, pr = t & 15;
  double pf[4];
  auto load_block = [&](int b) {
    const int k = 16 * b + pj;
    if (k < NP - 2) {
      const double* src = V + (size_t)k * NP + pr;
#pragma unroll
      for (int e = 0; e < 4; ++e) pf[e] = src[16 * e];
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) pf[e] = 0.0;
    }
  };
  load_block(NB - 1);
  for (int b = NB - 1; b >= 0; --b) {
    const int buf = b & 1;
    double* vb = S.vb[buf];
    double* tb = S.tb[buf];
    {
      const int k = 16 * b + pj;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = pr + 16 * e;
        vb[r * LV + pj] = r <= k ? 0.0 : pf[e];  // v_k is zero above row k + 1
        if (r == k) S.tau[buf][pj] = pf[e];
      }
    }
    if (b > 0) load_block(b - 1);
    lds_barrier();
    // wave 0: G = Vb^T Vb (rows >= 16 b), then Tb column by column:
    // Tb[j][j] = tau_j, Tb[0:j, j] = -tau_j Tb[0:j, 0:j] G[0:j, j]
    if (wave == 0) {
      d4 g = {0.0, 0.0, 0.0, 0.0};
      for (int kk = 4 * b; kk < 16; ++kk) {
        const double xv = vb[(4 * kk + lr) * LV + lc];
        g = mfma_f64(xv, xv, g);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) S.g[(lr + 4 * r) * LV + lc] = g[r];
      __builtin_amdgcn_wave_barrier();
      if (lane < 16) {
        double trr[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          asm volatile("" ::: "memory");  // G column by column, not all loads at once
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < j; ++k) s = fma(trr[k], S.g[k * LV + j], s);
          const double tj = S.tau[buf][j];
          trr[j] = lane < j ? -tj * s : (lane == j ? tj : 0.0);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) tb[lane * LV + j] = trr[j];
      }
    }
    // W = Vb^T Q[:, J]
    d4 w = {0.0, 0.0, 0.0, 0.0};
    if (J >= b) {
#pragma unroll
      for (int I = 0; I < NB; ++I)
        if (I >= b) {
#pragma unroll
          for (int kk = 0; kk < 4; ++kk)
            w = mfma_f64(vb[(16 * I + 4 * kk + lr) * LV + lc], q[I][kk], w);
        }
    }
    lds_barrier();  // Tb
    // W' = -Tb W;  Q[I][J] += Vb[I rows] W'
    if (J >= b) {
      d4 wp = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) wp = mfma_f64(-tb[lc * LV + 4 * kk + lr], w[kk], wp);
#pragma unroll
      for (int I = 0; I < NB; ++I)
        if (I >= b) {
#pragma unroll
          for (int kk = 0; kk < 4; ++kk)
            q[I] = mfma_f64(vb[(16 * I + lc) * LV + 4 * kk + lr], wp[kk], q[I]);
        }
    }
  }
  // every reflector row was read before the last barrier
#pragma unroll
  for (int I = 0; I < NB; ++I)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      V[(size_t)(16 * I + lr + 4 * r) * NP + (NP - 1 - (16 * J + lc))] = q[I][r];
}

// NP = 64 (33 <= N <= 64, and the observation-space zones of N = 128): the
// reduction on the lower triangle in full (one wave per zone), e_63 = 0,
// then Q by k_formq64
__global__ void k_fill_elast64(double* __restrict__ evec, int nzc) {
  const int zl = blockIdx.x * 256 + threadIdx.x;
  if (zl < nzc) evec[(size_t)63 * nzc + zl] = 0.0;
}

void launch_tridiag_lower64(const ChunkArgs& A, hipStream_t st) {
  hipLaunchKernelGGL((k_tridiag_lower<64, true>), dim3(A.nzc), dim3(64), 0, st, A.z0, A.zorder,
                     A.sel_ptr, A.mat, A.ymat, A.dvec, A.evec, A.nzc, (int64_t)64 * 64, 64, 64);
  hipLaunchKernelGGL(k_fill_elast64, dim3((A.nzc + 255) / 256), dim3(256), 0, st, A.evec, A.nzc);
  hipLaunchKernelGGL(k_formq64, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr, A.ymat);
}

// ---------------------------------------------------------------------------
// T = (Y Y^T) Omega (the operator of k_tmat: Omega = RotateVector(w, v) as
// the identity plus a rank-structured term), rows 64 y .. 64 y + 63 in
// workgroup y: wave w forms tile row I = 4 y + w of Y Y^T (16 tiles in its
// accumulators, Y staged through LDS 16 columns at a time, the next block
// loaded while the current one is multiplied), then its rows' products with
// the vectors of Omega, then T.  Zones without observations
// get T = I.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tmat256(int64_t z0, const int32_t* __restrict__ zorder,
                                                 int N, const int64_t* __restrict__ sel_ptr,
                                                 const double* __restrict__ umat,
                                                 double* __restrict__ mat,
                                                 const double* __restrict__ vec) {
  constexpr int NP = NPW, KC = 16, LY = KC + 1;
  __shared__ double Yc[2][NP * LY];
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;
  double* T = mat + (size_t)zl * NP * NP;
  const int rb = blockIdx.y;
  if (sel_ptr[z + 1] == sel_ptr[z]) {
    for (int e = threadIdx.x; e < 64 * NP; e += 256) {
      const int r = 64 * rb + e / NP, c = e % NP;
      T[(size_t)r * NP + c] = r == c ? 1.0 : 0.0;
    }
    return;
  }
  const double* Yz = umat + (size_t)zl * NP * NP;
  const double* v = vec + (size_t)zl * 4 * NP + 2 * NP;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, kk = lane >> 4;
  const int I = 4 * rb + wave;
  d4 acc[16];
#pragma unroll
  for (int J = 0; J < 16; ++J) acc[J] = d4{0, 0, 0, 0};
  // columns k0 .. k0 + 15 of Y in LDS buffer (k0 / 16) % 2; thread e holds
  // 16 values of the next block (rows e / 16 + 16 u, column e % 16) in
  // registers while the current block is multiplied
  double yn[16];
  auto load_y = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 16; ++u) yn[u] = Yz[(size_t)(threadIdx.x / KC + 16 * u) * NP + k0 + threadIdx.x % KC];
  };
  auto store_y = [&](int b) {
#pragma unroll
    for (int u = 0; u < 16; ++u) Yc[b][(threadIdx.x / KC + 16 * u) * LY + threadIdx.x % KC] = yn[u];
  };
  load_y(0);
  store_y(0);
  __syncthreads();
  for (int k0 = 0; k0 < NP; k0 += KC) {
    const int b = (k0 / KC) & 1;
    if (k0 + KC < NP) load_y(k0 + KC);
    asm volatile("" ::: "memory");  // compiler barrier: keeps the loads above here
    const double* Yb = Yc[b];
#pragma unroll
    for (int k4 = 0; k4 < KC; k4 += 4) {
      const double av = Yb[(16 * I + li) * LY + k4 + kk];
#pragma unroll
      for (int J = 0; J < 16; ++J) acc[J] = mfma_f64(av, Yb[(16 * J + li) * LY + k4 + kk], acc[J]);
    }
    if (k0 + KC < NP) store_y(b ^ 1);
    __syncthreads();
  }
  // register q of lane l: row 16 I + l/16 + 4 q, column 16 J + l%16
  const double wN = 1.0 / sqrt((double)N);
  const double vN = v[N - 1];
  const double sgv = vN >= 0 ? 1.0 : -1.0;
  const double alv = -1.0 / (fabs(vN) + 1.0);
  const double alw = -1.0 / (wN + 1.0);
  double vw = 0;  // v~ . w~
  for (int s = 0; s < N - 1; ++s) vw = fma(v[s], wN, vw);
  double vl[16];
#pragma unroll
  for (int J = 0; J < 16; ++J) {
    const int c = 16 * J + li;
    vl[J] = c < N ? v[c] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    // u0 = T0[r][N-1], u1 = T0 v, u2 = T0 w~, u3 = T0 z_v
    double u0 = 0.0, u1 = 0.0, u2 = 0.0, u3 = 0.0;
#pragma unroll
    for (int J = 0; J < 16; ++J) {
      const int c = 16 * J + li;
      const double tt = acc[J][q];
      u0 += c == N - 1 ? tt : 0.0;
      u1 = fma(tt, vl[J], u1);
      u2 = fma(tt, c < N - 1 ? wN : 0.0, u2);
      u3 = fma(tt, vl[J] + (c == N - 1 ? sgv : 0.0), u3);
    }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      u0 += shfl_xor_d(u0, o);
      u1 += shfl_xor_d(u1, o);
      u2 += shfl_xor_d(u2, o);
      u3 += shfl_xor_d(u3, o);
    }
    const int r = 16 * I + kk + 4 * q;
#pragma unroll
    for (int J = 0; J < 16; ++J) {
      const int s = 16 * J + li;
      double tt = acc[J][q];
      if (s < N) {
        const double ws = wN;
        const double zw = wN + (s == N - 1 ? 1.0 : 0.0);
        const double vt = (s < N - 1) ? vl[J] : 0.0;
        if (s == N - 1) tt -= u0;
        tt += u1 * ws + alw * u2 * zw + alv * u3 * vt + alv * alw * vw * u3 * zw;
      }
      T[(size_t)r * NP + s] = tt;
    }
  }
}

// ---------------------------------------------------------------------------
// Update: out = S_z [T | a] for the zone's rows, transposed on the MFMA as in
// k_update (A = [T|a]^T from LDS, B = S_z^T from global).  Workgroup g:
// zone 8 (g / 64) + g % 8 of the chunk, columns 32 cb .. 32 cb + 31 of T with
// cb = (g / 8) % 8, so the eight workgroups of a zone land on one XCD (the
// dispatcher deals workgroups to the XCDs in turn) and share its L2 for the
// zone's rows of S; 68 KB of LDS each, two per CU.  The members of the next
// 8 k-steps are loaded while those of the current 8 are multiplied.  Each workgroup forms
// S a itself (bit-identical in all eight), for the mean and the clamp.
//   MODE 0: xa = xf + Sf a, Sa = Sf T            (rrsqrt.F90:151,185,463)
//   MODE 1: Sf = (E - xbar)/sqrt(N-1) on the fly; xa = xbar + clamp(Sf a);
//           Ea = xa + sqrt(N-1) * (inflation * Sf T)  (assimilation.F90:3301-3320)
// TE is the type of S and out (double, or float on the mixed path).
// ---------------------------------------------------------------------------
template <int MODE, typename TE>
__global__ __launch_bounds__(256) void k_update256(int nzc, int64_t z0,
                                                   const int32_t* __restrict__ zorder, int N,
                                                   const int64_t* __restrict__ zone_ptr,
                                                   const double* __restrict__ mat,
                                                   const double* __restrict__ vec,
                                                   const double* __restrict__ X0,
                                                   const TE* __restrict__ S, int64_t ldS,
                                                   double* __restrict__ xa, TE* __restrict__ Out,
                                                   int64_t ldO, double scaling, double infl,
                                                   double maxcorr,
                                                   const double* __restrict__ maxcorr_v) {
#pragma clang fp contract(off)
  constexpr int NP = NPW, CB = 32, NCB = NP / CB, NBc = CB / 16, LW = CB + 1;
  __shared__ double Ws[NP * LW];
  __shared__ double as[NP];
  const int g = blockIdx.x;
  const int zl = (g / (8 * NCB)) * 8 + (g & 7), cb = (g >> 3) % NCB;
  if (zl >= nzc) return;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;
  const int64_t i0 = zone_ptr[z];
  const int64_t sz = zone_ptr[z + 1] - i0;
  if (sz <= 0) return;
  const double* T = mat + (size_t)zl * NP * NP;
  const double* a = vec + (size_t)zl * 4 * NP + NP;
  copy_batched<16>(threadIdx.x, NP * CB, 256,
                   [&](int e) { return T[(size_t)(e / CB) * NP + CB * cb + e % CB]; },
                   [&](int e, double t) { Ws[(e / CB) * LW + e % CB] = t; });
  as[threadIdx.x] = a[threadIdx.x];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, kk = lane >> 4;
  for (int64_t rb = 16 * wave; rb < sz; rb += 64) {
    d4 acc[NBc];
#pragma unroll
    for (int J = 0; J < NBc; ++J) acc[J] = d4{0, 0, 0, 0};
    d4 acca = {0, 0, 0, 0};
    const int64_t row = rb + li;
    const bool rok = row < sz;
    const double xb = (MODE == 1 && rok) ? X0[i0 + row] : 0.0;
    // the members of 8 k-steps at a time, the next 8 loaded while the
    // current ones are multiplied (a load per step, waited for at once,
    // would expose the memory latency 64 times)
    auto load8 = [&](int k0, TE* raw) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int kc = k0 + 4 * u + kk;
        const bool ok = rok && kc < N;
        raw[u] = S[ok ? (i0 + row) + (int64_t)kc * ldS : i0];
      }
    };
    TE cur[8];
    load8(0, cur);
    for (int k0 = 0; k0 < NP; k0 += 32) {
      TE nxt[8];
      load8(k0 + 32 < NP ? k0 + 32 : k0, nxt);
      asm volatile("" ::: "memory");  // compiler barrier: keeps the loads above here
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int kc = k0 + 4 * u + kk;
        const bool ok = rok && kc < N;
        const double sv = (double)cur[u];
        const double bval = ok ? ((MODE == 1) ? (sv - xb) / scaling : sv) : 0.0;
#pragma unroll
        for (int J = 0; J < NBc; ++J) acc[J] = mfma_f64(Ws[kc * LW + 16 * J + li], bval, acc[J]);
        acca = mfma_f64(li == 0 ? as[kc] : 0.0, bval, acca);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) cur[u] = nxt[u];
    }
    // S a of row (rb + li): column 0 of the a-tile, lanes 0..15, register 0
    const double dx = shfl_d(acca[0], li);
    double xar = 0.0;
    if (rok) {
      const double x0 = X0[i0 + row];
      xar = x0 + dx;
      if (MODE == 1) {
        const double mc = maxcorr_v ? maxcorr_v[i0 + row] : maxcorr;
        if (xar - mc > x0) xar = x0 + mc;
        if (xar < x0 - mc) xar = x0 - mc;
      }
      if (MODE == 0 && cb == 0 && kk == 0) xa[i0 + row] = xar;
    }
#pragma unroll
    for (int J = 0; J < NBc; ++J) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int c = CB * cb + 16 * J + kk + 4 * r;
        if (rok && c < N) {
          double o = acc[J][r];
          if (MODE == 1) {
            const double sa = (infl != 1.0) ? infl * o : o;
            o = xar + sa * scaling;
          }
          Out[(i0 + row) + (int64_t)c * ldO] = (TE)o;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
void launch_chunk256(const ChunkArgs& A, hipStream_t st, PhaseTimer* tm) {
  if (tm) tm->mark(PH_GRAM);
  if (A.prec == 1)
    hipLaunchKernelGGL((k_gram256<float>), dim3(A.nzc, 4), dim3(256), 0, st, A.z0, A.zorder,
                       A.sel_ptr, A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS32, A.mat, A.vec,
                       A.err);
  else
    hipLaunchKernelGGL((k_gram256<double>), dim3(A.nzc, 4), dim3(256), 0, st, A.z0, A.zorder,
                       A.sel_ptr, A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS, A.mat, A.vec,
                       A.err);
  if (tm) tm->mark(PH_EIG);
  hipLaunchKernelGGL((k_tridiag_lower<256>), dim3(A.nzc), dim3(512), 0, st, A.z0, A.zorder,
                     A.sel_ptr, A.mat, A.ymat, A.dvec, A.evec, A.nzc, (int64_t)256 * 256, 256, 256);
  launch_tridiag256_tail(A, st);
  hipLaunchKernelGGL(k_formq256, dim3(A.nzc, 4), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                     A.ymat, A.umat);
  launch_ql256(A, st);
  if (A.mode == 2) return;  // the likelihood has no kernels at this size
  if (tm) tm->mark(PH_TMAT);
  hipLaunchKernelGGL(k_tmat256, dim3(A.nzc, 4), dim3(256), 0, st, A.z0, A.zorder, A.N, A.sel_ptr,
                     A.umat, A.mat, A.vec);
  if (tm) tm->mark(PH_UPDATE);
  const unsigned nb = (unsigned)(((A.nzc + 7) / 8) * 64);  // 8 column blocks per zone
  if (A.prec == 1)
    hipLaunchKernelGGL((k_update256<1, float>), dim3(nb), dim3(256), 0, st, A.nzc, A.z0, A.zorder,
                       A.N, A.zone_ptr, A.mat, A.vec, A.X0, A.S32, A.ldS, A.xa, A.Out32, A.ldO,
                       A.scaling, A.infl, A.maxcorr, A.maxcorr_v);
  else if (A.mode == 0)
    hipLaunchKernelGGL((k_update256<0, double>), dim3(nb), dim3(256), 0, st, A.nzc, A.z0,
                       A.zorder, A.N, A.zone_ptr, A.mat, A.vec, A.X0, A.S, A.ldS, A.xa, A.Out,
                       A.ldO, A.scaling, A.infl, A.maxcorr, A.maxcorr_v);
  else
    hipLaunchKernelGGL((k_update256<1, double>), dim3(nb), dim3(256), 0, st, A.nzc, A.z0,
                       A.zorder, A.N, A.zone_ptr, A.mat, A.vec, A.X0, A.S, A.ldS, A.xa, A.Out,
                       A.ldO, A.scaling, A.infl, A.maxcorr, A.maxcorr_v);
  if (tm) tm->mark(PH_END);
}

}  // namespace oakg
