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
//   k_tmat128  T = (Y Y^T) Omega with Y Y^T on MFMA (operands from L2)
//   k_update   as for NP <= 64, with [T | a] read from global memory
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
  for (int64_t s = 0; s < nsteps; ++s) {
    const int64_t q = 4 * s + kq;
    double r = 0, rd = 0;
    int64_t l = 0;
    bool ok = q < p;
    if (ok) {
      l = sel_idx[p0 + q];
      if (l < 0 || l >= m) {
        atomicOr(err, OAKG_DERR_IDX);
        ok = false;
        l = 0;
      } else {
        const double w = sel_w[p0 + q];
        r = (w * w) * rho[l];
        rd = r * innov[l];
      }
    }
    double bv[8];
    const TS* row = HS + (size_t)l * P128 + jl;
#pragma unroll
    for (int J = 0; J < 8; ++J) bv[J] = ok ? (double)row[16 * J] : 0.0;
    const double aA = r * bv[IA], aB = r * bv[IB];
#pragma unroll
    for (int J = 0; J <= IA; ++J) accA[J] = mfma_f64(aA, bv[J], accA[J]);
#pragma unroll
    for (int J = 0; J <= IB; ++J) accB[J] = mfma_f64(aB, bv[J], accB[J]);
    if (W == 0) {
#pragma unroll
      for (int J = 0; J < 8; ++J) bacc[J] = fma(rd, bv[J], bacc[J]);
    }
  }
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
__global__ __launch_bounds__(256) void k_gram128(int64_t z0, const int64_t* __restrict__ sel_ptr,
                                                 const int32_t* __restrict__ sel_idx,
                                                 const double* __restrict__ sel_w, int64_t m,
                                                 const double* __restrict__ rho,
                                                 const double* __restrict__ innov,
                                                 const TS* __restrict__ HS,
                                                 double* __restrict__ mat, double* __restrict__ vec,
                                                 unsigned* __restrict__ err) {
  const int zl = blockIdx.x;
  const int64_t z = z0 + zl;
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

// ---------------------------------------------------------------------------
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

__global__ __launch_bounds__(256) void k_eig128(int64_t z0, int nzc, int N,
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
  const int64_t z = z0 + zl;
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
// T = (Y Y^T) Omega, NP = 128 (Omega as in analysis.hip's k_tmat)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tmat128(int64_t z0, int N, const int64_t* __restrict__ sel_ptr,
                                                 const double* __restrict__ ymat,
                                                 double* __restrict__ mat,
                                                 const double* __restrict__ vec) {
  __shared__ double U[4][P128];
  const int zl = blockIdx.x;
  const int64_t z = z0 + zl;
  double* T = mat + (size_t)zl * P128 * P128;
  if (sel_ptr[z + 1] == sel_ptr[z]) {
    for (int e = threadIdx.x; e < P128 * P128; e += 256) T[e] = (e / P128 == e % P128) ? 1.0 : 0.0;
    return;
  }
  const double* Y = ymat + (size_t)zl * P128 * P128;
  const double* v = vec + (size_t)zl * 4 * P128 + 2 * P128;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, kk = lane >> 4;
  // wave w: tile rows I = 2w, 2w+1, all tile columns
  d4 acc[2][8];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int J = 0; J < 8; ++J) acc[a][J] = d4{0, 0, 0, 0};
  for (int j0 = 0; j0 < P128; j0 += 4) {
    double bf[8];
#pragma unroll
    for (int J = 0; J < 8; ++J) bf[J] = Y[(16 * J + li) * P128 + j0 + kk];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const double af = bf[2 * wave + a];
#pragma unroll
      for (int J = 0; J < 8; ++J) acc[a][J] = mfma_f64(af, bf[J], acc[a][J]);
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int J = 0; J < 8; ++J)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        T[(16 * (2 * wave + a) + (lane >> 4) + 4 * r) * P128 + 16 * J + li] = acc[a][J][r];
  __syncthreads();
  const double wN = 1.0 / sqrt((double)N);
  const double vN = v[N - 1];
  const double sgv = vN >= 0 ? 1.0 : -1.0;
  const double alv = -1.0 / (fabs(vN) + 1.0);
  const double alw = -1.0 / (wN + 1.0);
  double vw = 0;
  for (int s = 0; s < N - 1; ++s) vw = fma(v[s], wN, vw);
  if (threadIdx.x < P128) {
    const int r = threadIdx.x;
    double u1 = 0, u2 = 0, u3 = 0;
    for (int q = 0; q < P128; ++q) {
      const double t = T[r * P128 + q];
      const double vq = q < N ? v[q] : 0.0;
      u1 = fma(t, vq, u1);
      u2 = fma(t, q < N - 1 ? wN : 0.0, u2);
      u3 = fma(t, vq + (q == N - 1 ? sgv : 0.0), u3);
    }
    U[0][r] = T[r * P128 + N - 1];
    U[1][r] = u1;
    U[2][r] = u2;
    U[3][r] = u3;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < P128 * P128; e += 256) {
    const int r = e / P128, s = e % P128;
    if (s < N) {
      double t = T[e];
      const double zw = wN + (s == N - 1 ? 1.0 : 0.0);
      const double vt = (s < N - 1) ? v[s] : 0.0;
      if (s == N - 1) t -= U[0][r];
      t += U[1][r] * wN + alw * U[2][r] * zw + alv * U[3][r] * vt + alv * alw * vw * U[3][r] * zw;
      T[e] = t;
    }
  }
}

// ---------------------------------------------------------------------------
// Update for NP = 128: out = Sf_z [T | a], [T | a] read from global memory.
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(256) void k_update128(int64_t z0, int N, const int64_t* __restrict__ zone_ptr,
                                                   const double* __restrict__ mat,
                                                   const double* __restrict__ vec,
                                                   const double* __restrict__ X0,
                                                   const double* __restrict__ S, int64_t ldS,
                                                   double* __restrict__ xa, double* __restrict__ Out,
                                                   int64_t ldO, double scaling, double infl,
                                                   double maxcorr) {
#pragma clang fp contract(off)
  constexpr int NB = 8;
  const int zl = blockIdx.x;
  const int64_t z = z0 + zl;
  const int64_t i0 = zone_ptr[z];
  const int64_t sz = zone_ptr[z + 1] - i0;
  const double* T = mat + (size_t)zl * P128 * P128;
  const double* a = vec + (size_t)zl * 4 * P128 + P128;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, kk = lane >> 4;
  for (int64_t rb = 16 * wave; rb < sz; rb += 64) {
    d4 acc[NB + 1];
#pragma unroll
    for (int J = 0; J <= NB; ++J) acc[J] = d4{0, 0, 0, 0};
    const int64_t row = rb + li;
    const bool rok = row < sz;
    const double xb = (MODE == 1 && rok) ? X0[i0 + row] : 0.0;
    for (int k0 = 0; k0 < P128; k0 += 4) {
      const int kc = k0 + kk;
      double bval = 0.0;
      if (rok && kc < N) {
        const double sv = S[(i0 + row) + (int64_t)kc * ldS];
        bval = (MODE == 1) ? (sv - xb) / scaling : sv;
      }
#pragma unroll
      for (int J = 0; J < NB; ++J) acc[J] = mfma_f64(T[kc * P128 + 16 * J + li], bval, acc[J]);
      acc[NB] = mfma_f64(li == 0 ? a[kc] : 0.0, bval, acc[NB]);
    }
    const double dx = shfl_d(acc[NB][0], li);
    double xar = 0.0;
    if (rok) {
      const double x0 = X0[i0 + row];
      xar = x0 + dx;
      if (MODE == 1) {
        if (xar - maxcorr > x0) xar = x0 + maxcorr;
        if (xar < x0 - maxcorr) xar = x0 - maxcorr;
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

void launch_chunk128(const ChunkArgs& A, hipStream_t st, PhaseTimer* tm) {
  if (tm) tm->mark(PH_GRAM);
  if (A.prec == 1)
    hipLaunchKernelGGL(k_gram128<float>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.sel_ptr,
                       A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS32, A.mat, A.vec, A.err);
  else
    hipLaunchKernelGGL(k_gram128<double>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.sel_ptr,
                       A.sel_idx, A.sel_w, A.m, A.rho, A.innov, A.HS, A.mat, A.vec, A.err);
  if (tm) tm->mark(PH_EIG);
  if (A.eig_method == 1) {
    launch_tql_pipeline(128, A, st);
    hipLaunchKernelGGL(k_eig128, dim3(A.nzc), dim3(256), 0, st, A.z0, A.nzc, A.N, A.sel_ptr,
                       A.mat, A.ymat, A.vec, A.ampl, A.sweeps, A.err, A.tol, A.max_sweeps,
                       (const int32_t*)A.tql_status);
  } else {
    hipLaunchKernelGGL(k_eig128, dim3(A.nzc), dim3(256), 0, st, A.z0, A.nzc, A.N, A.sel_ptr,
                       A.mat, A.ymat, A.vec, A.ampl, A.sweeps, A.err, A.tol, A.max_sweeps,
                       (const int32_t*)nullptr);
  }
  if (A.mode == 2) {  // innovation likelihood: adjoint instead of transform + update
    if (tm) tm->mark(PH_UPDATE);
    launch_nll_adjoint(128, A, st);
    if (tm) tm->mark(PH_END);
    return;
  }
  if (tm) tm->mark(PH_TMAT);
  hipLaunchKernelGGL(k_tmat128, dim3(A.nzc), dim3(256), 0, st, A.z0, A.N, A.sel_ptr, A.ymat, A.mat,
                     A.vec);
  if (tm) tm->mark(PH_UPDATE);
  if (A.prec == 1)
    launch_update_f32(128, A, st);
  else if (A.mode == 0)
    hipLaunchKernelGGL(k_update128<0>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.N, A.zone_ptr, A.mat,
                       A.vec, A.X0, A.S, A.ldS, A.xa, A.Out, A.ldO, A.scaling, A.infl, A.maxcorr);
  else
    hipLaunchKernelGGL(k_update128<1>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.N, A.zone_ptr, A.mat,
                       A.vec, A.X0, A.S, A.ldS, A.xa, A.Out, A.ldO, A.scaling, A.infl, A.maxcorr);
  if (tm) tm->mark(PH_END);
}

}  // namespace oakg
