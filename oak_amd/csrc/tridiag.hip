// FP64 symmetric eigensolver by Householder tridiagonalisation and implicit
// QL iterations (the method of LAPACK's dsyev = dsytrd + dorgtr + dsteqr,
// which OAK's symeig calls, matoper_inc.F90:986-1005), restructured for a
// batch of zones:
//
//   k_tridiag      one workgroup per zone, matrix in LDS: M = Q T Q^T with
//                  Householder reflectors (lower), then Q formed explicitly
//                  by backward accumulation.  Writes Q and (d, e).
//   k_tql          one lane per zone: implicit QL with Wilkinson shifts on
//                  (d, e) (Golub & Van Loan, 4th ed., section 8.3; Bowdler,
//                  Martin, Reinsch & Wilkinson 1968); the sequential
//                  recurrences of many zones run side by side.  Every plane
//                  rotation is logged; d ends up holding the eigenvalues.
//   k_tql_apply    one workgroup per zone: applies the logged rotations to
//                  the rows of Q in LDS, giving the eigenvectors U = Q Z.
//   k_eigvec_epi   per zone: a = U diag(1/mu) U^T b (+ one refinement step
//                  against M), v ~ U diag(mu^1/2) U^T 1, Y = U diag(mu^-1/4),
//                  with mu = max(lambda, 1) (rrsqrt.F90:137-142, 176-177).
#include "analysis.hpp"
#include "common.hpp"

namespace oakg {

__device__ __forceinline__ double wave_sum_d(double s) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) s += shfl_xor_d(s, o);
  return s;
}

// ---------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(256) void k_tridiag(int64_t z0, const int64_t* __restrict__ sel_ptr,
                                                 const double* __restrict__ mat,
                                                 double* __restrict__ qmat,
                                                 double* __restrict__ dvec, double* __restrict__ evec,
                                                 int nzc) {
  constexpr int LD = NP + 1;
  __shared__ double A[NP * LD];
  __shared__ double pv[NP];
  __shared__ double tau[NP];
  __shared__ double sc[4];
  const int zl = blockIdx.x;
  const int64_t z = z0 + zl;
  if (sel_ptr[z + 1] == sel_ptr[z]) return;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const double* M = mat + (size_t)zl * NP * NP;
  for (int e = tid; e < NP * NP; e += 256) A[(e / NP) * LD + e % NP] = M[e];
  __syncthreads();
  // ---- reduction to tridiagonal form, column by column ----
  for (int i = 0; i < NP - 2; ++i) {
    const int L = NP - 1 - i;  // rows/columns i+1 .. NP-1
    if (wave == 0) {
      double s = 0;
      for (int r = i + 2 + lane; r < NP; r += 64) {
        const double x = A[r * LD + i];
        s = fma(x, x, s);
      }
      s = wave_sum_d(s);
      if (lane == 0) {
        const double alpha = A[(i + 1) * LD + i];
        double t = 0.0, beta = alpha, scal = 0.0;
        if (s > 0.0) {
          const double nrm = sqrt(fma(alpha, alpha, s));
          beta = alpha >= 0.0 ? -nrm : nrm;
          t = (beta - alpha) / beta;
          scal = 1.0 / (alpha - beta);
        }
        sc[0] = t;
        sc[1] = beta;
        sc[2] = scal;
      }
    }
    __syncthreads();
    const double t = sc[0], scal = sc[2];
    if (t != 0.0)
      for (int r = i + 2 + tid; r < NP; r += 256) A[r * LD + i] *= scal;
    if (tid == 0) {
      dvec[(size_t)i * nzc + zl] = A[i * LD + i];
      evec[(size_t)i * nzc + zl] = sc[1];
      tau[i] = t;
      A[(i + 1) * LD + i] = 1.0;  // v(i+1) = 1
    }
    __syncthreads();
    if (t == 0.0) continue;
    // p = t A_sub v, two threads per row
    {
      const int rr = tid >> 1, h = tid & 1;
      double sum = 0;
      if (rr < L) {
        const int r = i + 1 + rr;
        const int half = (L + 1) >> 1;
        const int c0 = i + 1 + h * half;
        const int c1 = min(i + 1 + (h + 1) * half, NP);
        for (int c = c0; c < c1; ++c) sum = fma(A[r * LD + c], A[c * LD + i], sum);
      }
      sum += shfl_xor_d(sum, 1);
      if (rr < L && h == 0) pv[i + 1 + rr] = t * sum;
    }
    __syncthreads();
    if (wave == 0) {
      double s = 0;
      for (int r = i + 1 + lane; r < NP; r += 64) s = fma(pv[r], A[r * LD + i], s);
      s = wave_sum_d(s);
      if (lane == 0) sc[3] = -0.5 * t * s;
    }
    __syncthreads();
    const double K = sc[3];
    // A_sub -= v w^T + w v^T, w = p + K v
    for (int e = tid; e < L * L; e += 256) {
      const int r = i + 1 + e / L, c = i + 1 + e % L;
      const double vr = A[r * LD + i], vc = A[c * LD + i];
      const double wr = fma(K, vr, pv[r]), wc = fma(K, vc, pv[c]);
      A[r * LD + c] -= fma(vr, wc, wr * vc);
    }
    __syncthreads();
  }
  if (tid == 0) {
    dvec[(size_t)(NP - 2) * nzc + zl] = A[(NP - 2) * LD + NP - 2];
    evec[(size_t)(NP - 2) * nzc + zl] = A[(NP - 1) * LD + NP - 2];
    dvec[(size_t)(NP - 1) * nzc + zl] = A[(NP - 1) * LD + NP - 1];
    evec[(size_t)(NP - 1) * nzc + zl] = 0.0;
  }
  __syncthreads();
  // ---- Q = H_0 H_1 ... H_{NP-3}, backward accumulation in place ----
  if (tid < 4) A[(NP - 2 + tid / 2) * LD + NP - 2 + tid % 2] = (tid / 2 == tid % 2) ? 1.0 : 0.0;
  __syncthreads();
  for (int i = NP - 3; i >= 0; --i) {
    const int a0 = i + 1;
    // row/column a0 of the active block become the identity's
    for (int r = a0 + 1 + tid; r < NP; r += 256) {
      A[r * LD + a0] = 0.0;
      A[a0 * LD + r] = 0.0;
    }
    if (tid == 0) A[a0 * LD + a0] = 1.0;
    __syncthreads();
    const double ti = tau[i];
    if (ti == 0.0) continue;
    const int L = NP - a0;
    // y_c = sum_r v_r Q[r][c]  (two threads per column)
    {
      const int cc = tid >> 1, h = tid & 1;
      double sum = 0;
      if (cc < L) {
        const int c = a0 + cc;
        const int half = (L + 1) >> 1;
        const int r0 = a0 + h * half;
        const int r1 = min(a0 + (h + 1) * half, NP);
        for (int r = r0; r < r1; ++r) sum = fma(A[r * LD + i], A[r * LD + c], sum);
      }
      sum += shfl_xor_d(sum, 1);
      if (cc < L && h == 0) pv[a0 + cc] = ti * sum;
    }
    __syncthreads();
    for (int e = tid; e < L * L; e += 256) {
      const int r = a0 + e / L, c = a0 + e % L;
      A[r * LD + c] -= A[r * LD + i] * pv[c];
    }
    __syncthreads();
  }
  for (int r = 1 + tid; r < NP; r += 256) {
    A[r * LD] = 0.0;
    A[r] = 0.0;
  }
  if (tid == 0) A[0] = 1.0;
  __syncthreads();
  double* Q = qmat + (size_t)zl * NP * NP;
  for (int e = tid; e < NP * NP; e += 256) Q[e] = A[(e / NP) * LD + e % NP];
}

// ---------------------------------------------------------------------------
// Rotation log: per zone, records of two doubles.  A chain (one QL sweep on
// rows l..m) starts with a header {l + 65536 m, count} followed by `count`
// records {c, s} for indices i = m-1, m-2, ...
// ---------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(64) void k_tql(int nzc, int64_t z0, const int64_t* __restrict__ sel_ptr,
                                            double* __restrict__ dvec, double* __restrict__ evec,
                                            double2* __restrict__ rlog, int64_t cap,
                                            int32_t* __restrict__ status) {
  __shared__ double d[NP][64];
  __shared__ double e[NP][64];
  const int lane = threadIdx.x;
  const int zl = blockIdx.x * 64 + lane;
  const bool valid = zl < nzc && sel_ptr[z0 + zl + 1] > sel_ptr[z0 + zl];
  if (valid) {
    for (int k = 0; k < NP; ++k) {
      d[k][lane] = dvec[(size_t)k * nzc + zl];
      e[k][lane] = evec[(size_t)k * nzc + zl];
    }
  }
  int st = 0;
  int64_t pos = 0;
  double2* lg = rlog + (size_t)(valid ? zl : 0) * cap;
  const double eps = 2.220446049250313e-16;
  if (valid) {
    for (int l = 0; l < NP && st == 0; ++l) {
      int iter = 0;
      for (;;) {
        int m = l;
        for (; m < NP - 1; ++m) {
          const double dd = fabs(d[m][lane]) + fabs(d[m + 1][lane]);
          if (fabs(e[m][lane]) <= eps * dd) break;
        }
        if (m == l) break;
        if (++iter > 60) { st = 2; break; }
        if (pos + 1 + (m - l) > cap) { st = 1; break; }
        double g = (d[l + 1][lane] - d[l][lane]) / (2.0 * e[l][lane]);
        double r = sqrt(fma(g, g, 1.0));
        g = d[m][lane] - d[l][lane] + e[l][lane] / (g + (g >= 0 ? r : -r));
        double s = 1.0, c = 1.0, p = 0.0;
        const int64_t hdr = pos++;
        int cnt = 0;
        bool under = false;
        for (int i = m - 1; i >= l; --i) {
          const double f = s * e[i][lane], b = c * e[i][lane];
          r = sqrt(fma(f, f, g * g));
          e[i + 1][lane] = r;
          if (r == 0.0) {
            d[i + 1][lane] -= p;
            e[m][lane] = 0.0;
            under = true;
            break;
          }
          const double ri = 1.0 / r;
          s = f * ri;
          c = g * ri;
          g = d[i + 1][lane] - p;
          r = fma(d[i][lane] - g, s, 2.0 * c * b);
          p = s * r;
          d[i + 1][lane] = g + p;
          g = fma(c, r, -b);
          lg[pos++] = make_double2(c, s);
          ++cnt;
        }
        lg[hdr] = make_double2((double)l + 65536.0 * (double)m, (double)cnt);
        if (under) continue;
        d[l][lane] -= p;
        e[l][lane] = g;
        e[m][lane] = 0.0;
      }
    }
    for (int k = 0; k < NP; ++k) dvec[(size_t)k * nzc + zl] = d[k][lane];
  }
  if (zl < nzc) status[zl] = valid ? (st | (int)(pos << 2)) : 0;
}

// ---------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(256) void k_tql_apply(int64_t z0, const int64_t* __restrict__ sel_ptr,
                                                   double* __restrict__ qmat,
                                                   const double2* __restrict__ rlog, int64_t cap,
                                                   const int32_t* __restrict__ status) {
  constexpr int LD = NP + 1;
  __shared__ double Z[NP * LD];
  __shared__ double2 cs[NP];
  const int zl = blockIdx.x;
  const int64_t z = z0 + zl;
  if (sel_ptr[z + 1] == sel_ptr[z]) return;
  if (status[zl] & 3) return;  // failed zone: handled by the caller
  const int tid = threadIdx.x;
  double* Q = qmat + (size_t)zl * NP * NP;
  for (int e = tid; e < NP * NP; e += 256) Z[(e / NP) * LD + e % NP] = Q[e];
  const int64_t len = status[zl] >> 2;
  const double2* lg = rlog + (size_t)zl * cap;
  __syncthreads();
  int64_t pos = 0;
  while (pos < len) {
    const double2 h = lg[pos];
    const int lm = (int)h.x;
    const int l = lm & 65535, m = lm >> 16;
    const int cnt = (int)h.y;
    for (int k = tid; k < cnt; k += 256) cs[k] = lg[pos + 1 + k];
    __syncthreads();
    if (tid < NP && cnt > 0) {
      double* zr = Z + tid * LD;
      double carry = zr[m];
      for (int k = 0; k < cnt; ++k) {
        const int i = m - 1 - k;
        const double c = cs[k].x, s = cs[k].y;
        const double zi = zr[i];
        zr[i + 1] = fma(s, zi, c * carry);
        carry = fma(c, zi, -s * carry);
      }
      zr[m - cnt] = carry;
    }
    (void)l;
    pos += 1 + cnt;
    __syncthreads();
  }
  for (int e = tid; e < NP * NP; e += 256) Q[e] = Z[(e / NP) * LD + e % NP];
}

// ---------------------------------------------------------------------------
// Epilogue from eigenvectors U (rows in global memory) and eigenvalues d.
// ---------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(256) void k_eigvec_epi(int64_t z0, int N, const int64_t* __restrict__ sel_ptr,
                                                    const double* __restrict__ mat,
                                                    double* __restrict__ umat,
                                                    const double* __restrict__ dvec, int nzc,
                                                    double* __restrict__ vec,
                                                    double* __restrict__ ampl,
                                                    unsigned* __restrict__ err) {
  constexpr int H = 256 / NP;  // threads per column / row
  __shared__ double col[6][NP];
  __shared__ double part[H][NP];
  const int zl = blockIdx.x;
  const int64_t z = z0 + zl;
  const int tid = threadIdx.x;
  double* V = vec + (size_t)zl * 4 * NP;
  if (sel_ptr[z + 1] == sel_ptr[z]) {
    if (tid < NP) V[NP + tid] = 0.0;
    return;
  }
  const double* M = mat + (size_t)zl * NP * NP;
  double* U = umat + (size_t)zl * NP * NP;
  const int j = tid % NP, hh = tid / NP;
  constexpr int SEG = NP / H;
  auto colsum = [&](int ci) -> double {  // sum_r U[r][j] * col[ci][r]
    double s = 0;
    for (int r = hh * SEG; r < hh * SEG + SEG; ++r) s = fma(U[r * NP + j], col[ci][r], s);
    part[hh][j] = s;
    __syncthreads();
    double t = 0;
    for (int q = 0; q < H; ++q) t += part[q][j];
    __syncthreads();
    return t;
  };
  auto rowsum = [&](int ci) -> double {  // sum_c U[j][c] * col[ci][c]
    double s = 0;
    for (int c = hh * SEG; c < hh * SEG + SEG; ++c) s = fma(U[j * NP + c], col[ci][c], s);
    part[hh][j] = s;
    __syncthreads();
    double t = 0;
    for (int q = 0; q < H; ++q) t += part[q][j];
    __syncthreads();
    return t;
  };
  const double lam = dvec[(size_t)j * nzc + zl];  // eigenvalue of I + C for column j
  const double muc = fmax(lam, 1.0);
  if (hh == 0) {
    col[0][j] = V[j];                  // b
    col[1][j] = j < N ? 1.0 : 0.0;     // 1_N
  }
  __syncthreads();
  const double c1 = colsum(0);
  const double c2 = colsum(1);
  if (hh == 0) {
    col[2][j] = c1 / muc;
    col[3][j] = c2 * sqrt(muc);
  }
  __syncthreads();
  const double a0 = rowsum(2);
  const double vr = rowsum(3);
  // refinement a += U diag(1/mu) U^T (b - M a)
  if (hh == 0) col[4][j] = a0;
  __syncthreads();
  double ma = 0;
  for (int c = hh * SEG; c < hh * SEG + SEG; ++c) ma = fma(M[j * NP + c], col[4][c], ma);
  part[hh][j] = ma;
  __syncthreads();
  double mat_a = 0;
  for (int q = 0; q < H; ++q) mat_a += part[q][j];
  __syncthreads();
  if (hh == 0) col[5][j] = col[0][j] - mat_a;
  __syncthreads();
  const double c3 = colsum(5);
  if (hh == 0) col[2][j] = c3 / muc;
  __syncthreads();
  const double ar = a0 + rowsum(2);
  if (hh == 0) col[4][j] = vr * vr;
  __syncthreads();
  double nv = 0;
  for (int q = 0; q < NP; ++q) nv += col[4][q];
  const double vn = vr / sqrt(nv);
  if (hh == 0) col[5][j] = 1.0 / sqrt(sqrt(muc));
  __syncthreads();
  for (int e = tid; e < NP * NP; e += 256) U[e] = U[e] * col[5][e % NP];
  if (hh == 0) {
    V[NP + j] = ar;
    V[2 * NP + j] = vn;
    if (ampl && j < N) ampl[(size_t)z * N + j] = ar;
    if (ar != ar) atomicOr(err, OAKG_DERR_NAN);
  }
}

// ---------------------------------------------------------------------------
template <int NP>
static void tql_pipeline_t(const ChunkArgs& A, hipStream_t st) {
  hipLaunchKernelGGL(k_tridiag<NP>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.sel_ptr, A.mat, A.ymat,
                     A.dvec, A.evec, A.nzc);
  hipLaunchKernelGGL(k_tql<NP>, dim3((A.nzc + 63) / 64), dim3(64), 0, st, A.nzc, A.z0, A.sel_ptr,
                     A.dvec, A.evec, (double2*)A.rlog, A.rlog_cap, A.tql_status);
  hipLaunchKernelGGL(k_tql_apply<NP>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.sel_ptr, A.ymat,
                     (const double2*)A.rlog, A.rlog_cap, A.tql_status);
  hipLaunchKernelGGL(k_eigvec_epi<NP>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.N, A.sel_ptr, A.mat,
                     A.ymat, A.dvec, A.nzc, A.vec, A.ampl, A.err);
}

void launch_tql_pipeline(int NP, const ChunkArgs& A, hipStream_t st) {
  switch (NP) {
    case 16: tql_pipeline_t<16>(A, st); break;
    case 32: tql_pipeline_t<32>(A, st); break;
    case 64: tql_pipeline_t<64>(A, st); break;
    case 128: tql_pipeline_t<128>(A, st); break;
  }
}

}  // namespace oakg
