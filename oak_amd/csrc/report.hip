// Innovation statistics of OAK's report (assimilation.F90:3584-3628), written
// to the log for the forecast and the analysis (:3389-3391) and per observed
// variable (:3410-3430).  For d = yo - Hx, R = diag(rmse^2) wrapped by the
// exclusion mask D (0 = excluded; DCDCovar, assimilation.F90:3086-3092, so
// that R^-1 d = D (D d) / rmse^2, covariance.F90:612-619, 425-432) and
// ingrid = count(D != 0), per observation range:
//   ingrid, sqrt(sum(d^2, D != 0) / ingrid), sum(d, D != 0) / ingrid,
//   sqrt(sum(d R^-1 d) / ingrid), sqrt(|HS^T R^-1 d|^2 / ingrid).
//
// One pass over the observations: each workgroup takes a piece of at most
// kRepRows observations of one range, keeps R^-1 d of the piece in LDS and
// writes 4 + NP partial sums; a second kernel adds the pieces of each range
// in order (the result does not depend on the launch) and forms the
// statistics.  The pass reads HS once (m N values): it is bound by HBM.
//   layout 0: HS column-major with leading dimension ldHS (caller's HS);
//             lanes run over observations (coalesced), wave w takes the
//             columns k = w + 4 c
//   layout 1: HS row-major, NP values per observation in hs_col<NP> order
//             (the ensemble path's observation anomalies, FP64 or FP32);
//             lanes run over the columns of a row, wave w over rows w + 4 i
#include "../../include/oakg.h"
#include "analysis.hpp"
#include "common.hpp"

namespace oakg {

__device__ __forceinline__ double rinv_innov(double d, double r2, bool ex) {
  return ex ? 0.0 * ((0.0 * d) / r2) : d / r2;
}

template <int NP, int LAYOUT, typename T>
__global__ __launch_bounds__(256) void k_report_part(const RepPiece* __restrict__ pieces, int N,
                                                     const double* __restrict__ d,
                                                     const double* __restrict__ rmse2,
                                                     const uint8_t* __restrict__ excl,
                                                     const T* __restrict__ HS, int64_t ldHS,
                                                     double* __restrict__ part,
                                                     unsigned* __restrict__ err) {
#pragma clang fp contract(off)
  __shared__ double rr[kRepRows];
  __shared__ double red[4][NP];
  __shared__ double sc[4][4];
  const RepPiece pc = pieces[blockIdx.x];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int nr = (int)(pc.r1 - pc.r0);
  double cnt = 0, s2 = 0, s1 = 0, sr = 0;
  for (int i = t; i < nr; i += 256) {
    const int64_t l = pc.r0 + i;
    const double dl = d[l], q = rmse2[l];
    if (!(q > 0) || !isfinite(q)) atomicOr(err, OAKG_DERR_RMSE);
    const bool ex = excl && excl[l];
    const double ri = rinv_innov(dl, q, ex);
    rr[i] = ri;
    if (!ex) {
      cnt = cnt + 1.0;
      s2 = s2 + dl * dl;
      s1 = s1 + dl;
    }
    sr = sr + dl * ri;
  }
  for (int o = 32; o > 0; o >>= 1) {
    cnt = cnt + shfl_xor_d(cnt, o);
    s2 = s2 + shfl_xor_d(s2, o);
    s1 = s1 + shfl_xor_d(s1, o);
    sr = sr + shfl_xor_d(sr, o);
  }
  if (lane == 0) {
    sc[w][0] = cnt;
    sc[w][1] = s2;
    sc[w][2] = s1;
    sc[w][3] = sr;
  }
  __syncthreads();
  if (LAYOUT == 0) {
    constexpr int NC = NP / 4;
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.0;
    for (int i = lane; i < nr; i += 64) {
      const double ri = rr[i];
      const int64_t l = pc.r0 + i;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int k = w + 4 * c;
        if (k < N) acc[c] = acc[c] + (double)HS[l + (int64_t)k * ldHS] * ri;
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double v = acc[c];
      for (int o = 32; o > 0; o >>= 1) v = v + shfl_xor_d(v, o);
      if (lane == 0) red[0][w + 4 * c] = v;
    }
  } else {
    constexpr int NH = (NP + 63) / 64;
    double acc[NH];
#pragma unroll
    for (int h = 0; h < NH; ++h) acc[h] = 0.0;
    for (int i = w; i < nr; i += 4) {
      const double ri = rr[i];
      const T* row = HS + (pc.r0 + i) * NP;
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        const int k = lane + 64 * h;
        if (k < NP) acc[h] = acc[h] + (double)row[k] * ri;
      }
    }
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      const int k = lane + 64 * h;
      if (k < NP) red[w][k] = acc[h];
    }
  }
  __syncthreads();
  double* out = part + (size_t)blockIdx.x * (4 + NP);
  if (t < NP) {
    double v;
    if (LAYOUT == 0) {
      v = red[0][t];
    } else {
      const int k = hs_col<NP>(t);  // position of column t in a row
      v = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
    }
    out[4 + t] = v;
  }
  if (t < 4) out[t] = ((sc[0][t] + sc[1][t]) + sc[2][t]) + sc[3][t];
}

template <int NP>
__global__ __launch_bounds__(256) void k_report_finish(const int64_t* __restrict__ pbeg, int N,
                                                       const double* __restrict__ part,
                                                       double* __restrict__ stats) {
#pragma clang fp contract(off)
  __shared__ double v[4 + NP];
  const int r = blockIdx.x, t = threadIdx.x;
  const int64_t p0 = pbeg[r], p1 = pbeg[r + 1];
  if (t < 4 + N) {
    double s = 0;
    for (int64_t p = p0; p < p1; ++p) s = s + part[p * (4 + NP) + t];
    v[t] = s;
  }
  __syncthreads();
  if (t == 0) {
    double pp = 0;
    for (int k = 0; k < N; ++k) pp = pp + v[4 + k] * v[4 + k];
    const double g = v[0];
    double* o = stats + 5 * (size_t)r;
    o[0] = g;
    o[1] = sqrt(v[1] / g);
    o[2] = v[2] / g;
    o[3] = sqrt(v[3] / g);
    o[4] = sqrt(pp / g);
  }
}

template <int NP>
static void report_t(const ReportArgs& A, hipStream_t st) {
  if (A.npieces > 0) {
    if (A.layout == 0)
      hipLaunchKernelGGL((k_report_part<NP, 0, double>), dim3((unsigned)A.npieces), dim3(256), 0,
                         st, A.pieces, A.N, A.d, A.rmse2, A.excl, A.HS, A.ldHS, A.part, A.err);
    else if (A.HS32)
      hipLaunchKernelGGL((k_report_part<NP, 1, float>), dim3((unsigned)A.npieces), dim3(256), 0,
                         st, A.pieces, A.N, A.d, A.rmse2, A.excl, A.HS32, (int64_t)NP, A.part,
                         A.err);
    else
      hipLaunchKernelGGL((k_report_part<NP, 1, double>), dim3((unsigned)A.npieces), dim3(256), 0,
                         st, A.pieces, A.N, A.d, A.rmse2, A.excl, A.HS, (int64_t)NP, A.part,
                         A.err);
  }
  hipLaunchKernelGGL(k_report_finish<NP>, dim3((unsigned)A.nranges), dim3(256), 0, st, A.pbeg,
                     A.N, A.part, A.stats);
}

void launch_report(int NP, const ReportArgs& A, hipStream_t st) {
  switch (NP) {
    case 16: report_t<16>(A, st); break;
    case 32: report_t<32>(A, st); break;
    case 64: report_t<64>(A, st); break;
    case 128: report_t<128>(A, st); break;
  }
}

}  // namespace oakg
