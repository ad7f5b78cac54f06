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
// and forms the statistics.  The pass reads HS once (m N values): it is bound by HBM.
//   layout 0: HS column-major with leading dimension ldHS (caller's HS);
//             lanes run over observations (coalesced), wave w takes the
//             columns k = w + 4 c
//   layout 1: HS row-major, NP values per observation in hs_col<NP> order
//             (the ensemble path's observation anomalies, FP64 or FP32);
//             lanes run over the columns of a row, wave w over rows w + 4 i
// The partial sums of the pieces are added per range in a fixed order, so
// the bits of the result do not depend on the launch.
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
  // The loads of an iteration do not depend on a branch (out-of-range rows
  // and columns read a valid element and add nothing), so each wave keeps
  // all of them in flight at once.
  if (LAYOUT == 0) {
    // the columns of a wave in passes of CC (CC accumulators and CC loads in
    // flight per lane: few enough registers for six waves per SIMD)
    constexpr int NC = NP / 4;
    constexpr int CC = NC < 16 ? NC : 16;
#pragma unroll 1
    for (int c0 = 0; c0 < NC; c0 += CC) {
      double acc[CC];
#pragma unroll
      for (int c = 0; c < CC; ++c) acc[c] = 0.0;
      for (int i0 = 0; i0 < nr; i0 += 64) {
        const int i = i0 + lane;
        const bool ok = i < nr;
        const double ri = ok ? rr[i] : 0.0;
        const int64_t l = pc.r0 + (ok ? i : 0);
        double v[CC];
#pragma unroll
        for (int c = 0; c < CC; ++c) {
          const int k = min(w + 4 * (c0 + c), N - 1);
          v[c] = (double)HS[l + (int64_t)k * ldHS];
        }
#pragma unroll
        for (int c = 0; c < CC; ++c) acc[c] = acc[c] + (ok ? v[c] * ri : 0.0);
      }
#pragma unroll
      for (int c = 0; c < CC; ++c) {
        double v = acc[c];
        for (int o = 32; o > 0; o >>= 1) v = v + shfl_xor_d(v, o);
        if (lane == 0) red[0][w + 4 * (c0 + c)] = v;
      }
    }
  } else {
    constexpr int NH = (NP + 63) / 64;
    constexpr int U = NP >= 64 ? 8 : 16;  // rows per iteration of a wave
    double acc[NH];
#pragma unroll
    for (int h = 0; h < NH; ++h) acc[h] = 0.0;
    const int kl = min(lane, NP - 1);
    for (int i0 = w; i0 < nr; i0 += 4 * U) {
      double v[U][NH], ri[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = i0 + 4 * u;
        ok[u] = i < nr;
        ri[u] = ok[u] ? rr[i] : 0.0;
        const T* row = HS + (pc.r0 + (ok[u] ? i : 0)) * NP;
#pragma unroll
        for (int h = 0; h < NH; ++h) v[u][h] = (double)row[kl + 64 * h];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int h = 0; h < NH; ++h) acc[h] = acc[h] + (ok[u] ? v[u][h] * ri[u] : 0.0);
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

// One workgroup per range: 8 segments of its pieces x 128 columns, each
// thread with eight interleaved partial sums (independent loads); the
// segments are then combined in a fixed order.
template <int NP>
__global__ __launch_bounds__(1024) void k_report_finish(const int64_t* __restrict__ pbeg, int N,
                                                        const double* __restrict__ part,
                                                        double* __restrict__ stats) {
#pragma clang fp contract(off)
  constexpr int W = 4 + NP;  // partials per piece
  __shared__ double seg[8][W];
  __shared__ double v[W];
  const int r = blockIdx.x, t = threadIdx.x;
  const int cl = t & 127, sg = t >> 7;
  const int64_t p0 = pbeg[r], p1 = pbeg[r + 1];
  const int64_t len = (p1 - p0 + 7) / 8;
  const int64_t a = p0 + sg * len, b = min(p1, a + len);
  for (int col = cl; col < W; col += 128) {
    double s[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = 0.0;
    int64_t p = a;
    for (; p + 8 <= b; p += 8) {
      double x[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = part[(p + j) * W + col];
#pragma unroll
      for (int j = 0; j < 8; ++j) s[j] = s[j] + x[j];
    }
#pragma unroll
    for (int j = 0; j < 7; ++j)
      if (p + j < b) s[j] = s[j] + part[(p + j) * W + col];
    seg[sg][col] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
  }
  __syncthreads();
  if (t < W) {
    double x = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) x = x + seg[j][t];
    v[t] = x;
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
  hipLaunchKernelGGL(k_report_finish<NP>, dim3((unsigned)A.nranges), dim3(1024), 0, st, A.pbeg,
                     A.N, A.part, A.stats);
}

void launch_report(int NP, const ReportArgs& A, hipStream_t st) {
  switch (NP) {
    case 16: report_t<16>(A, st); break;
    case 32: report_t<32>(A, st); break;
    case 64: report_t<64>(A, st); break;
    case 128: report_t<128>(A, st); break;
    case 256: report_t<256>(A, st); break;
  }
}

}  // namespace oakg
