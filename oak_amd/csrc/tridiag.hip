// FP64 symmetric eigensolver by Householder tridiagonalisation and implicit
// QL iterations (the method of LAPACK's dsyev = dsytrd + dorgtr + dsteqr,
// which OAK's symeig calls, matoper_inc.F90:986-1005), restructured for a
// batch of zones:
//
//   k_tridiag      one workgroup per zone, matrix in registers (one tile
//                  per thread): M = Q T Q^T with Householder reflectors
//                  (lower), then Q formed explicitly by backward
//                  accumulation.  Writes Q and (d, e).
//   k_tql          one lane per zone (4-32 zones per wave): implicit QL with
//                  Wilkinson shifts on (d, e) (Golub & Van Loan, 4th ed.,
//                  section 8.3; Bowdler, Martin, Reinsch & Wilkinson 1968);
//                  the sequential recurrences of many zones run side by side.
//                  Every plane rotation is logged; d ends up holding the
//                  eigenvalues.
//   k_tql_apply    (N <= 32) one workgroup per zone: applies the logged
//                  rotations to the rows of Q in LDS, giving U = Q Z;
//                  k_tql_apply64 (33 <= N <= 64) the same on rows held whole
//                  in registers, k_tql_apply256 (N > 128) on 128-row blocks,
//                  each row split over two threads.
//   k_eigvec_epi   (N <= 64, N > 128) per zone: a = U diag(1/mu) U^T b (+ one
//                  refinement step against M), v ~ U diag(mu^1/2) U^T 1,
//                  Y = U diag(mu^-1/4), with mu = max(lambda, 1)
//                  (rrsqrt.F90:137-142, 176-177); mu itself is kept beside a
//                  and v for the likelihood.
//   k_tql_apply128 (65 <= N <= 128) the rotations on rows held in registers
//                  (each row split over two threads), then the epilogue of
//                  k_eigvec_epi.
// The reductions on the lower triangle only (k_tridiag_lower, with Q formed
// by k_formq64 at NP = 64) are in wide256.hip.
#include <type_traits>

#include "analysis.hpp"
#include "common.hpp"
#include "eigen_common.hpp"

namespace oakg {

// ---------------------------------------------------------------------------
// Householder tridiagonalisation (lower, as dsytd2) with the matrix in
// registers: thread t of 256 holds one T x T tile, T = NP/16.  Each step
// extracts column i through LDS; every wave forms the reflector's scalars
// from it; the tiles give their partial sums of A v and their share of
// v^T A v (so p.v needs no second reduction); p = tau A v is summed per row;
// and A -= v w^T + w v^T (w = p - tau/2 (p.v) v) is applied to the tiles that
// meet the trailing block.  Three barriers per step.  The reflectors go to
// the rows of qmat; Q is then accumulated backwards in the same tile layout
// and overwrites them.
// Up to three stages: once the trailing block is the lower-right half, it is
// passed through LDS to tiles of T/2 over that half, and at the last quarter
// to tiles of T/4, so the late steps (and the early steps of the backward Q
// accumulation, whose active block is as small) run on all four waves with a
// quarter (a sixteenth) of the work per thread instead of on wave 0 alone.
// T and Q are written in reversed order (d'_k = d_{NP-1-k}, columns of Q
// reversed): the reduction from the top leaves the large diagonal entries at
// the top of T, and QL, which deflates at the top of each block, converges
// slowly on such grading (tens of sweeps per eigenvalue for strongly
// observed zones; LAPACK's dsteqr switches to QR there).  Reversed, the
// small entries come first and QL needs a few sweeps per eigenvalue.
// ---------------------------------------------------------------------------
// The reduction indexes its row vectors with one pad entry per 8 rows,
// tri_rx(r) = r + r / 8: the tiles of the first stage start every 8 rows,
// and without the pad the reads of v and p by the 16 tile rows of a wave
// fall four to a bank.

// XF: side of the largest trailing block moved between stages (NP/2 from
// the first stage, NP/4 when a launch starts at the second).
template <int NP, int XF = NP / 2>
struct TriLds {
  static constexpr int PR = NP + NP / 8;  // padded length of a row vector
  double colw[4][PR]; // each wave's copy of v (reduction)
  double pub[2][PR];  // column i+1 before step i's update (double-buffered)
  double col[2][NP];  // Householder vector v (Q accumulation, double-buffered)
  double pw[2][PR];   // p (then y during the Q accumulation)
  union {
    double part[16][PR + 1];             // per-tile partial sums
    double xfer[XF][XF + 1];             // the trailing block between the stages
  };
  double red[2][4];   // per-wave shares of v^T A v (double-buffered)
  double tau[NP];
};

// Tile (tr, tc) of a 16 x 16 tile grid held by thread t, in decreasing order
// of min(tr, tc): the tiles that meet a trailing block [i+1, n)^2 are always
// threads 0 .. (16-k)^2 - 1 (k = the smallest active tile index), so whole
// waves fall idle as the block shrinks instead of a few lanes in every wave.
// Thread t lies on ring s = floor(sqrt t) (tiles with min index 15 - s),
// position j = t - s^2: j <= s on the ring's row, j > s on its column.
__device__ __forceinline__ void ring_tile(int t, int& tr, int& tc) {
  int s = (int)sqrtf((float)t);
  while (s * s > t) --s;
  while ((s + 1) * (s + 1) <= t) ++s;
  const int j = t - s * s, mn = 15 - s;
  if (j <= s) { tr = mn; tc = mn + j; }
  else { tr = mn + (j - s); tc = mn; }
}

// PART selects the steps a launch runs: 0 all of them (and, for NP < 128,
// the accumulation of Q); for NP = 128 the reduction can run as two
// launches, PART 1 the first stage (the 8 x 8 tiles that need most of the
// registers) ending with the trailing half-block [H2, NP)^2 stored in the
// rows H2 .. NP-1 of qmat (not yet holding reflectors: row r, column c - H2),
// then PART 2 the later stages from that block; or as three, PART 3 (steps
// 0 .. 31), PART 4 (steps 32 .. 63 on 6 x 6 tiles of the 96 x 96 trailing
// block) and PART 2.  The later launches need a fraction of the registers,
// so more zones share a CU through their latency-bound steps.  Every way
// runs the same operations on the same values: bit-identical T and
// reflectors.
template <int NP, int PART>
__global__ __launch_bounds__(256) void k_tridiag(int64_t z0, const int32_t* __restrict__ zorder, const int64_t* __restrict__ sel_ptr,
                                                 const double* __restrict__ mat,
                                                 double* __restrict__ qmat,
                                                 double* __restrict__ dvec, double* __restrict__ evec,
                                                 int nzc, int64_t zstride, int elast) {
  // zstride: doubles between the matrices (and the reflector blocks) of
  // consecutive zones, NP^2 except for the trailing blocks of NP = 256;
  // elast = 0 leaves e_{NP-1} to the caller (there, the coupling of the two
  // halves of the NP = 256 tridiagonal)
  // Stage s runs on tiles of Ts over the block [Hs, NP)^2 = 16 Ts rows:
  // T1 = NP/16 over the whole matrix, then T2 = T1/2 and T3 = T2/2 (as far
  // as the tiles go down to 1 x 1).
  constexpr int T1 = NP / 16;
  constexpr int T2 = T1 >= 2 ? T1 / 2 : T1;
  constexpr int T3 = T2 >= 2 ? T2 / 2 : T2;
  constexpr int H2 = NP - 16 * T2, H3 = NP - 16 * T3;
  __shared__ TriLds<NP, PART == 2 ? NP / 4 : (PART >= 3 ? 1 : NP / 2)> S;
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;  // the zone is solved by the other pipeline
  if (sel_ptr[z + 1] == sel_ptr[z]) return;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  int tr, tc;
  ring_tile(t, tr, tc);
  const int r1 = tr * T1, c1 = tc * T1;            // tile origins per stage
  const int r2 = H2 + tr * T2, c2 = H2 + tc * T2;
  const int r3 = H3 + tr * T3, c3 = H3 + tc * T3;
  const double* M = mat + (size_t)zl * zstride;
  double* V = qmat + (size_t)zl * zstride;

  // Steps i0 .. i1-1 of the reduction on tiles of T at (r0, c0); tile
  // column tc covers columns base + tc*T ...  Every wave carries the current
  // column i in registers (x[h] = A[lane + 64 h][i]) and forms the
  // reflector's scalars from it; v goes to the wave's own LDS copy
  // (S.colw[wave]) for the tiles.  Step i publishes column i+1 as it stands
  // before the step's update, and every wave then forms the updated column
  // itself, so a step needs two barriers.  The threads that sum the rows of
  // A v (the last NP, whose waves hold the fewest active tiles in the ring
  // order) also form v^T p; the last wave writes the reflector out.
  constexpr int NH = (NP + 63) / 64;
  constexpr int PW0 = (256 - NP) / 64;  // first wave of the row sums
  double x[NH];
  double* col = S.colw[wave];
  auto reduce_steps = [&](auto& a, auto Tk, int r0, int c0, int base, int i0, int i1) {
    constexpr int T = decltype(Tk)::value;
    // the padded index of a tile's row (column) ii: for T dividing 8 the T
    // rows lie in one group of 8, tri_rx(r0 + ii) = tri_rx(r0) + ii; tiles of
    // 6 can straddle a group
    const int R0 = tri_rx(r0), C0 = tri_rx(c0);
    auto RX = [&](int ii) { return 8 % T == 0 ? R0 + ii : tri_rx(r0 + ii); };
    auto CX = [&](int jj) { return 8 % T == 0 ? C0 + jj : tri_rx(c0 + jj); };
    for (int i = i0; i < i1; ++i) {
      const int b = i & 1;
      double* pw = S.pw[b];
      double* pub = S.pub[b];
      // (2) reflector H = I - tau v v^T with H x = beta e_1, x = A[i+1.., i].
      // Every wave forms the scalars (identical results), so no barrier is
      // needed before v is used.
      double tau = 0.0, scal = 0.0, beta;
      double v[NH];
      {
        double s = 0.0;
#pragma unroll
        for (int h = 0; h < NH; ++h) {
          const int r = lane + 64 * h;
          if (r >= i + 2 && r < NP) s = fma(x[h], x[h], s);
        }
        s = wave_sum_d(s);
        const int ia = i + 1, id = i;
        const double alpha = readlane_d((ia >> 6) ? x[NH - 1] : x[0], ia & 63);
        const double di = readlane_d((id >> 6) ? x[NH - 1] : x[0], id & 63);
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
        // v: 0 up to row i, 1 at i+1, x * scal below
#pragma unroll
        for (int h = 0; h < NH; ++h) {
          const int r = lane + 64 * h;
          v[h] = r > i + 1 ? x[h] * scal : (r == i + 1 ? 1.0 : 0.0);
          if (r < NP) col[tri_rx(r)] = v[h];
        }
        __builtin_amdgcn_wave_barrier();
        if (wave == 3) {
          // NP = 128: tau goes to the unused entry i of row i (read by k_formq128)
#pragma unroll
          for (int h = 0; h < NH; ++h) {
            const int r = lane + 64 * h;
            if (r < NP) V[(size_t)i * NP + r] = r == i ? (NP == 128 ? tau : 0.0) : v[h];
          }
          if (lane == 0) {
            dvec[(size_t)(NP - 1 - i) * nzc + zl] = di;
            evec[(size_t)(NP - 2 - i) * nzc + zl] = beta;
            S.tau[i] = tau;
          }
        }
      }
      // (3) per-tile partial sums of A v (an inactive tile has zero sums;
      // its v entries are then not used)
      const bool act = (r0 + T - 1 > i) && (c0 + T - 1 > i);
      double vR[T], vC[T];
#pragma unroll
      for (int jj = 0; jj < T; ++jj) vC[jj] = col[CX(jj)];
#pragma unroll
      for (int ii = 0; ii < T; ++ii) vR[ii] = col[RX(ii)];
#pragma unroll
      for (int ii = 0; ii < T; ++ii) {
        double ps = 0;
        if (act) {
#pragma unroll
          for (int jj = 0; jj < T; ++jj) ps = fma(a[ii][jj], vC[jj], ps);
        }
        S.part[tc][RX(ii)] = ps;
      }
      // column i+1 before this step's update, from the threads of its tile column
      if (tc == (i + 1 - base) / T) {
        const int j0 = (i + 1 - base) % T;
#pragma unroll
        for (int jj = 0; jj < T; ++jj)
          if (jj == j0) {
#pragma unroll
            for (int ii = 0; ii < T; ++ii) pub[RX(ii)] = a[ii][jj];
          }
      }
      lds_barrier();
      // (4) p = tau A v and the waves' shares of v^T p.  Rows above the block
      // (later stages) read stale sums and get p = 0 (they are <= i).
      if (wave >= PW0) {
        const int row = t - (256 - NP), rowx = tri_rx(row);
        double vp = 0.0;
        if (row >= 0) {
          double s4[4];
#pragma unroll
          for (int qq = 0; qq < 4; ++qq) s4[qq] = S.part[qq][rowx];
#pragma unroll
          for (int qq = 4; qq < 16; ++qq) s4[qq & 3] += S.part[qq][rowx];
          const double s = (s4[0] + s4[1]) + (s4[2] + s4[3]);
          const double pr = row > i ? tau * s : 0.0;
          pw[rowx] = pr;
          vp = col[rowx] * pr;
        }
        vp = wave_sum_d(vp);
        if (lane == 0) S.red[b][wave] = vp;
      }
      lds_barrier();
      // (5) A -= v w^T + w v^T on the tiles that meet the trailing block;
      // w = p + K v, K = -tau/2 v^T p
      double vpt = 0.0;
#pragma unroll
      for (int w = PW0; w < 4; ++w) vpt += S.red[b][w];
      const double K = -0.5 * tau * vpt;
      if (act) {
        double wR[T], wC[T];
#pragma unroll
        for (int ii = 0; ii < T; ++ii) wR[ii] = fma(K, vR[ii], pw[RX(ii)]);
#pragma unroll
        for (int jj = 0; jj < T; ++jj) wC[jj] = fma(K, vC[jj], pw[CX(jj)]);
#pragma unroll
        for (int ii = 0; ii < T; ++ii)
#pragma unroll
          for (int jj = 0; jj < T; ++jj)
            a[ii][jj] = fma(-vR[ii], wC[jj], fma(-wR[ii], vC[jj], a[ii][jj]));
      }
      // column i+1 after the update (the same operations as on the tiles:
      // v_{i+1} = 1, w_{i+1} = p_{i+1} + K)
      {
        const double w1 = fma(K, 1.0, pw[tri_rx(i + 1)]);
#pragma unroll
        for (int h = 0; h < NH; ++h) {
          const int r = lane + 64 * h;
          if (r < NP) {
            const double w = fma(K, v[h], pw[tri_rx(r)]);
            const double nc = fma(-v[h], w1, fma(-w, 1.0, pub[tri_rx(r)]));
            x[h] = r > i ? nc : 0.0;
          }
        }
      }
    }
  };

  // Backward accumulation Q <- H_k Q for k = k_hi .. k_lo on tiles of T at
  // (r0, c0); the reflector rows are read one step ahead into vnext.
  double vnext = 0.0;
  auto q_steps = [&](auto& q, auto Tk, int r0, int c0, int k_hi, int k_lo) {
    constexpr int T = decltype(Tk)::value;
    for (int k = k_hi; k >= k_lo; --k) {
      const int b = k & 1;
      double* vb = S.col[b];
      double* yb = S.pw[b];
      if (t < NP) vb[t] = vnext;
      if (k > 0 && t < NP) vnext = V[(size_t)(k - 1) * NP + t];  // prefetch
      lds_barrier();
      const bool act = (r0 + T - 1 > k) && (c0 + T - 1 > k);
      double vR[T];
#pragma unroll
      for (int ii = 0; ii < T; ++ii) vR[ii] = act ? vb[r0 + ii] : 0.0;
#pragma unroll
      for (int jj = 0; jj < T; ++jj) {
        double s = 0;
        if (act) {
#pragma unroll
          for (int ii = 0; ii < T; ++ii) s = fma(vR[ii], q[ii][jj], s);
        }
        S.part[tr][c0 + jj] = s;
      }
      lds_barrier();
      if (t < NP) {
        double s = 0;
#pragma unroll
        for (int qq = 0; qq < 16; ++qq) s += S.part[qq][t];
        yb[t] = S.tau[k] * s;
      }
      lds_barrier();
      if (act) {
#pragma unroll
        for (int ii = 0; ii < T; ++ii)
#pragma unroll
          for (int jj = 0; jj < T; ++jj) q[ii][jj] = fma(-vR[ii], yb[c0 + jj], q[ii][jj]);
      }
    }
  };

  // The trailing block [Hd, NP)^2 from tiles of Ts at (rs, cs) to tiles of
  // Td at (rd, cd) through LDS; destination tiles outside the block are set
  // to the identity.  The callers' last reads of part (aliasing xfer) ended
  // before the last barrier of their loop, and part is written again only
  // after the first barrier of the next one.
  auto move_block = [&](auto& src, auto Ts, int rs, int cs, auto& dst, auto Td, int rd, int cd,
                        int Hd) {
    constexpr int TS = decltype(Ts)::value, TD = decltype(Td)::value;
    if (rs >= Hd && cs >= Hd) {
#pragma unroll
      for (int ii = 0; ii < TS; ++ii)
#pragma unroll
        for (int jj = 0; jj < TS; ++jj) S.xfer[rs - Hd + ii][cs - Hd + jj] = src[ii][jj];
    }
    lds_barrier();
    const bool in = rd >= Hd && cd >= Hd;
#pragma unroll
    for (int ii = 0; ii < TD; ++ii)
#pragma unroll
      for (int jj = 0; jj < TD; ++jj)
        dst[ii][jj] = in ? S.xfer[rd - Hd + ii][cd - Hd + jj] : (rd + ii == cd + jj ? 1.0 : 0.0);
    lds_barrier();  // xfer read before part is written again
  };
  auto copy_tile = [&](auto& src, auto& dst, auto Tk) {
    constexpr int TT = decltype(Tk)::value;
#pragma unroll
    for (int ii = 0; ii < TT; ++ii)
#pragma unroll
      for (int jj = 0; jj < TT; ++jj) dst[ii][jj] = src[ii][jj];
  };
  using K1 = std::integral_constant<int, T1>;
  using K2 = std::integral_constant<int, T2>;
  using K3 = std::integral_constant<int, T3>;

  static_assert(PART == 0 || (NP == 128 && T2 < T1), "the staged launches are for NP = 128");
  // three launches: steps 0 .. 31 on the 8 x 8 tiles, ending with the block
  // [32, NP)^2 in rows 32 .. NP-1 of qmat (PART 3); steps 32 .. H2-1 on 6 x 6
  // tiles of that block (PART 4, 16 x 6 = 96), ending with [H2, NP)^2 as
  // PART 1 leaves it; then PART 2
  constexpr int H6 = NP - 16 * 6;
  if constexpr (PART == 4) {
    using K6 = std::integral_constant<int, 6>;
    const int r6 = H6 + tr * 6, c6 = H6 + tc * 6;
    double a6[6][6];
#pragma unroll
    for (int ii = 0; ii < 6; ++ii)
#pragma unroll
      for (int jj = 0; jj < 6; ++jj) a6[ii][jj] = V[(size_t)(r6 + ii) * NP + (c6 + jj - H6)];
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      const int r = lane + 64 * h;
      x[h] = (r >= H6 && r < NP) ? V[(size_t)r * NP] : 0.0;
    }
    __syncthreads();  // every read of the block ends before row H6 gets its reflector
    reduce_steps(a6, K6{}, r6, c6, H6, H6, H2);
#pragma unroll
    for (int ii = 0; ii < 6; ++ii)
#pragma unroll
      for (int jj = 0; jj < 6; ++jj)
        if (r6 + ii >= H2 && c6 + jj >= H2)
          V[(size_t)(r6 + ii) * NP + (c6 + jj - H2)] = a6[ii][jj];
    return;
  }
  // ---- reduction ----
  double a3[T3][T3];
  {
    double a2[T2][T2];
    if constexpr (PART == 2) {
      // the trailing block and its first column from the rows of qmat
#pragma unroll
      for (int ii = 0; ii < T2; ++ii)
#pragma unroll
        for (int jj = 0; jj < T2; ++jj) a2[ii][jj] = V[(size_t)(r2 + ii) * NP + (c2 + jj - H2)];
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        const int r = lane + 64 * h;
        x[h] = (r >= H2 && r < NP) ? V[(size_t)r * NP] : 0.0;
      }
      __syncthreads();  // every read of the block ends before row H2 gets its reflector
    } else {
      double a[T1][T1];
#pragma unroll
      for (int ii = 0; ii < T1; ++ii)
#pragma unroll
        for (int jj = 0; jj < T1; ++jj) a[ii][jj] = M[(r1 + ii) * NP + c1 + jj];
#pragma unroll
    for (int h = 0; h < NH; ++h) {  // column 0
      const int r = lane + 64 * h;
      x[h] = r < NP ? M[(size_t)r * NP] : 0.0;
    }
      if constexpr (PART == 3) {
        reduce_steps(a, K1{}, r1, c1, 0, 0, H6);
#pragma unroll
        for (int ii = 0; ii < T1; ++ii)
#pragma unroll
          for (int jj = 0; jj < T1; ++jj)
            if (r1 + ii >= H6 && c1 + jj >= H6)
              V[(size_t)(r1 + ii) * NP + (c1 + jj - H6)] = a[ii][jj];
        return;
      }
      if constexpr (T2 < T1) {
        reduce_steps(a, K1{}, r1, c1, 0, 0, H2);
        if constexpr (PART == 1) {
          // the trailing block to rows H2 .. NP-1 of qmat (the tiles, updated
          // by step H2-1, hold it; the column x carries is its column 0)
          if (r1 >= H2 && c1 >= H2) {
#pragma unroll
            for (int ii = 0; ii < T1; ++ii)
#pragma unroll
              for (int jj = 0; jj < T1; ++jj) V[(size_t)(r1 + ii) * NP + (c1 + jj - H2)] = a[ii][jj];
          }
          return;
        }
        move_block(a, K1{}, r1, c1, a2, K2{}, r2, c2, H2);
      } else {
        copy_tile(a, a2, K1{});
      }
    }
    if constexpr (T3 < T2) {
      reduce_steps(a2, K2{}, r2, c2, H2, H2, H3);
      move_block(a2, K2{}, r2, c2, a3, K3{}, r3, c3, H3);
    } else {
      copy_tile(a2, a3, K2{});
    }
  }
  reduce_steps(a3, K3{}, r3, c3, H3, H3, NP - 2);
  // trailing 2 x 2 block
#pragma unroll
  for (int ii = 0; ii < T3; ++ii)
#pragma unroll
    for (int jj = 0; jj < T3; ++jj) {
      const int r = r3 + ii, c = c3 + jj;
      if (c == NP - 2 && r == NP - 2) dvec[(size_t)1 * nzc + zl] = a3[ii][jj];
      if (c == NP - 2 && r == NP - 1) evec[(size_t)0 * nzc + zl] = a3[ii][jj];
      if (c == NP - 1 && r == NP - 1) dvec[(size_t)0 * nzc + zl] = a3[ii][jj];
    }
  if (t == 0 && elast) evec[(size_t)(NP - 1) * nzc + zl] = 0.0;
  if constexpr (NP == 128) return;  // Q is formed by k_formq128 on MFMA

  // ---- Q = H_0 H_1 ... H_{NP-3}, accumulated backwards: Q <- H_k Q ----
  // While k >= Hs, Q differs from I only on the block [Hs, NP)^2.
  double q[T1][T1];
  {
    double q2[T2][T2];
    {
      double q3[T3][T3];
#pragma unroll
      for (int ii = 0; ii < T3; ++ii)
#pragma unroll
        for (int jj = 0; jj < T3; ++jj) q3[ii][jj] = (r3 + ii == c3 + jj) ? 1.0 : 0.0;
      __syncthreads();  // reflectors written by wave 0 are visible to the workgroup
      if (t < NP) vnext = V[(size_t)(NP - 3) * NP + t];
      q_steps(q3, K3{}, r3, c3, NP - 3, H3);
      if constexpr (T3 < T2) {
        move_block(q3, K3{}, r3, c3, q2, K2{}, r2, c2, H3);
        q_steps(q2, K2{}, r2, c2, H3 - 1, H2);
      } else {
        copy_tile(q3, q2, K3{});
      }
    }
    if constexpr (T2 < T1) {
      move_block(q2, K2{}, r2, c2, q, K1{}, r1, c1, H2);
      q_steps(q, K1{}, r1, c1, H2 - 1, 0);
    } else {
      copy_tile(q2, q, K2{});
    }
  }
  lds_barrier();
#pragma unroll
  for (int ii = 0; ii < T1; ++ii)
#pragma unroll
    for (int jj = 0; jj < T1; ++jj) V[(size_t)(r1 + ii) * NP + (NP - 1 - (c1 + jj))] = q[ii][jj];
}

// ---------------------------------------------------------------------------
// Q = H_0 H_1 ... H_125 for NP = 128 from the reflectors k_tridiag leaves in
// the rows of qmat (row k: v_k, with tau_k in its otherwise zero entry k),
// accumulated backwards in blocks of 16 reflectors in compact WY form
// (LAPACK dlarft / dlarfb): H_k0 ... H_k0+15 = I - Vb Tb Vb^T with Tb upper
// triangular, and Q <- Q - Vb (Tb (Vb^T Q)), all products on
// v_mfma_f64_16x16x4f64.  Q stays in the accumulators: wave w holds tile
// columns w and 7 - w (8 tiles of 16 x 16 each, so the waves get similar
// work as the active block grows); block b touches only tiles (I, J) with
// I, J >= b.  The D register kk of a tile is the B operand of the k-step
// over its rows 4 kk .. 4 kk + 3, so the three products chain in registers.
// The next block's reflectors are loaded while a block is applied.  Q is
// written with its columns reversed, as k_tridiag writes it for NP < 128.
// ---------------------------------------------------------------------------
struct FormQLds {
  // padded row of Vb and Tb: with 18 doubles the MFMA operand reads of 64
  // lanes (16 rows x 4, or 16 columns x 4) and the staging writes below fall
  // on distinct bank pairs as far as 64 lanes allow
  static constexpr int LV = 18;
  double vb[2][128 * LV];        // Vb[r][j]: row r of reflector k0 + j
  double tb[2][16 * LV];         // Tb
  double g[16 * LV];             // Vb^T Vb
  double tau[2][16];
};

__global__ __launch_bounds__(256, 2) void k_formq128(int64_t z0, const int32_t* __restrict__ zorder, const int64_t* __restrict__ sel_ptr,
                                                  double* __restrict__ qmat) {
  constexpr int NP = 128, NB = 8, LV = FormQLds::LV;
  __shared__ FormQLds S;
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;  // the zone is solved by the other pipeline
  if (sel_ptr[z + 1] == sel_ptr[z]) return;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int lr = lane >> 4, lc = lane & 15;
  double* V = qmat + (size_t)zl * NP * NP;
  const int J0 = wave, J1 = 7 - wave;
  d4 q[2][8];
#pragma unroll
  for (int jj = 0; jj < 2; ++jj)
#pragma unroll
    for (int I = 0; I < 8; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        q[jj][I][r] = (I == (jj ? J1 : J0) && lr + 4 * r == lc) ? 1.0 : 0.0;
  // thread t moves reflector pj = t / 16 of a block, rows pr + 16 e (e < 8):
  // 16 threads read 16 consecutive values of a reflector row, and their
  // LDS stores (row stride LV) spread over the banks
  const int pj = t >> 4, pr = t & 15;
  double pf[8];
  auto load_block = [&](int b) {
    const int k = 16 * b + pj;
    if (k < NP - 2) {
      const double* src = V + (size_t)k * NP + pr;
#pragma unroll
      for (int e = 0; e < 8; ++e) pf[e] = src[16 * e];
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) pf[e] = 0.0;
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
      for (int e = 0; e < 8; ++e) {
        const int r = pr + 16 * e;
        vb[r * LV + pj] = r <= k ? 0.0 : pf[e];  // v_k is zero above row k + 1
        if (r == k) S.tau[buf][pj] = pf[e];
      }
    }
    if (b > 0) load_block(b - 1);
    lds_barrier();
    // (1) wave 0: G = Vb^T Vb (rows >= 16 b), then Tb column by column:
    //     Tb[j][j] = tau_j, Tb[0:j, j] = -tau_j Tb[0:j, 0:j] G[0:j, j]
    if (wave == 0) {
      d4 g = {0.0, 0.0, 0.0, 0.0};
      for (int kk = 4 * b; kk < 32; ++kk) {
        const double x = vb[(4 * kk + lr) * LV + lc];
        g = mfma_f64(x, x, g);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) S.g[(lr + 4 * r) * LV + lc] = g[r];
      __builtin_amdgcn_wave_barrier();
      if (lane < 16) {
        double tr[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          asm volatile("" ::: "memory");  // G column by column, not all loads at once
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < j; ++k) s = fma(tr[k], S.g[k * LV + j], s);
          const double tj = S.tau[buf][j];
          tr[j] = lane < j ? -tj * s : (lane == j ? tj : 0.0);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) tb[lane * LV + j] = tr[j];
      }
    }
    // (2) W_J = Vb^T Q[:, J] for the wave's tile columns J >= b
    d4 w[2];
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      w[jj] = d4{0.0, 0.0, 0.0, 0.0};
      if ((jj ? J1 : J0) >= b) {
#pragma unroll
        for (int I = 0; I < 8; ++I)
          if (I >= b) {
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
              w[jj] = mfma_f64(vb[(16 * I + 4 * kk + lr) * LV + lc], q[jj][I][kk], w[jj]);
          }
      }
    }
    lds_barrier();  // Tb
    // (3) W'_J = -Tb W_J;  Q[I][J] += Vb[I rows] W'_J
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      if ((jj ? J1 : J0) >= b) {
        d4 wp = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) wp = mfma_f64(-tb[lc * LV + 4 * kk + lr], w[jj][kk], wp);
#pragma unroll
        for (int I = 0; I < 8; ++I)
          if (I >= b) {
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
              q[jj][I] = mfma_f64(vb[(16 * I + lc) * LV + 4 * kk + lr], wp[kk], q[jj][I]);
          }
      }
    }
  }
  // every reflector row was read before the last barrier
#pragma unroll
  for (int jj = 0; jj < 2; ++jj)
#pragma unroll
    for (int I = 0; I < 8; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        V[(size_t)(16 * I + lr + 4 * r) * NP + (NP - 1 - (16 * (jj ? J1 : J0) + lc))] = q[jj][I][r];
}



// Implicit QL with Wilkinson shifts on (d, e), one lane per zone (EISPACK
// tql2 ordering: eigenvalue l converges at the top of the block l..m).  The
// end m of the unreduced block is carried from one sweep to the next: the
// negligibility test of e_j is made as soon as e_j is final in the sweep, so
// a full rescan is only needed after a 1 x 1 deflation.  ABS = true (the
// mixed path, whose error bound is 1e-4): e_j is negligible when
// |e_j| <= tol = eps ||T|| with ||T|| = max_k (|d_k| + |e_k|), EISPACK tql2's
// absolute test.  T comes from a Householder reduction whose backward error
// is already of order eps ||M||, and every matrix solved here is I plus a
// positive semidefinite one, so the absolute test keeps the solve backward
// stable; it saves the iterations that a test relative to d_j spends on
// rounding-level e_j next to eigenvalues near 1 (the N - p unit eigenvalues
// of a zone with p < N observations most of all).  ABS = false (FP64): the
// relative test of LAPACK's dsteqr, |e_j| <= eps (|d_j| + |d_{j+1}|), except
// between two diagonal entries within 4 tol of 1, where it is tol: there the
// iterations only refine rounding noise inside the unit eigenvalue cluster
// (any basis of it gives the same analysis), and the rest of the spectrum is
// solved as dsteqr would.
template <int NP, int ZW, bool ABS>
__global__ __launch_bounds__(64) void k_tql(int nzc, int64_t z0, const int32_t* __restrict__ zorder, const int64_t* __restrict__ sel_ptr,
                                            double* __restrict__ dvec, double* __restrict__ evec,
                                            double2* __restrict__ rlog, int64_t cap,
                                            int32_t* __restrict__ status,
                                            unsigned* __restrict__ err) {
  // ZW zones per wave (lanes ZW..63 idle): the sweeps of different zones
  // differ in length and number, and a wave runs the longest of its lanes,
  // so fewer zones per wave waste less and spread the work over more SIMDs
  __shared__ double d[NP][ZW];
  __shared__ double e[NP][ZW];
  const int lane = threadIdx.x;
  if (lane >= ZW) return;
  const int zl = blockIdx.x * ZW + lane;
  int64_t zq = zl < nzc ? zone_at(zorder, z0 + zl) : 0;
  const bool own = zq >= 0;  // negative: solved by the other pipeline
  if (!own) zq = 0;
  const bool valid = zl < nzc && own && sel_ptr[zq + 1] > sel_ptr[zq];
  if (valid) {
    copy_batched<16>(0, NP, 1, [&](int k) { return dvec[(size_t)k * nzc + zl]; },
                     [&](int k, double x) { d[k][lane] = x; });
    copy_batched<16>(0, NP, 1, [&](int k) { return evec[(size_t)k * nzc + zl]; },
                     [&](int k, double x) { e[k][lane] = x; });
  }
  int st = 0;
  int64_t pos = 0;
  double2* lg = rlog + (valid ? rlog_off(zl, cap) : 0);
  const double eps = 2.220446049250313e-16;
  if (valid) {
    double tn = fabs(d[NP - 1][lane]);
    for (int k = 0; k < NP - 1; ++k) tn = fmax(tn, fabs(d[k][lane]) + fabs(e[k][lane]));
    const double tol = eps * tn;
    // the deflation threshold of an off-diagonal entry between dj and dj1
    auto thr = [&](double dj, double dj1) {
      if (ABS) return tol;
      const bool unit = fabs(dj - 1.0) <= 4.0 * tol && fabs(dj1 - 1.0) <= 4.0 * tol;
      return unit ? tol : eps * (fabs(dj) + fabs(dj1));
    };
    auto negl = [&](int j) { return fabs(e[j][lane]) <= thr(d[j][lane], d[j + 1][lane]); };
    auto scan = [&](int from) {
      int m = from;
      for (; m < NP - 1; ++m)
        if (negl(m)) break;
      return m;
    };
    int l = 0, iter = 0;
    int m = scan(0);
    for (;;) {
      if (m == l) {
        ++l;
        if (l >= NP - 1) break;
        iter = 0;
        m = scan(l);
        continue;
      }
      if (++iter > 60) { st = 2; break; }
      if (pos + 1 + (m - l) > cap) { st = 1; break; }
      const double el = e[l][lane], dl = d[l][lane];
      double g = (d[l + 1][lane] - dl) / (2.0 * el);
      const double rg = sqrt(fma(g, g, 1.0));
      g = d[m][lane] - dl + el / (g + (g >= 0 ? rg : -rg));
      double s = 1.0, c = 1.0, p = 0.0;
      const int64_t hdr = pos++;
      int i = m - 1;
      double ei = e[i][lane], di = d[i][lane], dip1 = d[i + 1][lane];
      double dnext2 = 0.0;  // new d_{i+2}
      int jmin = m;
      bool under = false;
      // wait for the loads above here: left to the compiler, the wait lands
      // in the loop header and runs on every step
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
      // Branch-free step: h = 0 exactly (the block splits at i + 1) means
      // f = g = 0, so the same formulas with 1/r taken as 1 give
      // e_{i+1} = 0 and d_{i+1} = d_{i+1} - p; the step then logs nothing
      // and ends the sweep.
      for (; i >= l && !under; --i) {
        // operands of the next step, loaded ahead (unconditionally: a load
        // under a branch is waited for at the join; unused when i == l)
        const int im = i > 0 ? i - 1 : 0;
        const double ein = e[im][lane], din = d[im][lane];
        const double f = s * ei, bb = c * ei;
        const double h = fma(f, f, g * g);
        under = h == 0.0;
        // r = sqrt(h), s' = f / r, c' = g / r; with g2 = d_{i+1} - p,
        // (d_i - g2) s' + 2 c' bb = X / r, X formed while 1/r is computed
        const double g2 = dip1 - p;
        const double X = fma(di - g2, f, 2.0 * g * bb);
        const double ri = rsqrt_acc(under ? 1.0 : h);
        const double r = h * ri;
        e[i + 1][lane] = r;
        s = f * ri;
        c = g * ri;
        const double rr = X * ri;
        p = s * rr;
        const double dn = g2 + p;
        d[i + 1][lane] = dn;
        g = fma(c, rr, -bb);
        lg[rlog_rec(pos)] = make_double2(c, s);
        pos += under ? 0 : 1;
        const bool small = fabs(r) <= thr(dn, dnext2);
        jmin = (!under & (i + 1 < m) & small) ? i + 1 : jmin;  // no short circuit: no branch
        dnext2 = dn;
        ei = ein;
        dip1 = di;
        di = din;
      }
      lg[rlog_rec(hdr)] = make_double2((double)l + 65536.0 * (double)m, (double)(pos - hdr - 1));
      if (under) {  // the loop stepped once past the split at i + 1
        e[m][lane] = 0.0;
        m = i + 2;
        continue;
      }
      const double dl2 = dl - p;
      d[l][lane] = dl2;
      e[l][lane] = g;
      e[m][lane] = 0.0;
      if (fabs(g) <= thr(dl2, dnext2)) {
        ++l;
        if (l >= NP - 1) break;
        iter = 0;
        m = jmin > l ? jmin : scan(l);
      } else {
        m = jmin;
      }
    }
    for (int k = 0; k < NP; ++k) dvec[(size_t)k * nzc + zl] = d[k][lane];
  }
  if (zl < nzc) status[zl] = valid ? (st | (int)(pos << 2)) : 0;
  if (st) atomicAdd(err + 1, 1u);
}

// ---------------------------------------------------------------------------
// U = Q Z: the logged rotations applied to the rows of Q held in LDS, one
// thread per row.  The header and records of the next chain are loaded
// while the current one is applied.
// ---------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(NP < 64 ? 64 : NP) void k_tql_apply(int64_t z0, const int32_t* __restrict__ zorder,
                                                                 const int64_t* __restrict__ sel_ptr,
                                                                 double* __restrict__ qmat,
                                                                 const double2* __restrict__ rlog,
                                                                 int64_t cap,
                                                                 const int32_t* __restrict__ status) {
  constexpr int LD = NP + 1;
  constexpr int BS = NP < 64 ? 64 : NP;
  constexpr int PAD = 8;  // room for the look-ahead reads below column 0 of row 0
  __shared__ double Zs[PAD + NP * LD];
  __shared__ double2 buf[2][NP + 12];
  double* Z = Zs + PAD;
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;  // the zone is solved by the other pipeline
  if (sel_ptr[z + 1] == sel_ptr[z]) return;
  const int stz = status[zl];
  if (stz & 3) return;  // failed zone: solved again by the fallback
  const int tid = threadIdx.x;
  double* Q = qmat + (size_t)zl * NP * NP;
  copy_batched<16>(tid, NP * NP, BS, [&](int e) { return Q[e]; },
                   [&](int e, double x) { Z[(e / NP) * LD + e % NP] = x; });
  const int64_t len = stz >> 2;
  const double2* lg = rlog + rlog_off(zl, cap);
  if (tid < NP && tid < len) buf[0][tid] = lg[rlog_rec(tid)];
  __syncthreads();
  int64_t pos = 0;
  int b = 0;
  while (pos < len) {
    const double2 h = buf[b][0];
    const int m = __builtin_amdgcn_readfirstlane(((int)h.x) >> 16);
    const int cnt = __builtin_amdgcn_readfirstlane((int)h.y);
    const int64_t nxt = pos + 1 + cnt;
    const bool pf = tid < NP && nxt + tid < len;
    double2 pv = make_double2(0.0, 0.0);
    if (pf) pv = lg[rlog_rec(nxt + tid)];
    if (tid < NP) {
      // column i of this row is p[0]; rotations i = m-1, m-2, ... with
      // (c, s) = CS(0), CS(1), ...  Groups of 4 rotations in two register
      // sets; the operands of the next group are read while the current one
      // is applied.  The look-ahead reads reach at most 4 columns below the
      // chain (into the row before, or PAD) and 11 records past it (buf).
      typedef __attribute__((address_space(3))) const double lds_d;
      double* p = Z + tid * LD + (m - 1);
      lds_d* cd = (lds_d*)(buf[b] + 1);
      asm volatile("" : "+v"(cd));  // one VGPR base for the broadcast reads
      double carry = p[1];
      int k = 0;
#define CS(J) make_double2(cd[2 * (J)], cd[2 * (J) + 1])
#define OAK_ROT(Q, ZV, DST)                  \
  DST = fma(Q.y, ZV, Q.x * carry);           \
  carry = fma(Q.x, ZV, -Q.y * carry);
      if (cnt >= 8) {
        double a0 = p[0], a1 = p[-1], a2 = p[-2], a3 = p[-3];
        double2 qa0 = CS(0), qa1 = CS(1), qa2 = CS(2), qa3 = CS(3);
        for (; k + 8 <= cnt; k += 8) {
          const double b0 = p[-4], b1 = p[-5], b2 = p[-6], b3 = p[-7];
          const double2 qb0 = CS(4), qb1 = CS(5), qb2 = CS(6), qb3 = CS(7);
          OAK_ROT(qa0, a0, p[1])
          OAK_ROT(qa1, a1, p[0])
          OAK_ROT(qa2, a2, p[-1])
          OAK_ROT(qa3, a3, p[-2])
          a0 = p[-8]; a1 = p[-9]; a2 = p[-10]; a3 = p[-11];
          qa0 = CS(8); qa1 = CS(9); qa2 = CS(10); qa3 = CS(11);
          OAK_ROT(qb0, b0, p[-3])
          OAK_ROT(qb1, b1, p[-4])
          OAK_ROT(qb2, b2, p[-5])
          OAK_ROT(qb3, b3, p[-6])
          p -= 8;
          cd += 16;
        }
      }
      for (; k < cnt; ++k) {
        const double zv = p[0];
        const double2 q = CS(0);
        OAK_ROT(q, zv, p[1])
        p -= 1;
        cd += 2;
      }
#undef OAK_ROT
#undef CS
      p[1] = carry;  // column m - cnt
    }
    if (pf) buf[b ^ 1][tid] = pv;
    __syncthreads();
    pos = nxt;
    b ^= 1;
  }
  for (int e = tid; e < NP * NP; e += BS) Q[e] = Z[(e / NP) * LD + e % NP];
}

// ---------------------------------------------------------------------------
// U = Q Z for NP = 64 with the rows in registers: thread r holds row r of Q
// whole (64 values), so a chain needs no other thread; it runs in unrolled,
// branch-free segments of 16 positions with the identity (1, 0) outside the
// chain, the (c, s) of each position from a table in LDS (the scheme of
// k_tql_apply128's groups).  The operations on a row are those of
// k_tql_apply.  One wave per zone.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_tql_apply64(int64_t z0, const int32_t* __restrict__ zorder,
                                                     const int64_t* __restrict__ sel_ptr,
                                                     double* __restrict__ qmat,
                                                     const double2* __restrict__ rlog,
                                                     int64_t cap,
                                                     const int32_t* __restrict__ status) {
  constexpr int NP = 64;
  __shared__ double2 buf[2][NP + 12];
  __shared__ double2 pr[NP];
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;  // the zone is solved by the other pipeline
  if (sel_ptr[z + 1] == sel_ptr[z]) return;
  const int stz = status[zl];
  if (stz & 3) return;  // failed zone: solved again by the fallback
  const int tid = threadIdx.x;
  double* Qr = qmat + (size_t)zl * NP * NP + (size_t)tid * NP;
  double x[NP];
  {
    const double2* q2 = reinterpret_cast<const double2*>(Qr);
#pragma unroll
    for (int c = 0; c < NP / 2; ++c) {
      const double2 v = q2[c];
      x[2 * c] = v.x;
      x[2 * c + 1] = v.y;
    }
  }
  const int64_t len = stz >> 2;
  const double2* lg = rlog + rlog_off(zl, cap);
  if (tid < len) buf[0][tid] = lg[rlog_rec(tid)];
  __syncthreads();
  int64_t pos = 0;
  int b = 0;
  while (pos < len) {
    const double2 hd = buf[b][0];
    const int m = __builtin_amdgcn_readfirstlane(((int)hd.x) >> 16);
    const int cnt = __builtin_amdgcn_readfirstlane((int)hd.y);
    const int l = m - cnt;
    const int64_t nxt = pos + 1 + cnt;
    const bool pf = nxt + tid < len;
    double2 pv = make_double2(0.0, 0.0);
    if (pf) pv = lg[rlog_rec(nxt + tid)];
    // positions m-1 .. l: rotation of columns P and P+1 with the record of
    // position P
    pr[tid] = (tid >= l && tid <= m - 1) ? buf[b][1 + (m - 1 - tid)] : make_double2(1.0, 0.0);
    auto segment = [&](auto top) {
      constexpr int T = decltype(top)::value;
      if (m - 1 < T - 15 || l > T) return;
      double2 qv[16];
#pragma unroll
      for (int t = 0; t < 16; ++t)
        if (T - t >= 0) qv[t] = pr[T - t];
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const int J = T - t;
        if (J >= 0) {
          const double hh = x[J + 1], zj = x[J];
          x[J + 1] = fma(qv[t].y, zj, qv[t].x * hh);
          x[J] = fma(qv[t].x, zj, -qv[t].y * hh);
        }
      }
    };
    segment(std::integral_constant<int, 62>{});
    segment(std::integral_constant<int, 46>{});
    segment(std::integral_constant<int, 30>{});
    segment(std::integral_constant<int, 14>{});
    if (pf) buf[b ^ 1][tid] = pv;
    __syncthreads();
    pos = nxt;
    b ^= 1;
  }
  {
    double2* q2 = reinterpret_cast<double2*>(Qr);
#pragma unroll
    for (int c = 0; c < NP / 2; ++c) q2[c] = make_double2(x[2 * c], x[2 * c + 1]);
  }
}

// ---------------------------------------------------------------------------
// U = Q Z for NP = 256 with the rows in registers, the scheme of
// k_tql_apply128 (below) at twice the width: workgroup y takes rows
// 128 y .. 128 y + 127, row r's columns 0..127 in thread r (waves 0-1, the
// lower group) and columns 128..255 in thread 128 + r (waves 2-3, the upper
// group), 128 values a thread.  A chain acts on the upper columns
// (positions >= 128), couples columns 127 and 128 (position 127), then acts
// on the lower columns; the groups run one chain apart, each in unrolled,
// branch-free segments of 16 positions with the identity outside the chain.
// One workgroup per CU (the row halves take 256 registers a lane), every SIMD
// issuing.
// ---------------------------------------------------------------------------
struct Apply256Lds {
  double2 buf[3][256 + 12];  // records of chains t+1, t, t-1
  double2 pr[4][128];        // per wave: (c, s) of its group's positions
  double xc[2][2][128];      // [step parity][0: column 127, 1: column 128] per row
};

__global__ __launch_bounds__(256) void k_tql_apply256(int64_t z0, const int32_t* __restrict__ zorder,
                                                       const int64_t* __restrict__ sel_ptr,
                                                       double* __restrict__ qmat,
                                                       const double2* __restrict__ rlog,
                                                       int64_t cap,
                                                       const int32_t* __restrict__ status,
                                                       unsigned* __restrict__ err) {
  constexpr int NP = 256, H = 128;
  __shared__ Apply256Lds S;
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;
  if (sel_ptr[z + 1] == sel_ptr[z]) return;
  const int tid = threadIdx.x;
  const int stz = status[zl];
  if (stz & 3) {
    if (blockIdx.y == 0 && tid == 0) atomicOr(err, OAKG_DERR_NOCONV);
    return;
  }
  const int g = tid >> 7, r = tid & 127, w = tid >> 6, lane = tid & 63;
  double* Qr = qmat + (size_t)zl * NP * NP + (size_t)(H * blockIdx.y + r) * NP + H * g;
  double x[H];  // row H y + r, columns H g .. H g + H - 1
  {
    const double2* q2 = reinterpret_cast<const double2*>(Qr);
#pragma unroll
    for (int c = 0; c < H / 2; ++c) {
      const double2 v = q2[c];
      x[2 * c] = v.x;
      x[2 * c + 1] = v.y;
    }
  }
  const int64_t len = stz >> 2;
  const double2* lg = rlog + rlog_off(zl, cap);
  if (tid < len) S.buf[0][tid] = lg[rlog_rec(tid)];
  __syncthreads();
  int64_t pos = 0;
  int bt = 0, bp = 2, par = 0;
  bool has_prev = false;
  int m_prev = 0, l_prev = 0;
  double2* pr = S.pr[w];
  while (true) {
    const bool has_cur = pos < len;
    if (!has_cur && !has_prev) break;
    int m = 0, l = 0;
    int64_t nxt = pos;
    if (has_cur) {
      const double2 hd = S.buf[bt][0];
      m = __builtin_amdgcn_readfirstlane(((int)hd.x) >> 16);
      const int cnt = __builtin_amdgcn_readfirstlane((int)hd.y);
      l = m - cnt;
      nxt = pos + 1 + cnt;
    }
    const int bn = bt == 2 ? 0 : bt + 1;
    const bool pf = has_cur && nxt + tid < len;
    double2 pv = make_double2(0.0, 0.0);
    if (pf) pv = lg[rlog_rec(nxt + tid)];
    // coupling rotation (position 127) of chain t-1
    if (has_prev && l_prev <= H - 1 && m_prev - 1 >= H - 1) {
      const double2 q = S.buf[bp][1 + (m_prev - 1) - (H - 1)];
      const double chi = g == 1 ? x[0] : S.xc[par ^ 1][1][r];
      const double clo = g == 0 ? x[H - 1] : S.xc[par ^ 1][0][r];
      const double nhi = fma(q.y, clo, q.x * chi);
      const double nlo = fma(q.x, clo, -q.y * chi);
      if (g == 1) x[0] = nhi;
      else x[H - 1] = nlo;
    }
    if (g == 1) {
      if (has_cur && m - 1 >= H) {
        // chain t on columns 128..255: positions 254..128 in eight unrolled
        // segments of 16, the identity (1, 0) outside l..m-1
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int P = H + 64 * u + lane;
          pr[64 * u + lane] =
              (P >= l && P <= m - 1) ? S.buf[bt][1 + (m - 1 - P)] : make_double2(1.0, 0.0);
        }
        auto segment = [&](auto top) {
          constexpr int T = decltype(top)::value;
          if (m - 1 < T - 15 || l > T) return;
          double2 qv[16];
#pragma unroll
          for (int t = 0; t < 16; ++t)
            if (T - t >= H) qv[t] = pr[T - t - H];
#pragma unroll
          for (int t = 0; t < 16; ++t) {
            const int J = T - t;
            if (J >= H) {
              const double hh = x[J + 1 - H], zj = x[J - H];
              x[J + 1 - H] = fma(qv[t].y, zj, qv[t].x * hh);
              x[J - H] = fma(qv[t].x, zj, -qv[t].y * hh);
            }
          }
        };
        segment(std::integral_constant<int, 254>{});
        segment(std::integral_constant<int, 238>{});
        segment(std::integral_constant<int, 222>{});
        segment(std::integral_constant<int, 206>{});
        segment(std::integral_constant<int, 190>{});
        segment(std::integral_constant<int, 174>{});
        segment(std::integral_constant<int, 158>{});
        segment(std::integral_constant<int, 142>{});
      }
      S.xc[par][1][r] = x[0];
    } else {
      if (has_prev && l_prev <= H - 2) {
        // chain t-1 on columns 0..127: positions 126..l
        const int mp = m_prev, lp = l_prev;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int P = 64 * u + lane;
          pr[P] = (P >= lp && P <= mp - 1 && P <= H - 2) ? S.buf[bp][1 + (mp - 1 - P)]
                                                         : make_double2(1.0, 0.0);
        }
        auto segment_lo = [&](auto top) {
          constexpr int T = decltype(top)::value;
          if (mp - 1 < T - 15 || lp > T) return;
          double2 qv[16];
#pragma unroll
          for (int t = 0; t < 16; ++t)
            if (T - t >= 0) qv[t] = pr[T - t];
#pragma unroll
          for (int t = 0; t < 16; ++t) {
            const int J = T - t;
            if (J >= 0) {
              const double hh = x[J + 1], zj = x[J];
              x[J + 1] = fma(qv[t].y, zj, qv[t].x * hh);
              x[J] = fma(qv[t].x, zj, -qv[t].y * hh);
            }
          }
        };
        segment_lo(std::integral_constant<int, 126>{});
        segment_lo(std::integral_constant<int, 110>{});
        segment_lo(std::integral_constant<int, 94>{});
        segment_lo(std::integral_constant<int, 78>{});
        segment_lo(std::integral_constant<int, 62>{});
        segment_lo(std::integral_constant<int, 46>{});
        segment_lo(std::integral_constant<int, 30>{});
        segment_lo(std::integral_constant<int, 14>{});
      }
      S.xc[par][0][r] = x[H - 1];
    }
    if (pf) S.buf[bn][tid] = pv;
    __syncthreads();
    has_prev = has_cur;
    m_prev = m;
    l_prev = l;
    bp = bt;
    bt = bn;
    pos = nxt;
    par ^= 1;
  }
  {
    double2* q2 = reinterpret_cast<double2*>(Qr);
#pragma unroll
    for (int c = 0; c < H / 2; ++c) q2[c] = make_double2(x[2 * c], x[2 * c + 1]);
  }
}

// ---------------------------------------------------------------------------
// U = Q Z for NP = 128 with the rows in registers.  A chain (one QL sweep,
// from its first position m-1 down to l) runs as unrolled, branch-free
// segments of 16 positions, each segment it meets with the identity at the
// positions it does not cover (a jump into the middle of an unrolled sequence
// is not structured control flow, and the compiler's rewriting of it copies
// the registers around); the (c, s) of a segment come from a per-wave,
// position-indexed table in LDS.  The arithmetic of every rotation is that of
// k_tql_apply.  The LDS of the rotation phase is reused by the load of Q and
// by the epilogue, to stage a half of U for its column sums.
// ---------------------------------------------------------------------------
// k_tql_apply128 (N > 64): the logged rotations applied to the rows of Q,
// then the epilogue of k_eigvec_epi.  Each row of Q is split over two
// threads, row r's columns 0..63 in thread r (waves 0-1, the lower group) and
// columns 64..127 in thread 128 + r (waves 2-3, the upper group): 64 values a
// thread, all in VGPRs.  A chain of rotations (positions m-1 down to l) acts
// on the upper columns (positions >= 64), couples columns 63 and 64 (position
// 63), then acts on the lower columns (positions <= 62).  The groups run one
// chain apart: in step t the upper group applies chain t while the lower
// group applies chain t-1, both after the coupling rotation of chain t-1,
// which each group computes from its own column and the other group's column
// published at the end of step t-1 (same inputs and operations in both
// groups, so the two copies agree bit for bit).  The records of chains t+1,
// t and t-1 are in three buffers.
struct Apply128Lds {
  union {
    double z[8 + 128 * 65];  // a staged half of U, row stride 65 (load and epilogue)
    struct {
      double2 buf[3][128 + 12];  // records of chains t+1, t, t-1
      double2 pr[4][64];         // per wave: (c, s) of its group's positions
      double xc[2][2][128];      // [step parity][0: column 63, 1: column 64] per row
    } r;
  };
  double ev[5][128];    // epilogue vectors
  double part[2][4][64];  // partial column sums of four row quarters
  double hd[2][2][128];   // [0: x, 1: y][group] partial row products
};

__global__ __launch_bounds__(256) void k_tql_apply128(int64_t z0, const int32_t* __restrict__ zorder, int N, const int64_t* __restrict__ sel_ptr,
                                                       const double* __restrict__ mat,
                                                       double* __restrict__ qmat,
                                                       const double2* __restrict__ rlog,
                                                       int64_t cap,
                                                       const int32_t* __restrict__ status,
                                                       const double* __restrict__ dvec, int nzc,
                                                       double* __restrict__ vec,
                                                       double* __restrict__ ampl,
                                                       unsigned* __restrict__ err, int refine) {
  constexpr int NP = 128, H = 64, LD = 65;
  __shared__ Apply128Lds S;
  double* Z = S.z + 8;
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;  // the zone is solved by the other pipeline
  const int tid = threadIdx.x;
  const int g = tid >> 7, r = tid & 127, w = tid >> 6, lane = tid & 63;
  if (sel_ptr[z + 1] == sel_ptr[z]) {  // no observations: a = 0
    if (tid < NP) vec[(size_t)zl * 4 * NP + NP + tid] = 0.0;
    return;
  }
  const int stz = status[zl];
  if (stz & 3) return;  // failed zone: solved again by the fallback
  double* Q = qmat + (size_t)zl * NP * NP;
  double x[H];  // row r, columns 64 g .. 64 g + 63
  copy_batched<16>(tid, NP * H, 256, [&](int e) { return Q[(e >> 6) * NP + H + (e & 63)]; },
                   [&](int e, double v) { Z[(e >> 6) * LD + (e & 63)] = v; });
  __syncthreads();
  if (g == 1) {
#pragma unroll
    for (int c = 0; c < H; ++c) x[c] = Z[r * LD + c];
  }
  __syncthreads();
  copy_batched<16>(tid, NP * H, 256, [&](int e) { return Q[(e >> 6) * NP + (e & 63)]; },
                   [&](int e, double v) { Z[(e >> 6) * LD + (e & 63)] = v; });
  __syncthreads();
  if (g == 0) {
#pragma unroll
    for (int c = 0; c < H; ++c) x[c] = Z[r * LD + c];
  }
  __syncthreads();  // Z is now reused by the record buffers
  const int64_t len = stz >> 2;
  const double2* lg = rlog + rlog_off(zl, cap);
  if (tid < len && tid < NP) S.r.buf[0][tid] = lg[rlog_rec(tid)];
  __syncthreads();
  int64_t pos = 0;
  int bt = 0, bp = 2, par = 0;
  bool has_prev = false;
  int m_prev = 0, l_prev = 0;
  double2* pr = S.r.pr[w];
  while (true) {
    const bool has_cur = pos < len;
    if (!has_cur && !has_prev) break;
    int m = 0, l = 0;
    int64_t nxt = pos;
    if (has_cur) {
      const double2 hd = S.r.buf[bt][0];
      m = __builtin_amdgcn_readfirstlane(((int)hd.x) >> 16);
      const int cnt = __builtin_amdgcn_readfirstlane((int)hd.y);
      l = m - cnt;
      nxt = pos + 1 + cnt;
    }
    const int bn = bt == 2 ? 0 : bt + 1;
    const bool pf = has_cur && tid < NP && nxt + tid < len;
    double2 pv = make_double2(0.0, 0.0);
    if (pf) pv = lg[rlog_rec(nxt + tid)];
    // coupling rotation (position 63) of chain t-1
    if (has_prev && l_prev <= H - 1 && m_prev - 1 >= H - 1) {
      const double2 q = S.r.buf[bp][1 + (m_prev - 1) - (H - 1)];
      const double c64 = g == 1 ? x[0] : S.r.xc[par ^ 1][1][r];
      const double c63 = g == 0 ? x[H - 1] : S.r.xc[par ^ 1][0][r];
      const double n64 = fma(q.y, c63, q.x * c64);
      const double n63 = fma(q.x, c63, -q.y * c64);
      if (g == 1) x[0] = n64;
      else x[H - 1] = n63;
    }
    if (g == 1) {
      if (has_cur && m - 1 >= H) {
        // chain t on columns 64..127: positions 126..64 in four unrolled
        // segments of 16 with the identity (1, 0) outside l..m-1
        {
          const int P = H + lane;
          pr[lane] = (P >= l && P <= m - 1) ? S.r.buf[bt][1 + (m - 1 - P)] : make_double2(1.0, 0.0);
        }
        auto segment = [&](auto top) {
          constexpr int T = decltype(top)::value;
          if (m - 1 < T - 15 || l > T) return;
          double2 qv[16];
#pragma unroll
          for (int t = 0; t < 16; ++t)
            if (T - t >= H) qv[t] = pr[T - t - H];
#pragma unroll
          for (int t = 0; t < 16; ++t) {
            const int J = T - t;
            if (J >= H) {
              const double hh = x[J + 1 - H], zj = x[J - H];
              x[J + 1 - H] = fma(qv[t].y, zj, qv[t].x * hh);
              x[J - H] = fma(qv[t].x, zj, -qv[t].y * hh);
            }
          }
        };
        segment(std::integral_constant<int, 126>{});
        segment(std::integral_constant<int, 110>{});
        segment(std::integral_constant<int, 94>{});
        segment(std::integral_constant<int, 78>{});
      }
      S.r.xc[par][1][r] = x[0];
    } else {
      if (has_prev && l_prev <= H - 2) {
        // chain t-1 on columns 0..63: positions 62..l
        const int mp = m_prev, lp = l_prev;
        pr[lane] = (lane >= lp && lane <= mp - 1 && lane <= H - 2) ? S.r.buf[bp][1 + (mp - 1 - lane)]
                                                                   : make_double2(1.0, 0.0);
        auto segment_lo = [&](auto top) {
          constexpr int T = decltype(top)::value;
          if (mp - 1 < T - 15 || lp > T) return;
          double2 qv[16];
#pragma unroll
          for (int t = 0; t < 16; ++t)
            if (T - t >= 0) qv[t] = pr[T - t];
#pragma unroll
          for (int t = 0; t < 16; ++t) {
            const int J = T - t;
            if (J >= 0) {
              const double hh = x[J + 1], zj = x[J];
              x[J + 1] = fma(qv[t].y, zj, qv[t].x * hh);
              x[J] = fma(qv[t].x, zj, -qv[t].y * hh);
            }
          }
        };
        segment_lo(std::integral_constant<int, 62>{});
        segment_lo(std::integral_constant<int, 46>{});
        segment_lo(std::integral_constant<int, 30>{});
        segment_lo(std::integral_constant<int, 14>{});
      }
      S.r.xc[par][0][r] = x[H - 1];
    }
    if (pf) S.r.buf[bn][tid] = pv;
    __syncthreads();
    has_prev = has_cur;
    m_prev = m;
    l_prev = l;
    bp = bt;
    bt = bn;
    pos = nxt;
    par ^= 1;
  }
  // ---- epilogue: the quantities of k_eigvec_epi from the half rows in hand ----
  double* mu = S.ev[0];  // mu_j = max(lambda_j, 1) of column j
  double* bv = S.ev[1];  // b
  double* xv = S.ev[2];  // U^T b / mu, then U^T (b - M a0) / mu
  double* yv = S.ev[3];  // U^T 1_N sqrt(mu), then mu^-1/4
  double* av = S.ev[4];  // a0, then b - M a0, then vr^2
  double* V = vec + (size_t)zl * 4 * NP;
  if (tid < NP) {
    mu[tid] = fmax(dvec[(size_t)tid * nzc + zl], 1.0);
    bv[tid] = V[tid];
  }
  const int cc = tid & 63, q4 = tid >> 6;
  auto stage = [&](int grp) {
    if (g == grp) {
#pragma unroll
      for (int c = 0; c < H; ++c) Z[r * LD + c] = x[c];
    }
    __syncthreads();
  };
  // U^T b and U^T 1_N over the staged half (column offset co), rows in quarters
  auto colsum2 = [&](int co) {
    double s1[4] = {0, 0, 0, 0}, s2[4] = {0, 0, 0, 0};
#pragma unroll 2
    for (int r0 = 32 * q4; r0 < 32 * q4 + 32; r0 += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int rr = r0 + u;
        const double zz = Z[rr * LD + cc];
        s1[u] = fma(zz, bv[rr], s1[u]);
        s2[u] += rr < N ? zz : 0.0;
      }
    }
    S.part[0][q4][cc] = (s1[0] + s1[1]) + (s1[2] + s1[3]);
    S.part[1][q4][cc] = (s2[0] + s2[1]) + (s2[2] + s2[3]);
    __syncthreads();
    if (tid < 64) {
      const double c1 = (S.part[0][0][tid] + S.part[0][1][tid]) + (S.part[0][2][tid] + S.part[0][3][tid]);
      const double c2 = (S.part[1][0][tid] + S.part[1][1][tid]) + (S.part[1][2][tid] + S.part[1][3][tid]);
      xv[co + tid] = c1 / mu[co + tid];
      yv[co + tid] = c2 * sqrt(mu[co + tid]);
    }
    __syncthreads();
  };
  auto colsum1 = [&](int co) {
    double s1[4] = {0, 0, 0, 0};
#pragma unroll 2
    for (int r0 = 32 * q4; r0 < 32 * q4 + 32; r0 += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) s1[u] = fma(Z[(r0 + u) * LD + cc], av[r0 + u], s1[u]);
    }
    S.part[0][q4][cc] = (s1[0] + s1[1]) + (s1[2] + s1[3]);
    __syncthreads();
    if (tid < 64)
      xv[co + tid] = ((S.part[0][0][tid] + S.part[0][1][tid]) + (S.part[0][2][tid] + S.part[0][3][tid])) /
                     mu[co + tid];
    __syncthreads();
  };
  // half-row products (U x)_r over the group's columns, into hd[.][g][r]
  auto rowdot_h = [&](const double* xa, const double* ya) {
    double tx = 0, ty = 0;
#pragma unroll
    for (int c0 = 0; c0 < H; c0 += 16) {
      double2 xx[8], yy[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        xx[u] = reinterpret_cast<const double2*>(xa + H * g + c0)[u];
        if (ya) yy[u] = reinterpret_cast<const double2*>(ya + H * g + c0)[u];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        tx = fma(x[c0 + 2 * u], xx[u].x, tx);
        tx = fma(x[c0 + 2 * u + 1], xx[u].y, tx);
        if (ya) {
          ty = fma(x[c0 + 2 * u], yy[u].x, ty);
          ty = fma(x[c0 + 2 * u + 1], yy[u].y, ty);
        }
      }
    }
    S.hd[0][g][r] = tx;
    if (ya) S.hd[1][g][r] = ty;
  };
  stage(0);
  colsum2(0);
  stage(1);
  colsum2(H);  // the upper half stays staged
  rowdot_h(xv, yv);
  __syncthreads();
  double a0 = 0, vr = 0;
  if (tid < NP) {
    a0 = S.hd[0][0][tid] + S.hd[0][1][tid];
    vr = S.hd[1][0][tid] + S.hd[1][1][tid];
    av[tid] = a0;
  }
  __syncthreads();
  // refinement a = a0 + U diag(1/mu) U^T (b - M a0) (FP64 path; the mixed
  // path's bound of 1e-4 is far above the cross-talk it removes); M
  // symmetric: row r of M a0 from column r, half of the columns per group
  // (32 loads in flight)
  double ar = a0;
  if (refine) {
    {
      const double* M = mat + (size_t)zl * NP * NP;
      double mq[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int c0 = 0; c0 < H; c0 += 32) {
        double mv[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) mv[u] = M[(size_t)(H * g + c0 + u) * NP + r];
#pragma unroll
        for (int u = 0; u < 32; ++u) mq[u & 7] = fma(mv[u], av[H * g + c0 + u], mq[u & 7]);
      }
      S.hd[0][g][r] = ((mq[0] + mq[1]) + (mq[2] + mq[3])) + ((mq[4] + mq[5]) + (mq[6] + mq[7]));
    }
    __syncthreads();
    if (tid < NP) av[tid] = bv[tid] - (S.hd[0][0][tid] + S.hd[0][1][tid]);
    __syncthreads();
    colsum1(H);  // upper half still staged
    stage(0);
    colsum1(0);
    rowdot_h(xv, nullptr);
    __syncthreads();
    if (tid < NP) ar = a0 + (S.hd[0][0][tid] + S.hd[0][1][tid]);
  } else {
    stage(0);  // the lower half, for Y below
  }
  if (tid < NP) {
    av[tid] = vr * vr;
    yv[tid] = 1.0 / sqrt(sqrt(mu[tid]));
  }
  __syncthreads();
  if (tid < NP) {
    double nq[4] = {0, 0, 0, 0};
#pragma unroll 8
    for (int q = 0; q < NP; q += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) nq[u] += av[q + u];
    }
    const double nv = (nq[0] + nq[1]) + (nq[2] + nq[3]);
    const double vn = vr / sqrt(nv);
    V[NP + tid] = ar;
    V[2 * NP + tid] = vn;
    V[3 * NP + tid] = mu[tid];
    if (ampl && tid < N) ampl[(size_t)z * N + tid] = ar;
    if (ar != ar) atomicOr(err, OAKG_DERR_NAN);
  }
  // Y = U diag(mu^-1/4): the lower half is staged, then the upper half
  for (int e = tid; e < NP * H; e += 256)
    Q[(e >> 6) * NP + (e & 63)] = Z[(e >> 6) * LD + (e & 63)] * yv[e & 63];
  __syncthreads();
  if (g == 1) {
#pragma unroll
    for (int c = 0; c < H; ++c) Z[r * LD + c] = x[c] * yv[H + c];
  }
  __syncthreads();
  for (int e = tid; e < NP * H; e += 256) Q[(e >> 6) * NP + H + (e & 63)] = Z[(e >> 6) * LD + (e & 63)];
}

// ---------------------------------------------------------------------------
// Epilogue from eigenvectors U (rows in global memory) and eigenvalues d.
// ---------------------------------------------------------------------------
template <int NP, int BS = 256>
__global__ __launch_bounds__(BS) void k_eigvec_epi(int64_t z0, const int32_t* __restrict__ zorder, int N, const int64_t* __restrict__ sel_ptr,
                                                    const double* __restrict__ mat,
                                                    double* __restrict__ umat,
                                                    const double* __restrict__ dvec, int nzc,
                                                    double* __restrict__ vec,
                                                    double* __restrict__ ampl,
                                                    unsigned* __restrict__ err) {
  constexpr int H = BS / NP;  // threads per column / row
  __shared__ double col[6][NP];
  __shared__ double part[H][NP];
  const int zl = blockIdx.x;
  const int64_t z = zone_at(zorder, z0 + zl);
  if (z < 0) return;  // the zone is solved by the other pipeline
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
    lds_barrier();
    double t = 0;
    for (int q = 0; q < H; ++q) t += part[q][j];
    lds_barrier();
    return t;
  };
  auto rowsum = [&](int ci) -> double {  // sum_c U[j][c] * col[ci][c]
    double s = 0;
    for (int c = hh * SEG; c < hh * SEG + SEG; ++c) s = fma(U[j * NP + c], col[ci][c], s);
    part[hh][j] = s;
    lds_barrier();
    double t = 0;
    for (int q = 0; q < H; ++q) t += part[q][j];
    lds_barrier();
    return t;
  };
  const double lam = dvec[(size_t)j * nzc + zl];  // eigenvalue of I + C for column j
  const double muc = fmax(lam, 1.0);
  if (hh == 0) {
    col[0][j] = V[j];                  // b
    col[1][j] = j < N ? 1.0 : 0.0;     // 1_N
  }
  lds_barrier();
  const double c1 = colsum(0);
  const double c2 = colsum(1);
  if (hh == 0) {
    col[2][j] = c1 / muc;
    col[3][j] = c2 * sqrt(muc);
  }
  lds_barrier();
  const double a0 = rowsum(2);
  const double vr = rowsum(3);
  // refinement a += U diag(1/mu) U^T (b - M a)
  if (hh == 0) col[4][j] = a0;
  lds_barrier();
  double ma = 0;
  for (int c = hh * SEG; c < hh * SEG + SEG; ++c) ma = fma(M[j * NP + c], col[4][c], ma);
  part[hh][j] = ma;
  lds_barrier();
  double mat_a = 0;
  for (int q = 0; q < H; ++q) mat_a += part[q][j];
  lds_barrier();
  if (hh == 0) col[5][j] = col[0][j] - mat_a;
  lds_barrier();
  const double c3 = colsum(5);
  if (hh == 0) col[2][j] = c3 / muc;
  lds_barrier();
  const double ar = a0 + rowsum(2);
  if (hh == 0) col[4][j] = vr * vr;
  lds_barrier();
  double nv = 0;
  for (int q = 0; q < NP; ++q) nv += col[4][q];
  const double vn = vr / sqrt(nv);
  if (hh == 0) col[5][j] = 1.0 / sqrt(sqrt(muc));
  lds_barrier();
  for (int e = tid; e < NP * NP; e += BS) U[e] = U[e] * col[5][e % NP];
  if (hh == 0) {
    V[NP + j] = ar;
    V[2 * NP + j] = vn;
    V[3 * NP + j] = muc;
    if (ampl && j < N) ampl[(size_t)z * N + j] = ar;
    if (ar != ar) atomicOr(err, OAKG_DERR_NAN);
  }
}

// ---------------------------------------------------------------------------
template <int NP>
static void tql_pipeline_t(const ChunkArgs& A, hipStream_t st) {
  int split = 0;
  if constexpr (NP == 128) split = A.tri_split;
  if (split == 2) {
    hipLaunchKernelGGL((k_tridiag<128, 3>), dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                       A.mat, A.ymat, A.dvec, A.evec, A.nzc, (int64_t)NP * NP, 1);
    hipLaunchKernelGGL((k_tridiag<128, 4>), dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                       A.mat, A.ymat, A.dvec, A.evec, A.nzc, (int64_t)NP * NP, 1);
    hipLaunchKernelGGL((k_tridiag<128, 2>), dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                       A.mat, A.ymat, A.dvec, A.evec, A.nzc, (int64_t)NP * NP, 1);
  } else if (split == 4) {
    // steps 0..63 on the lower triangle only, then the trailing block by the
    // lower-triangle kernel at NP = 64, one wave per zone (wide256.hip)
    launch_tridiag_lower128(A, st);
    launch_tridiag_lower128_tail(A, st);
  } else if (split == 3) {
    // steps 0..63 on the lower triangle only (wide256.hip), then as split 1
    launch_tridiag_lower128(A, st);
    hipLaunchKernelGGL((k_tridiag<128, 2>), dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                       A.mat, A.ymat, A.dvec, A.evec, A.nzc, (int64_t)NP * NP, 1);
  } else if (split == 1) {
    hipLaunchKernelGGL((k_tridiag<128, 1>), dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                       A.mat, A.ymat, A.dvec, A.evec, A.nzc, (int64_t)NP * NP, 1);
    hipLaunchKernelGGL((k_tridiag<128, 2>), dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                       A.mat, A.ymat, A.dvec, A.evec, A.nzc, (int64_t)NP * NP, 1);
  } else if (NP == 64) {
    // the lower-triangle reduction on one wave per zone, Q on MFMA
    launch_tridiag_lower64(A, st);
  } else {
    hipLaunchKernelGGL((k_tridiag<NP, 0>), dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                       A.mat, A.ymat, A.dvec, A.evec, A.nzc, (int64_t)NP * NP, 1);
  }
  if constexpr (NP == 128)
    hipLaunchKernelGGL(k_formq128, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr, A.ymat);
  // zones per QL wave: a tuning knob (the rotation log layout groups 64
  // zones whatever the wave width)
  // (the deflation test: absolute for the mixed path, relative outside the
  // unit eigenvalue cluster for FP64)
#define OAK_TQL(ZW)                                                                                       \
  do {                                                                                                    \
    if (A.prec == 1)                                                                                      \
      hipLaunchKernelGGL((k_tql<NP, ZW, true>), dim3((A.nzc + ZW - 1) / ZW), dim3(64), 0, st, A.nzc,     \
                         A.z0, A.zorder, A.sel_ptr, A.dvec, A.evec, (double2*)A.rlog, A.rlog_cap,         \
                         A.tql_status, A.err);                                                            \
    else                                                                                                  \
      hipLaunchKernelGGL((k_tql<NP, ZW, false>), dim3((A.nzc + ZW - 1) / ZW), dim3(64), 0, st, A.nzc,    \
                         A.z0, A.zorder, A.sel_ptr, A.dvec, A.evec, (double2*)A.rlog, A.rlog_cap,         \
                         A.tql_status, A.err);                                                            \
  } while (0)
  if (A.ql_zw == 4)
    OAK_TQL(4);
  else if (A.ql_zw == 8)
    OAK_TQL(8);
  else if (A.ql_zw == 32)
    OAK_TQL(32);
  else
    OAK_TQL(16);
#undef OAK_TQL
  if constexpr (NP == 128) {  // the epilogue runs at the end of the apply kernel
    hipLaunchKernelGGL(k_tql_apply128, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.N, A.sel_ptr, A.mat,
                       A.ymat, (const double2*)A.rlog, A.rlog_cap, A.tql_status, A.dvec, A.nzc,
                       A.vec, A.ampl, A.err, A.prec == 1 ? 0 : 1);
  } else {
    if constexpr (NP == 64)
      hipLaunchKernelGGL(k_tql_apply64, dim3(A.nzc), dim3(64), 0, st, A.z0, A.zorder, A.sel_ptr,
                         A.ymat, (const double2*)A.rlog, A.rlog_cap, A.tql_status);
    else
      hipLaunchKernelGGL(k_tql_apply<NP>, dim3(A.nzc), dim3(NP < 64 ? 64 : NP), 0, st, A.z0,
                         A.zorder, A.sel_ptr, A.ymat, (const double2*)A.rlog, A.rlog_cap,
                         A.tql_status);
    hipLaunchKernelGGL(k_eigvec_epi<NP>, dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.N, A.sel_ptr,
                       A.mat, A.ymat, A.dvec, A.nzc, A.vec, A.ampl, A.err);
  }
}

// NP = 256, steps 128..253 of the Householder reduction: the trailing
// 128 x 128 block that k_tridiag_lower<256> leaves in rows 128..191 of each zone's
// reflector block, reduced by the NP = 128 kernels in two launches (their
// reflectors into rows 192..255, zone stride 256^2).  Their d and e land
// in entries 0..127 and 0..126 of the reversed tridiagonal, as the steps of
// one 256 reduction would write them.
void launch_tridiag256_tail(const ChunkArgs& A, hipStream_t st) {
  constexpr int64_t ZS = 256 * 256;
  double* blk = A.ymat + 128 * 256;
  double* v2 = A.ymat + 192 * 256;
  hipLaunchKernelGGL((k_tridiag<128, 1>), dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                     blk, v2, A.dvec, A.evec, A.nzc, ZS, 0);
  hipLaunchKernelGGL((k_tridiag<128, 2>), dim3(A.nzc), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                     blk, v2, A.dvec, A.evec, A.nzc, ZS, 0);
}

// NP = 256: QL on the chunk's (d, e), the rotations on Q (umat) with the
// rows in registers, then the epilogue (a, v, mu and Y in umat).
void launch_ql256(const ChunkArgs& A, hipStream_t st) {
  constexpr int NP = 256;
#define OAK_TQL256(ZW)                                                                                    \
  do {                                                                                                    \
    if (A.prec == 1)                                                                                      \
      hipLaunchKernelGGL((k_tql<NP, ZW, true>), dim3((A.nzc + ZW - 1) / ZW), dim3(64), 0, st, A.nzc,     \
                         A.z0, A.zorder, A.sel_ptr, A.dvec, A.evec, (double2*)A.rlog, A.rlog_cap,         \
                         A.tql_status, A.err);                                                            \
    else                                                                                                  \
      hipLaunchKernelGGL((k_tql<NP, ZW, false>), dim3((A.nzc + ZW - 1) / ZW), dim3(64), 0, st, A.nzc,    \
                         A.z0, A.zorder, A.sel_ptr, A.dvec, A.evec, (double2*)A.rlog, A.rlog_cap,         \
                         A.tql_status, A.err);                                                            \
  } while (0)
  if (A.ql_zw == 4)
    OAK_TQL256(4);
  else if (A.ql_zw == 8)
    OAK_TQL256(8);
  else if (A.ql_zw == 32)
    OAK_TQL256(32);
  else
    OAK_TQL256(16);
#undef OAK_TQL256
  hipLaunchKernelGGL(k_tql_apply256, dim3(A.nzc, 2), dim3(256), 0, st, A.z0, A.zorder, A.sel_ptr,
                     A.umat, (const double2*)A.rlog, A.rlog_cap, A.tql_status, A.err);
  hipLaunchKernelGGL((k_eigvec_epi<NP, 1024>), dim3(A.nzc), dim3(1024), 0, st, A.z0, A.zorder, A.N,
                     A.sel_ptr, A.mat, A.umat, A.dvec, A.nzc, A.vec, A.ampl, A.err);
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
