// Launch interfaces of the per-zone analysis kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace oakg {

enum Phase { PH_PREP = 0, PH_SELECT, PH_GRAM, PH_EIG, PH_TMAT, PH_UPDATE, PH_END, PH_COUNT };

// Records an event at each phase boundary; per-phase device time is the sum
// of the intervals that start with a mark of that phase.
struct PhaseTimer {
  hipStream_t st = nullptr;
  std::vector<hipEvent_t> pool;
  std::vector<int> phase;
  size_t used = 0;
  void reset(hipStream_t s) { st = s; used = 0; phase.clear(); }
  void mark(int ph) {
    if (used == pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      pool.push_back(e);
    }
    (void)hipEventRecord(pool[used], st);
    phase.push_back(ph);
    ++used;
  }
  // accumulates milliseconds per phase into ms[PH_COUNT]; call after a sync
  void collect(double* ms) {
    for (int i = 0; i < PH_COUNT; ++i) ms[i] = 0;
    for (size_t i = 0; i + 1 < used; ++i) {
      float t = 0;
      if (hipEventElapsedTime(&t, pool[i], pool[i + 1]) == hipSuccess) ms[phase[i]] += t;
    }
  }
  ~PhaseTimer() {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

struct PrepLocArgs {
  int64_t m;
  int N;
  const double* HSf;
  int64_t ldHSf;
  const double *Hxf, *yo, *rmse2;
  const uint8_t* excl;
  double *HS, *innov, *rho;
  unsigned* err;
  double hs_scale;  // HS = hs_scale * HSf (1 for locAnalysis, the inflation for the likelihood)
};

struct PrepEnsArgs {
  int64_t n, m;
  int N;
  const double* E;
  int64_t ldE;
  const int64_t* row_ptr;
  const int32_t* col;
  const double* val;
  const double* Hshift;
  const double *yo, *rmse2;
  const uint8_t* excl;
  double scaling;
  double *xbar, *HEc, *HS, *innov, *rho;
  unsigned* err;
  // mixed precision (FP32 state): E32, Hshift32 in, HS32 out
  const float* E32;
  const float* Hshift32;
  float* HS32;
};

// The zone at position pos of the zone loop: zones run in the order of
// zorder when the loop has one (zones grouped by observation count), else in
// their own order.
__device__ __forceinline__ int64_t zone_at(const int32_t* __restrict__ zorder, int64_t pos) {
  return zorder ? (int64_t)zorder[pos] : pos;
}

struct ChunkArgs {
  int64_t z0;               // first position of the chunk in the zone loop
  const int32_t* zorder;    // position -> zone, or NULL (position = zone)
  int nzc, N, mode;  // mode 0 = locAnalysis, 1 = ensemble
  int64_t m;
  const int64_t* zone_ptr;
  const int64_t* sel_ptr;
  const int32_t* sel_idx;
  const double* sel_w;
  const double *rho, *innov, *HS;
  double *mat, *ymat, *vec, *ampl;
  double* umat;  // NP = 256: Q, then U, then Y (mat: M then T; ymat: the reflectors)
  int32_t* sweeps;
  // eigensolver: 0 = one-sided Jacobi, 1 = Householder tridiagonalisation + implicit QL
  int eig_method;
  int ql_zw;                // zones per wave of the QL kernel (4, 8, 16 or 32)
  int tri_split;            // NP = 128: the Householder reduction as two launches
  unsigned* persist_wq;     // NP = 128: work counter of the persistent Gram (NULL: one
                            // workgroup per zone)
  int persist_item;         // its scheduling: 1 a zone per workgroup, 0 each wave its quarter
  int persist_grid;         // its workgroups (persist_grid128)
  double *dvec, *evec;      // tridiagonal / eigenvalues, NP x nzc
  void* rlog;               // QL rotation log, rlog_cap records of 16 B per zone
  int64_t rlog_cap;
  int32_t* tql_status;      // per zone: bit0 log overflow, bit1 no convergence, >>2 log length
  unsigned* err;
  double tol;
  int max_sweeps;
  const double* X0;  // xf (mode 0) or xbar (mode 1)
  const double* S;   // Sf (mode 0) or E (mode 1)
  int64_t ldS;
  double* xa;   // mode 0
  double* Out;  // Sa (mode 0) or Ea (mode 1)
  int64_t ldO;
  double scaling, infl, maxcorr;
  const double* maxcorr_v;  // per-element clamp [n] (zone-permuted) or NULL: maxcorr for all
  // mixed precision: 1 = FP32 S (HS32), FP32 members (S32) and FP32 Out32
  int prec;
  const float* HS32;
  const float* S32;
  float* Out32;
  // mode 2 (innovation likelihood and its adjoint): per-zone L, per-zone
  // outputs J_z, dJ/dL_z, dJ/dalpha_z and per-observation dJ/dlog sigma
  const double* corr;
  double *nll_J, *nll_gL, *nll_gA, *nll_gth;
};

// Innovation statistics of report (report.hip): pieces of at most kRepRows
// observations, each inside one observation range.
constexpr int kRepRows = 512;
struct RepPiece {
  int64_t r0, r1;
};
struct ReportArgs {
  int N;
  int layout;  // 0: HS column-major (ldHS); 1: row-major NP-padded in hs_col order
  int64_t npieces;
  int nranges;
  const RepPiece* pieces;
  const int64_t* pbeg;  // [nranges+1] first piece of each range
  const double *d, *rmse2;
  const uint8_t* excl;
  const double* HS;
  const float* HS32;  // layout 1 in FP32 (mixed path)
  int64_t ldHS;
  double* part;   // npieces x (4 + NP)
  double* stats;  // nranges x 5
  unsigned* err;
};

void launch_prep_loc(int NP, const PrepLocArgs& A, hipStream_t st);
void launch_prep_ens(int NP, const PrepEnsArgs& A, hipStream_t st);
void launch_chunk(int NP, const ChunkArgs& A, hipStream_t st, PhaseTimer* tm);
// NP = 128 in two parts (Gram + eigensolve, then transform + update), so
// that the zones solved in observation space can run in between
// Workgroups of the persistent Gram that fit on the current device at once
// (0 when the runtime cannot say).
int persist_grid128();
void launch_solve128(const ChunkArgs& A, hipStream_t st, PhaseTimer* tm);
void launch_apply128(const ChunkArgs& A, hipStream_t st, PhaseTimer* tm);
// eigensolver of a chunk for NP <= 64 (QL pipeline or Jacobi, with fallback)
void launch_eig_small(int NP, const ChunkArgs& A, hipStream_t st);
// zones with 1 <= p < 64 at NP = 128 in observation space (dual.hip): A has
// the dual order and the NP = 64 buffers; ymat128/vec128/ampl receive W, a, v
void launch_dual128(const ChunkArgs& A, int N, double* ymat128, double* vec128, double* ampl,
                    hipStream_t st, PhaseTimer* tm);
void launch_chunk128(const ChunkArgs& A, hipStream_t st, PhaseTimer* tm);
void launch_tql_pipeline(int NP, const ChunkArgs& A, hipStream_t st);
// NP = 256 (wide256.hip): the whole chunk; QL, rotations and epilogue of
// its eigensolver (tridiag.hip)
void launch_chunk256(const ChunkArgs& A, hipStream_t st, PhaseTimer* tm);
void launch_ql256(const ChunkArgs& A, hipStream_t st);
void launch_tridiag256_tail(const ChunkArgs& A, hipStream_t st);
// NP = 128, OAKG_TUNE_TRIDIAG_SPLIT 3: steps 0..63 of the reduction on the
// lower-triangle register tiles (wide256.hip)
void launch_tridiag_lower128(const ChunkArgs& A, hipStream_t st);
// OAKG_TUNE_TRIDIAG_SPLIT 4: steps 64..125 by the lower-triangle kernel at
// NP = 64 on the trailing block (after launch_tridiag_lower128)
void launch_tridiag_lower128_tail(const ChunkArgs& A, hipStream_t st);
// NP = 64: reduction on the lower triangle and Q on MFMA (wide256.hip)
void launch_tridiag_lower64(const ChunkArgs& A, hipStream_t st);
void launch_prep_ens_f32(int NP, const PrepEnsArgs& A, hipStream_t st);
void launch_update_f32(int NP, const ChunkArgs& A, hipStream_t st);
void launch_nll_adjoint(int NP, const ChunkArgs& A, hipStream_t st);
void launch_report(int NP, const ReportArgs& A, hipStream_t st);
void launch_nll_totals(int64_t nz, const double* Jz, const double* gA, double* out,
                       hipStream_t st);

}  // namespace oakg
