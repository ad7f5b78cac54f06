// Internal declarations shared by the host-side translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "select_cells.hpp"

namespace oakg {

constexpr int kMaxCells = 1 << 22;  // cap of the observation bin grid


// Device views of the call's inputs (all device pointers).
struct ZonesDev {
  int64_t nzones;
  const int64_t* zone_ptr;
  const double* zpos;    // nz x 4
  const double* corr;    // nz
  const double* maxlen;  // nz
  int metric, loctype;
};

struct ObsDev {
  int64_t m;
  const double *x, *y, *z, *t;
  const double* yo;
  const double* rmse2;
  const uint8_t* excluded;
};

struct SelDev {
  const int64_t* ptr;  // nz+1
  const int32_t* idx;
  const double* w;
  const int64_t* counts;  // nz (list length per zone)
};

// Grow-only device buffers, one per role, owned by a context.
class Workspace {
 public:
  enum Slot {
    // internal buffers
    COUNTS, SEL_PTR, SEL_IDX, SEL_W, ZORDER, ZORDER_P, ZORDER_D, PERSIST_WQ, MAT_D, YMAT_D, VEC_D, GRID, KEY, KEY2, IDX, SIDX, CELL_CNT, CELL_PTR, SX, SY,
    SORT_TMP, SCAN_TMP, SCAN_TMP2, HS, INNOV, RHO, XBAR, HEC, MAT, YMAT, VEC, SWEEPS, DVEC, EVEC, RLOG, UMAT, TQL_STATUS, HS32, NLL_Z, NLL_GTH,
    // staged copies of host inputs and outputs
    Z_PTR, Z_POS, Z_CORR, Z_MAXLEN, O_X, O_Y, O_Z, O_T, O_YO, O_RMSE2, O_EXCL, IN_S, IN_HSF,
    IN_XF, IN_HXF, H_RP, H_COL, H_VAL, H_SHIFT, OUT_XA, OUT_S, OUT_AMPL, U_SEL_PTR, U_SEL_IDX,
    U_SEL_W, MAXCORR, OP_BOXES, OP_TET, OP_VAR, OP_X, OP_Y, OP_Z, OP_T, OP_SLOTC, OP_SLOTW, OP_CNT, OP_RP, OP_VALID, OP_COL, OP_VAL, OP_STATUS, OP_GC, OP_GM, OP_GS, PR_IDX, PR_IN, PR_OUT, REP_PIECES, REP_PBEG, REP_PART, REP_STATS, REP_D, REP_HS, NSLOTS
  };
  Workspace() : ptr_(NSLOTS, nullptr), cap_(NSLOTS, 0) {}
  ~Workspace() { release(); }
  void* get(Slot s, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (cap_[s] >= bytes) return ptr_[s];
    if (ptr_[s]) (void)hipFree(ptr_[s]);
    ptr_[s] = nullptr;
    cap_[s] = 0;
    size_t b = bytes + bytes / 8;  // some headroom for the next call
    if (hipMalloc(&ptr_[s], b) != hipSuccess) {
      ptr_[s] = nullptr;
      return nullptr;
    }
    cap_[s] = b;
    return ptr_[s];
  }
  void release() {
    for (int i = 0; i < NSLOTS; ++i)
      if (ptr_[i]) (void)hipFree(ptr_[i]);
    for (int i = 0; i < NSLOTS; ++i) { ptr_[i] = nullptr; cap_[i] = 0; }
  }

 private:
  std::vector<void*> ptr_;
  std::vector<size_t> cap_;
};

hipError_t run_selection(Workspace& ws, const ZonesDev& zd, const ObsDev& od, int localise,
                         hipStream_t st, unsigned* d_err, SelDev& sel, int64_t* total);
hipError_t scan_counts(Workspace& ws, const int64_t* counts, int64_t* sel_ptr, int64_t nz,
                       hipStream_t st);

// Fill kernels (used instead of hipMemsetAsync so that the sequence can be
// captured into a graph).
void fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t st);
void fill_u32(unsigned* p, int64_t n, unsigned v, hipStream_t st);
void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st);
void fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t st);
void launch_permute_rows(bool gather, int elem, int64_t n, int N, const int64_t* idx,
                         const void* in, int64_t ldi, void* out, int64_t ldo, int64_t lim,
                         unsigned* err, hipStream_t st);

}  // namespace oakg
