// Observation operator from coordinates (interp.hip): grid, box tree and
// launch interfaces.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace oakg {

// A model variable's grid on the host (for the box tree): n dims, nodes per
// dim, node coordinates per dim (first index fastest).
struct GridHost {
  int n;
  int64_t shape[4];
  const double* coord[4];
};

// The same on the device, with the mask (1 = land) and the node -> state
// column map (-1: not in the state; NULL: the node index).
struct GridDev {
  int n;
  int64_t shape[4];
  const double* coord[4];
  const uint8_t* mask;
  const int64_t* state;
};

// A box of the search tree: node index range, coordinate bounds, children.
struct BoxDev {
  int32_t imin[4], imax[4];
  double xmin[4], xmax[4];
  int32_t child0, nchild;
};

struct ObsPosDev {
  const int32_t* var;
  const double *x, *y, *z, *t;
};

int tet_count(int n);
std::vector<double> tet_table(int n);
std::vector<BoxDev> build_box_tree(const GridHost& g);

void launch_obs_interp(int n, int64_t m, int var, const GridDev& g, const BoxDev* boxes,
                       const double* tc, int nbth, const ObsPosDev& o, int slot_stride,
                       int32_t* slot_col, double* slot_w, int64_t* counts, uint8_t* valid,
                       unsigned* status, hipStream_t st);
void launch_obs_compact(int64_t m, int slot_stride, const int64_t* row_ptr,
                        const int32_t* slot_col, const double* slot_w, int32_t* col, double* val,
                        hipStream_t st);

}  // namespace oakg
