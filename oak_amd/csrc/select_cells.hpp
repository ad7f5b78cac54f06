// Range of observation bins that a zone's cutoff disc can touch
// (selectObservations, assimilation.F90:3683-3771, binned).  Host and device:
// the CPU test suite compiles this header with g++ to check the ranges of
// zones outside the observations' bounding box.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define OAKG_HD __host__ __device__
#else
#define OAKG_HD
#endif

namespace oakg {

struct GridInfo {
  double x0, y0, h;  // lower-left corner of the bin grid and bin side
  int gx, gy;        // bins per row and rows
};

// Bins [cx0, cx1] x [cy0, cy1] that a disc of radius c around (zx, zy) can
// touch, with a margin that covers the rounding of the bin arithmetic (exact
// selection is decided by the distance test itself).  Returns false when the
// disc misses the grid: then no bin is to be read.  On true, every index is
// inside the grid (0 <= cx0 <= cx1 < gx, 0 <= cy0 <= cy1 < gy).
OAKG_HD inline bool zone_cells(const GridInfo& g, double zx, double zy, double c, int& cx0,
                               int& cx1, int& cy0, int& cy1) {
  const double a0 = (zx - c - g.x0) / g.h, a1 = (zx + c - g.x0) / g.h;
  const double b0 = (zy - c - g.y0) / g.h, b1 = (zy + c - g.y0) / g.h;
  const double eps = 1e-7;
  // a NaN coordinate makes every comparison below false: no bins
  if (!(a1 + eps >= 0.0) || !(b1 + eps >= 0.0) || !(a0 - eps < (double)g.gx) ||
      !(b0 - eps < (double)g.gy)) {
    cx0 = cy0 = 0;
    cx1 = cy1 = -1;
    return false;
  }
  // the clamps are applied in floating point, so that a far-away zone cannot
  // overflow the conversion to int
  cx0 = (int)fmin(fmax(floor(a0 - eps), 0.0), (double)(g.gx - 1));
  cx1 = (int)fmin(fmax(floor(a1 + eps), 0.0), (double)(g.gx - 1));
  cy0 = (int)fmin(fmax(floor(b0 - eps), 0.0), (double)(g.gy - 1));
  cy1 = (int)fmin(fmax(floor(b1 + eps), 0.0), (double)(g.gy - 1));
  return cx0 <= cx1 && cy0 <= cy1;
}

}  // namespace oakg
