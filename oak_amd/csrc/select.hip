// Observation selection on the GPU: assimilation.F90:3683-3771 (selectObservations)
// evaluated for every zone at once.
//
// Cartesian horizontal localisation (metric 0, loctype 1) bins the observations
// once into a uniform cell grid (cell side >= the largest cutoff; counting sort
// by cell, stable in observation index) and each zone scans only the cells its
// cutoff disc touches.  Every other mode scans all observations.  In both cases
// the test is OAK's own: d <= maxLength with d computed as in distance(), so
// the selected set is identical; the taper weight is exp(-(d/L)^2).
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "common.hpp"
#include "internal.hpp"

namespace oakg {

// ---------------------------------------------------------------------------
// Summary of the geometry: bounding box of the observations, largest cutoff,
// and a check of rmse^2 > 0 (assimilation.F90:2097-2102).  One block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_geo_summary(int64_t m, const double* __restrict__ ox,
                                                      const double* __restrict__ oy,
                                                      const double* __restrict__ rmse2,
                                                      int64_t nz, const double* __restrict__ maxlen,
                                                      GridInfo* __restrict__ gi,
                                                      unsigned* __restrict__ err) {
  __shared__ double sh[4][1024];
  double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
  double cmax = 0;
  bool bad = false;
  for (int64_t l = threadIdx.x; l < m; l += blockDim.x) {
    double x = ox[l], y = oy[l];
    xmin = fmin(xmin, x); xmax = fmax(xmax, x);
    ymin = fmin(ymin, y); ymax = fmax(ymax, y);
    double r = rmse2[l];
    if (!(r > 0) || !isfinite(r)) bad = true;
  }
  for (int64_t z = threadIdx.x; z < nz; z += blockDim.x) cmax = fmax(cmax, maxlen[z]);
  if (bad) atomicOr(err, OAKG_DERR_RMSE);
  int t = threadIdx.x;
  sh[0][t] = xmin; sh[1][t] = -xmax; sh[2][t] = ymin; sh[3][t] = -ymax;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s)
      for (int q = 0; q < 4; ++q) sh[q][t] = fmin(sh[q][t], sh[q][t + s]);
    __syncthreads();
  }
  double r0 = sh[0][0], r1 = -sh[1][0], r2 = sh[2][0], r3 = -sh[3][0];
  __syncthreads();
  sh[0][t] = -cmax;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) sh[0][t] = fmin(sh[0][t], sh[0][t + s]);
    __syncthreads();
  }
  if (t == 0) {
    double c = -sh[0][0];
    if (m == 0) { r0 = r2 = 0; r1 = r3 = 0; }
    double ex = r1 - r0, ey = r3 - r2;
    double h = fmax(c, fmax(ex, ey) / 2048.0);
    if (!(h > 0) || !isfinite(h)) h = 1.0;
    // keep the grid within kMaxCells cells
    for (int it = 0; it < 200; ++it) {
      double gx = floor(ex / h) + 1, gy = floor(ey / h) + 1;
      if (gx * gy <= (double)kMaxCells) break;
      h *= 1.25;
    }
    gi->x0 = r0;
    gi->y0 = r2;
    gi->h = h;
    gi->gx = (int)(floor(ex / h) + 1);
    gi->gy = (int)(floor(ey / h) + 1);
  }
}

__global__ void k_obs_cell(int64_t m, const double* __restrict__ ox, const double* __restrict__ oy,
                           const GridInfo* __restrict__ gi, uint32_t* __restrict__ key,
                           int32_t* __restrict__ idx, int32_t* __restrict__ cell_cnt) {
  int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= m) return;
  GridInfo g = *gi;
  double fx = floor((ox[l] - g.x0) / g.h), fy = floor((oy[l] - g.y0) / g.h);
  int cx = isfinite(fx) ? (int)fmin(fmax(fx, 0.0), (double)(g.gx - 1)) : 0;
  int cy = isfinite(fy) ? (int)fmin(fmax(fy, 0.0), (double)(g.gy - 1)) : 0;
  cx = min(max(cx, 0), g.gx - 1);
  cy = min(max(cy, 0), g.gy - 1);
  uint32_t k = (uint32_t)cy * (uint32_t)g.gx + (uint32_t)cx;
  key[l] = k;
  idx[l] = (int32_t)l;
  atomicAdd(&cell_cnt[k], 1);
}

__global__ void k_gather_sorted(int64_t m, const int32_t* __restrict__ sidx,
                                const double* __restrict__ ox, const double* __restrict__ oy,
                                double* __restrict__ sx, double* __restrict__ sy) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m) return;
  int32_t l = sidx[p];
  sx[p] = ox[l];
  sy[p] = oy[l];
}

// One wave per zone.  FILL = false: count; FILL = true: write the list.
template <bool FILL>
__global__ __launch_bounds__(256) void k_select_binned(
    int64_t nz, const double* __restrict__ zpos, const double* __restrict__ corr,
    const double* __restrict__ maxlen, const GridInfo* __restrict__ gi,
    const int32_t* __restrict__ cell_ptr, const int32_t* __restrict__ sidx,
    const double* __restrict__ sx, const double* __restrict__ sy, int64_t* __restrict__ counts,
    const int64_t* __restrict__ sel_ptr, int32_t* __restrict__ sel_idx,
    double* __restrict__ sel_w) {
  int64_t z = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
  int lane = threadIdx.x % 64;
  if (z >= nz) return;
  GridInfo g = *gi;
  double zx = zpos[4 * z], zy = zpos[4 * z + 1];
  double c = maxlen[z], L = corr[z];
  int cx0, cx1, cy0, cy1;
  // a zone whose disc misses the bin grid gets the empty range (cy0 > cy1)
  zone_cells(g, zx, zy, c, cx0, cx1, cy0, cy1);
  int64_t cnt = 0;
  int64_t base = FILL ? sel_ptr[z] : 0;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int cy = cy0; cy <= cy1; ++cy) {
    int32_t p0 = cell_ptr[(int64_t)cy * g.gx + cx0];
    int32_t p1 = cell_ptr[(int64_t)cy * g.gx + cx1 + 1];
    for (int32_t pb = p0; pb < p1; pb += 64) {
      int32_t p = pb + lane;
      bool sel = false;
      double d = 0;
      if (p < p1) {
        d = oak_distance(0, sx[p], sy[p], zx, zy);
        sel = d <= c;
      }
      if (!FILL) {
        cnt += sel;
      } else {
        uint64_t mask = __ballot(sel);
        if (sel) {
          int64_t pos = base + __popcll(mask & lt_mask);
          sel_idx[pos] = sidx[p];
          sel_w[pos] = oak_taper(d, L);
        }
        base += __popcll(mask);
      }
    }
  }
  if (!FILL) {
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if (lane == 0) counts[z] = cnt;
  }
}

// One wave per zone, every observation in index order (spherical metrics,
// vertical/time localisation, and localise_obs = .false.).
template <bool FILL>
__global__ __launch_bounds__(256) void k_select_scan(
    int64_t nz, int metric, int loctype, int localise, const double* __restrict__ zpos,
    const double* __restrict__ corr, const double* __restrict__ maxlen, int64_t m,
    const double* __restrict__ ox, const double* __restrict__ oy, const double* __restrict__ oz,
    const double* __restrict__ ot, int64_t* __restrict__ counts,
    const int64_t* __restrict__ sel_ptr, int32_t* __restrict__ sel_idx,
    double* __restrict__ sel_w) {
  int64_t z = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
  int lane = threadIdx.x % 64;
  if (z >= nz) return;
  const double* zp = zpos + 4 * z;
  double c = maxlen[z], L = corr[z];
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  if (!FILL) {
    int64_t cnt = 0;
    for (int64_t lb = 0; lb < m; lb += 64) {
      int64_t l = lb + lane;
      if (l < m) {
        double d = loctype == 1 ? oak_distance(metric, ox[l], oy[l], zp[0], zp[1])
                   : loctype == 2 ? fabs(oz[l] - zp[2])
                                  : fabs(ot[l] - zp[3]);
        cnt += d <= c;
      }
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if (lane == 0) counts[z] = localise ? cnt : (cnt > 0 ? m : 0);
    return;
  }
  int64_t base = sel_ptr[z];
  if (sel_ptr[z + 1] == base) return;
  for (int64_t lb = 0; lb < m; lb += 64) {
    int64_t l = lb + lane;
    bool sel = false;
    double d = 0;
    if (l < m) {
      d = loctype == 1 ? oak_distance(metric, ox[l], oy[l], zp[0], zp[1])
          : loctype == 2 ? fabs(oz[l] - zp[2])
                         : fabs(ot[l] - zp[3]);
      sel = localise ? (d <= c) : true;
    }
    uint64_t mask = __ballot(sel);
    if (sel) {
      int64_t pos = base + __popcll(mask & lt_mask);
      sel_idx[pos] = (int32_t)l;
      sel_w[pos] = oak_taper(d, L);
    }
    base += __popcll(mask);
  }
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// sel_ptr[nz+1] = exclusive scan of counts[nz] (counts[nz] holds 0 on entry).
hipError_t scan_counts(Workspace& ws, const int64_t* counts, int64_t* sel_ptr, int64_t nz,
                       hipStream_t st) {
  size_t tmp = 0;
  hipError_t e = rocprim::exclusive_scan(nullptr, tmp, counts, sel_ptr, (int64_t)0,
                                         (size_t)(nz + 1), rocprim::plus<int64_t>(), st);
  if (e != hipSuccess) return e;
  void* t = ws.get(Workspace::SCAN_TMP, tmp);
  if (!t) return hipErrorOutOfMemory;
  return rocprim::exclusive_scan(t, tmp, counts, sel_ptr, (int64_t)0, (size_t)(nz + 1),
                                 rocprim::plus<int64_t>(), st);
}

// Runs the selection for all zones into the context's selection buffers.
// Returns hipSuccess and sets *total (host) to the number of list entries.
hipError_t run_selection(Workspace& ws, const ZonesDev& zd, const ObsDev& od, int localise,
                         hipStream_t st, unsigned* d_err, SelDev& sel, int64_t* total) {
  const int64_t nz = zd.nzones, m = od.m;
  int64_t* counts = (int64_t*)ws.get(Workspace::COUNTS, sizeof(int64_t) * (nz + 1));
  int64_t* sel_ptr = (int64_t*)ws.get(Workspace::SEL_PTR, sizeof(int64_t) * (nz + 1));
  if (!counts || !sel_ptr) return hipErrorOutOfMemory;
  hipError_t e;
  fill_i64(counts + nz, 1, 0, st);
  bool binned = (zd.metric == 0 && zd.loctype == 1 && localise && m > 0);
  GridInfo* gi = (GridInfo*)ws.get(Workspace::GRID, sizeof(GridInfo));
  int32_t *cell_ptr = nullptr, *sidx = nullptr;
  double *sx = nullptr, *sy = nullptr;
  if (binned) {
    hipLaunchKernelGGL(k_geo_summary, dim3(1), dim3(1024), 0, st, m, od.x, od.y, od.rmse2, nz,
                       zd.maxlen, gi, d_err);
    uint32_t* key = (uint32_t*)ws.get(Workspace::KEY, sizeof(uint32_t) * m);
    uint32_t* key2 = (uint32_t*)ws.get(Workspace::KEY2, sizeof(uint32_t) * m);
    int32_t* idx = (int32_t*)ws.get(Workspace::IDX, sizeof(int32_t) * m);
    sidx = (int32_t*)ws.get(Workspace::SIDX, sizeof(int32_t) * m);
    int32_t* cnt = (int32_t*)ws.get(Workspace::CELL_CNT, sizeof(int32_t) * (kMaxCells + 1));
    cell_ptr = (int32_t*)ws.get(Workspace::CELL_PTR, sizeof(int32_t) * (kMaxCells + 1));
    sx = (double*)ws.get(Workspace::SX, sizeof(double) * m);
    sy = (double*)ws.get(Workspace::SY, sizeof(double) * m);
    if (!key || !key2 || !idx || !sidx || !cnt || !cell_ptr || !sx || !sy)
      return hipErrorOutOfMemory;
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk((int64_t)kMaxCells + 1, 256)), dim3(256), 0, st, cnt,
                       (int64_t)kMaxCells + 1, 0);
    hipLaunchKernelGGL(k_obs_cell, dim3(nblk(m, 256)), dim3(256), 0, st, m, od.x, od.y, gi, key,
                       idx, cnt);
    size_t tmp = 0;
    int end_bit = 0;
    while ((1u << end_bit) < (unsigned)kMaxCells) ++end_bit;
    if ((e = rocprim::radix_sort_pairs(nullptr, tmp, key, key2, idx, sidx, (size_t)m, 0, end_bit,
                                       st)) != hipSuccess)
      return e;
    void* t = ws.get(Workspace::SORT_TMP, tmp);
    if (!t) return hipErrorOutOfMemory;
    if ((e = rocprim::radix_sort_pairs(t, tmp, key, key2, idx, sidx, (size_t)m, 0, end_bit, st)) !=
        hipSuccess)
      return e;
    size_t tmp2 = 0;
    if ((e = rocprim::exclusive_scan(nullptr, tmp2, cnt, cell_ptr, (int32_t)0,
                                     (size_t)kMaxCells + 1, rocprim::plus<int32_t>(), st)) !=
        hipSuccess)
      return e;
    void* t2 = ws.get(Workspace::SCAN_TMP2, tmp2);
    if (!t2) return hipErrorOutOfMemory;
    if ((e = rocprim::exclusive_scan(t2, tmp2, cnt, cell_ptr, (int32_t)0, (size_t)kMaxCells + 1,
                                     rocprim::plus<int32_t>(), st)) != hipSuccess)
      return e;
    hipLaunchKernelGGL(k_gather_sorted, dim3(nblk(m, 256)), dim3(256), 0, st, m, sidx, od.x, od.y,
                       sx, sy);
    hipLaunchKernelGGL(k_select_binned<false>, dim3(nblk(nz, 4)), dim3(256), 0, st, nz, zd.zpos,
                       zd.corr, zd.maxlen, gi, cell_ptr, sidx, sx, sy, counts, nullptr, nullptr,
                       nullptr);
  } else {
    if (m > 0)
      hipLaunchKernelGGL(k_select_scan<false>, dim3(nblk(nz, 4)), dim3(256), 0, st, nz, zd.metric,
                         zd.loctype, localise, zd.zpos, zd.corr, zd.maxlen, m, od.x, od.y, od.z,
                         od.t, counts, nullptr, nullptr, nullptr);
    else
      fill_i64(counts, nz, 0, st);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = scan_counts(ws, counts, sel_ptr, nz, st)) != hipSuccess) return e;
  int64_t tot = 0;
  if ((e = hipMemcpyAsync(&tot, sel_ptr + nz, sizeof(int64_t), hipMemcpyDeviceToHost, st)) !=
      hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  int32_t* sel_idx = (int32_t*)ws.get(Workspace::SEL_IDX, sizeof(int32_t) * (tot > 0 ? tot : 1));
  double* sel_w = (double*)ws.get(Workspace::SEL_W, sizeof(double) * (tot > 0 ? tot : 1));
  if (!sel_idx || !sel_w) return hipErrorOutOfMemory;
  if (tot > 0) {
    if (binned)
      hipLaunchKernelGGL(k_select_binned<true>, dim3(nblk(nz, 4)), dim3(256), 0, st, nz, zd.zpos,
                         zd.corr, zd.maxlen, gi, cell_ptr, sidx, sx, sy, counts, sel_ptr, sel_idx,
                         sel_w);
    else
      hipLaunchKernelGGL(k_select_scan<true>, dim3(nblk(nz, 4)), dim3(256), 0, st, nz, zd.metric,
                         zd.loctype, localise, zd.zpos, zd.corr, zd.maxlen, m, od.x, od.y, od.z,
                         od.t, counts, sel_ptr, sel_idx, sel_w);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  sel.ptr = sel_ptr;
  sel.idx = sel_idx;
  sel.w = sel_w;
  sel.counts = counts;
  *total = tot;
  return hipSuccess;
}

}  // namespace oakg
