// Observation operator from coordinates on the GPU: genObservationOper
// (assimilation.F90:2471-2656) with cinterp (ndgrid.F90:1183-1257), the
// databox search of ndgrid_inc.F90:534-990 and the hyper-tetrahedral
// interpolation of ndgrid.F90:357-665, followed by packSparseMatrix
// (assimilation.F90:4132-4208) and the zone permutation of the state columns
// (assimilation.F90:2404-2409).
//
// One thread per observation: a depth-first walk of the grid's box tree
// (built once on the host, flattened), the cell test of the candidate
// leaves (InCube: the first simplex of the cell's split that holds the
// point), then the 2^n corner weights of the cell.  The arithmetic of the
// simplex test and weights runs in the same order as the host restatement
// (sums in index order, LU with partial pivoting, no contraction), so
// indices and weights are reproducible bit for bit.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "interp.hpp"

namespace oakg {

constexpr double kTetTol = 1e-8;  // interp_tetrahedron, ndgrid.F90:474

static int factorial_i(int n) {
  int f = 1;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

int tet_count(int n) { return factorial_i(n) * (1 << (n - 1)); }

// split (ndgrid.F90:357-435) into the slab starting at simplex l0 of the
// column-major table tc(2^n, n+1, nbth): entry (j, k, l) at j + 2^n (k + (n+1) l).
static void split_tab(int n, int subn, const int* fixeddim, const int* sel, double* tc, int l0) {
  const int twon = 1 << n, ldk = twon, ldl = twon * (n + 1);
  if (subn == 1) {
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < twon; ++j) tc[j + ldk * k + ldl * l0] = 0.0;
    int k = 0;
    for (int j = 0; j < twon; ++j)
      if (sel[j]) tc[j + ldk * (k++) + ldl * l0] = 1.0;
    return;
  }
  int cnt = 0;
  for (int j = 0; j < twon; ++j) cnt += sel[j] != 0;
  const int nb = factorial_i(subn) * (1 << (subn - 1));
  const int sub_nb = factorial_i(subn - 1) * (1 << (subn - 2));
  for (int l = 0; l < nb; ++l)
    for (int j = 0; j < twon; ++j) tc[j + ldk * subn + ldl * (l0 + l)] = sel[j] ? 1.0 / cnt : 0.0;
  int m = 0;
  for (int i = 0; i < n; ++i) {
    if (fixeddim[i]) continue;
    int fd[4] = {0, 0, 0, 0}, sub[16];
    for (int d = 0; d < n; ++d) fd[d] = fixeddim[d];
    fd[i] = 1;
    for (int side = 0; side <= 1; ++side) {
      for (int k = 0; k < twon; ++k) sub[k] = sel[k] && (((k >> i) & 1) == side);
      split_tab(n, subn - 1, fd, sub, tc, l0 + m);
      m += sub_nb;
    }
  }
}

std::vector<double> tet_table(int n) {
  std::vector<double> tc((size_t)(1 << n) * (n + 1) * tet_count(n), 0.0);
  int fixed[4] = {0, 0, 0, 0}, sel[16];
  for (int j = 0; j < (1 << n); ++j) sel[j] = 1;
  split_tab(n, n, fixed, sel, tc.data(), 0);
  return tc;
}

// The box tree of locate_databox: box [imin, imax] of node indices with the
// bounds of its nodes' coordinates; a box with a side above 1 has the
// sub-boxes (im = (imin + imax) / 2; bit d of m - 1 set: [imin, im] along d,
// else [im, imax]) with a side >= 1, stored contiguously in m order; a box
// with no side above 1 is a cell.  Bounds come up from the leaves (the
// sub-boxes cover their parent's nodes).
struct TreeBuilder {
  const GridHost& g;
  std::vector<BoxDev>& boxes;
  int64_t lin(const int64_t* p) const {
    int64_t l = 0, off = 1;
    for (int d = 0; d < g.n; ++d) {
      l += p[d] * off;
      off *= g.shape[d];
    }
    return l;
  }
  void build(int64_t id) {
    const int n = g.n;
    int64_t imin[4], imax[4];
    for (int d = 0; d < 4; ++d) {
      imin[d] = boxes[id].imin[d];
      imax[d] = boxes[id].imax[d];
    }
    bool split = false;
    for (int d = 0; d < n; ++d) split |= imax[d] - imin[d] > 1;
    if (!split) {
      double lo[4], hi[4];
      for (int d = 0; d < n; ++d) {
        lo[d] = HUGE_VAL;
        hi[d] = -HUGE_VAL;
      }
      int64_t cnt[4], tot = 1;
      for (int d = 0; d < n; ++d) {
        cnt[d] = imax[d] - imin[d] + 1;
        tot *= cnt[d];
      }
      for (int64_t q = 0; q < tot; ++q) {
        int64_t p[4], r = q;
        for (int d = 0; d < n; ++d) {
          p[d] = imin[d] + r % cnt[d];
          r /= cnt[d];
        }
        const int64_t l = lin(p);
        for (int d = 0; d < n; ++d) {
          const double x = g.coord[d][l];
          if (x < lo[d]) lo[d] = x;
          if (x > hi[d]) hi[d] = x;
        }
      }
      for (int d = 0; d < n; ++d) {
        boxes[id].xmin[d] = lo[d];
        boxes[id].xmax[d] = hi[d];
      }
      boxes[id].child0 = -1;
      boxes[id].nchild = 0;
      return;
    }
    int64_t im[4];
    for (int d = 0; d < n; ++d) im[d] = (imin[d] + imax[d]) / 2;
    const int64_t first = (int64_t)boxes.size();
    int nk = 0;
    for (int q = 0; q < (1 << n); ++q) {
      BoxDev b{};
      bool keep = false;
      for (int d = 0; d < 4; ++d) {
        if (d >= n) {
          b.imin[d] = b.imax[d] = 0;
          continue;
        }
        if ((q >> d) & 1) {
          b.imin[d] = (int32_t)imin[d];
          b.imax[d] = (int32_t)im[d];
        } else {
          b.imin[d] = (int32_t)im[d];
          b.imax[d] = (int32_t)imax[d];
        }
        keep |= b.imax[d] - b.imin[d] >= 1;
      }
      if (!keep) continue;
      boxes.push_back(b);
      ++nk;
    }
    boxes[id].child0 = (int32_t)first;
    boxes[id].nchild = nk;
    for (int q = 0; q < nk; ++q) build(first + q);
    for (int d = 0; d < n; ++d) {
      double lo = HUGE_VAL, hi = -HUGE_VAL;
      for (int q = 0; q < nk; ++q) {
        const BoxDev& c = boxes[first + q];
        if (c.xmin[d] < lo) lo = c.xmin[d];
        if (c.xmax[d] > hi) hi = c.xmax[d];
      }
      boxes[id].xmin[d] = lo;
      boxes[id].xmax[d] = hi;
    }
  }
};

std::vector<BoxDev> build_box_tree(const GridHost& g) {
  std::vector<BoxDev> boxes;
  BoxDev root{};
  for (int d = 0; d < 4; ++d) {
    root.imin[d] = 0;
    root.imax[d] = d < g.n ? (int32_t)(g.shape[d] - 1) : 0;
  }
  boxes.push_back(root);
  TreeBuilder tb{g, boxes};
  tb.build(0);
  return boxes;
}

// ---------------------------------------------------------------------------
// Device side
// ---------------------------------------------------------------------------
// Barycentric coordinates c of xi in the simplex X (n x (n+1), column-major)
// (interp_tetrahedron, ndgrid.F90:464-629, non-degenerate branch).  Returns
// 0 inside, 1 outside, 2 degenerate (|det| <= tol).
template <int N>
__device__ int tet_coeff(const double* X, const double* xi, double* c) {
#pragma clang fp contract(off)
  constexpr int M = N + 1;
  double A[M][M], b[M], xc[N];
#pragma unroll
  for (int d = 0; d < N; ++d) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) s = s + X[d + N * k];
    xc[d] = s / M;
  }
#pragma unroll
  for (int k = 0; k < M; ++k) A[0][k] = 1.0;
#pragma unroll
  for (int d = 0; d < N; ++d)
#pragma unroll
    for (int k = 0; k < M; ++k) A[d + 1][k] = X[d + N * k] - xc[d];
  b[0] = 1.0;
#pragma unroll
  for (int d = 0; d < N; ++d) b[d + 1] = xi[d] - xc[d];
  double det = 1.0;
  bool zero_piv = false;
#pragma unroll
  for (int j = 0; j < M; ++j) {
    if (zero_piv) break;
    int p = j;
    double best = fabs(A[j][j]);
#pragma unroll
    for (int i = j + 1; i < M; ++i)
      if (fabs(A[i][j]) > best) {
        best = fabs(A[i][j]);
        p = i;
      }
    if (p != j) {
#pragma unroll
      for (int i = j + 1; i < M; ++i)
        if (i == p) {
#pragma unroll
          for (int k = 0; k < M; ++k) {
            const double t = A[j][k];
            A[j][k] = A[i][k];
            A[i][k] = t;
          }
          const double t = b[j];
          b[j] = b[i];
          b[i] = t;
        }
      det = -det;
    }
    const double piv = A[j][j];
    det = det * piv;
    if (piv == 0.0) {
      zero_piv = true;
      continue;
    }
#pragma unroll
    for (int i = j + 1; i < M; ++i) {
      const double l = A[i][j] / piv;
#pragma unroll
      for (int k = j + 1; k < M; ++k) A[i][k] = A[i][k] - l * A[j][k];
      b[i] = b[i] - l * b[j];
    }
  }
  if (!(fabs(det) > kTetTol)) return 2;
#pragma unroll
  for (int i = M - 1; i >= 0; --i) {
    double s = b[i];
#pragma unroll
    for (int k = i + 1; k < M; ++k) s = s - A[i][k] * c[k];
    c[i] = s / A[i][i];
  }
  int out = 0;
#pragma unroll
  for (int k = 0; k < M; ++k)
    if (!(0.0 - kTetTol <= c[k] && c[k] <= 1.0 + kTetTol)) out = 1;
  return out;
}

// interp_cube (ndgrid.F90:636-665): the first simplex of the split holding
// xi; its 2^n corner weights into w (if not null).  0 inside, 1 outside,
// 2 degenerate.
template <int N>
__device__ int cube_coeff(const double* __restrict__ tc, int nbth, const double* px,
                          const double* xi, double* w) {
#pragma clang fp contract(off)
  constexpr int TWON = 1 << N, M = N + 1;
  for (int l = 0; l < nbth; ++l) {
    const double* t = tc + (size_t)TWON * M * l;
    double X[N * M], c[M];
#pragma unroll
    for (int d = 0; d < N; ++d)
#pragma unroll
      for (int k = 0; k < M; ++k) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < TWON; ++j) s = s + px[d + N * j] * t[j + TWON * k];
        X[d + N * k] = s;
      }
    const int r = tet_coeff<N>(X, xi, c);
    if (r == 2) return 2;
    if (r == 0) {
      if (w) {
#pragma unroll
        for (int j = 0; j < TWON; ++j) {
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < M; ++k) s = s + t[j + TWON * k] * c[k];
          w[j] = s;
        }
      }
      return 0;
    }
  }
  return 1;
}

// Corner points of the cell with lower corner ind; corner j takes index
// min(ind + 1, shape - 1) along the set bits of j (InCube); node indices out.
template <int N>
__device__ void cell_corners(const GridDev& g, const int32_t* ind, double* px, int64_t* node) {
  constexpr int TWON = 1 << N;
#pragma unroll
  for (int j = 0; j < TWON; ++j) {
    int64_t l = 0, off = 1;
#pragma unroll
    for (int d = 0; d < N; ++d) {
      int64_t p = ind[d];
      if ((j >> d) & 1) p = ind[d] + 1 < g.shape[d] - 1 ? ind[d] + 1 : g.shape[d] - 1;
      l += p * off;
      off *= g.shape[d];
    }
    if (node) node[j] = l;
#pragma unroll
    for (int d = 0; d < N; ++d) px[d + N * j] = g.coord[d][l];
  }
}

// Pass 1: per observation of variable var, the cell (depth-first search of
// the box tree), the corner weights, and the entries of its row of H in the
// slots [stride l, stride l + 2^n): state column (or -1: no entry) and
// weight (the slots of every observation start out as -1).
// counts[l] = number of entries, valid[l] = validobs, status bit 1 for a
// degenerate cell (OAKG_EUNSUP) and bit 2 for a search deeper than the stack.
template <int N>
__global__ __launch_bounds__(256) void k_obs_interp(int64_t m, int var, GridDev g,
                                                    const BoxDev* __restrict__ boxes,
                                                    const double* __restrict__ tc, int nbth,
                                                    const int32_t* __restrict__ obs_var,
                                                    const double* __restrict__ ox,
                                                    const double* __restrict__ oy,
                                                    const double* __restrict__ oz,
                                                    const double* __restrict__ ot, int stride,
                                                    int32_t* __restrict__ slot_col,
                                                    double* __restrict__ slot_w,
                                                    int64_t* __restrict__ counts,
                                                    uint8_t* __restrict__ valid,
                                                    unsigned* __restrict__ status) {
  constexpr int TWON = 1 << N;
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= m) return;
  if (obs_var[l] != var) return;
  double xi[N];
  const double* oc[4] = {ox, oy, oz, ot};
#pragma unroll
  for (int d = 0; d < N; ++d) xi[d] = oc[d][l];
  // depth-first search: stack of (box, next child)
  constexpr int kDepth = 48;
  int32_t st_box[kDepth];
  int32_t st_next[kDepth];
  int depth = -1;
  int32_t found[4] = {0, 0, 0, 0};
  bool hit = false;
  auto inside_bounds = [&](const BoxDev& b) {
    bool in = true;
#pragma unroll
    for (int d = 0; d < N; ++d) in = in && !(xi[d] < b.xmin[d] || xi[d] > b.xmax[d]);
    return in;
  };
  auto test_cell = [&](const BoxDev& b) -> int {
    double px[N * TWON];
    int32_t ind[N];
#pragma unroll
    for (int d = 0; d < N; ++d) ind[d] = b.imin[d];
    cell_corners<N>(g, ind, px, nullptr);
    return cube_coeff<N>(tc, nbth, px, xi, nullptr);
  };
  {
    const BoxDev root = boxes[0];
    if (inside_bounds(root)) {
      if (root.nchild == 0) {
        const int r = test_cell(root);
        if (r == 2) atomicOr(status, 1u);
        if (r == 0) {
          hit = true;
#pragma unroll
          for (int d = 0; d < N; ++d) found[d] = root.imin[d];
        }
      } else {
        depth = 0;
        st_box[0] = 0;
        st_next[0] = 0;
      }
    }
  }
  while (depth >= 0 && !hit) {
    const BoxDev& parent = boxes[st_box[depth]];
    if (st_next[depth] >= parent.nchild) {
      --depth;
      continue;
    }
    const int32_t cid = parent.child0 + st_next[depth]++;
    const BoxDev child = boxes[cid];
    if (!inside_bounds(child)) continue;
    if (child.nchild == 0) {
      const int r = test_cell(child);
      if (r == 2) {
        atomicOr(status, 1u);
        depth = -1;
        break;
      }
      if (r == 0) {
        hit = true;
#pragma unroll
        for (int d = 0; d < N; ++d) found[d] = child.imin[d];
      }
      continue;
    }
    if (depth + 1 >= kDepth) {
      atomicOr(status, 2u);
      depth = -1;
      break;
    }
    ++depth;
    st_box[depth] = cid;
    st_next[depth] = 0;
  }
  int64_t cnt = 0;
  uint8_t ok = 0;
  if (hit) {
    // cinterp: masked corners put the point out of the grid
    double px[N * TWON], w[TWON];
    int64_t node[TWON];
    cell_corners<N>(g, found, px, node);
    bool masked = false;
    if (g.mask)
#pragma unroll
      for (int j = 0; j < TWON; ++j) masked = masked || g.mask[node[j]] != 0;
    if (!masked) {
      const int r = cube_coeff<N>(tc, nbth, px, xi, w);
      if (r == 2) atomicOr(status, 1u);
      if (r == 0) {
        // packSparseMatrix: entries of nodes outside the state are dropped;
        // validobs is the validity of the observation's last entry
#pragma unroll
        for (int j = 0; j < TWON; ++j) {
          const int64_t s = g.state ? g.state[node[j]] : node[j];
          slot_col[l * stride + j] = s >= 0 ? (int32_t)s : -1;
          slot_w[l * stride + j] = w[j];
          cnt += s >= 0;
          ok = s >= 0;
        }
      }
    }
  }
  if (!(hit && cnt > 0)) {
#pragma unroll
    for (int j = 0; j < TWON; ++j) slot_col[l * stride + j] = -1;
  }
  counts[l] = cnt;
  valid[l] = ok;
}

// Pass 2: compaction of the slots into CSR rows.
__global__ void k_obs_compact(int64_t m, int stride, const int64_t* __restrict__ row_ptr,
                              const int32_t* __restrict__ slot_col,
                              const double* __restrict__ slot_w, int32_t* __restrict__ col,
                              double* __restrict__ val) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= m) return;
  int64_t p = row_ptr[l];
  for (int j = 0; j < stride; ++j) {
    const int32_t s = slot_col[l * stride + j];
    if (s >= 0) {
      col[p] = s;
      val[p] = slot_w[l * stride + j];
      ++p;
    }
  }
}

void launch_obs_interp(int n, int64_t m, int var, const GridDev& g, const BoxDev* boxes,
                       const double* tc, int nbth, const ObsPosDev& o, int slot_stride,
                       int32_t* slot_col, double* slot_w, int64_t* counts, uint8_t* valid,
                       unsigned* status, hipStream_t st) {
  const dim3 grid((unsigned)((m + 255) / 256)), blk(256);
#define OAK_INTERP(NN)                                                                            \
  hipLaunchKernelGGL(k_obs_interp<NN>, grid, blk, 0, st, m, var, g, boxes, tc, nbth, o.var, o.x, \
                     o.y, o.z, o.t, slot_stride, slot_col, slot_w, counts, valid, status)
  switch (n) {
    case 1: OAK_INTERP(1); break;
    case 2: OAK_INTERP(2); break;
    case 3: OAK_INTERP(3); break;
    case 4: OAK_INTERP(4); break;
  }
#undef OAK_INTERP
}

void launch_obs_compact(int64_t m, int slot_stride, const int64_t* row_ptr,
                        const int32_t* slot_col, const double* slot_w, int32_t* col, double* val,
                        hipStream_t st) {
  if (m > 0)
    hipLaunchKernelGGL(k_obs_compact, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, m,
                       slot_stride, row_ptr, slot_col, slot_w, col, val);
}

}  // namespace oakg
