/*
 * oak_oracle_interp.c -- CPU restatement of OAK's generation of the
 * observation operator from coordinates (genObservationOper,
 * assimilation.F90:2471-2656, with cinterp and the hyper-tetrahedral
 * interpolation of ndgrid.F90:357-665, 1183-1257, the grid search of
 * ndgrid_inc.F90:534-990, and packSparseMatrix, assimilation.F90:4132-4208).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path; it is
 * built into the oracle library and never linked by the product.
 *
 * Pinning: test/test_ndgrid.F90's known answers (a linear field
 * f = 2x + 4y is interpolated exactly on regular and curvilinear grids, the
 * point in the middle of the domain is inside), the same for linear fields in
 * 1 to 4 dimensions, the partition of the unit cube by the tetrahedra
 * (volumes add up to 1, each corner weight table a convex combination),
 * coefficients >= -tol summing to 1, points outside the grid or next to a
 * masked node flagged out (tests/test_oracle_interp.py).
 *
 * Scope: grids with at least two nodes in every dimension.  A dimension of
 * one node makes every tetrahedron of the cell degenerate; OAK then takes
 * the SVD branch of interp_tetrahedron (ndgrid.F90:514-627), which is not
 * restated here: such cells return OKO_EUNSUP.
 *
 * Operation order: every sum runs over its index in increasing order, with
 * no fused multiply-add (the library builds with -ffp-contract=off), and the
 * (n+1) x (n+1) system of interp_tetrahedron is solved by LU with partial
 * pivoting and substitution.  OAK forms the inverse with LAPACK getrf/getri
 * (matoper_inc.F90:570-607): the coefficients agree with it to rounding.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OKO_OK 0
#define OKO_EARG 1
#define OKO_ENOMEM 4
#define OKO_EUNSUP 5

#define OKO_TOL 1e-8 /* interp_tetrahedron, ndgrid.F90:474 */

static int factorial(int n) {
  int f = 1;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

/* split (ndgrid.F90:357-435) into the slab tc(:, 1:subn+1, l0:...) of the
 * column-major table tc(twon, n+1, nbth): element (j, k, l) (0-based) at
 * j + twon (k + (n+1) l). */
static void split_rec(int n, int subn, const int* fixeddim, const int* selection, double* tc,
                      int l0) {
  const int twon = 1 << n, ldk = twon, ldl = twon * (n + 1);
  if (subn == 1) {
    /* the slab passed down holds vertex columns 1:2 of one tetrahedron */
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < twon; ++j) tc[j + ldk * k + ldl * l0] = 0.0;
    int k = 0;
    for (int j = 0; j < twon; ++j)
      if (selection[j]) {
        tc[j + ldk * k + ldl * l0] = 1.0;
        ++k;
      }
    return;
  }
  int cnt = 0;
  for (int j = 0; j < twon; ++j) cnt += selection[j] != 0;
  const int nbth = factorial(subn) * (1 << (subn - 1));
  const int subnbth = factorial(subn - 1) * (1 << (subn - 2));
  /* the middle point of the selected corners, vertex subn+1 of every tetrahedron */
  for (int l = 0; l < nbth; ++l)
    for (int j = 0; j < twon; ++j)
      tc[j + ldk * subn + ldl * (l0 + l)] = selection[j] ? 1.0 / cnt : 0.0;
  int m = 0;
  int fd[4], sub[16];
  for (int i = 0; i < n; ++i) {
    if (fixeddim[i]) continue;
    memcpy(fd, fixeddim, sizeof(int) * (size_t)n);
    fd[i] = 1;
    for (int side = 0; side <= 1; ++side) {
      for (int k = 0; k < twon; ++k) sub[k] = selection[k] && (((k >> i) & 1) == side);
      split_rec(n, subn - 1, fd, sub, tc, l0 + m);
      m += subnbth;
    }
  }
}

/* The table of init_basegrid (ndgrid.F90:869-900): tc(2^n, n+1, n! 2^(n-1)). */
int oko_tetrahedra(int n, double* tc) {
  if (n < 1 || n > 4) return OKO_EARG;
  int fixed[4] = {0, 0, 0, 0}, sel[16];
  for (int j = 0; j < (1 << n); ++j) sel[j] = 1;
  split_rec(n, n, fixed, sel, tc, 0);
  return OKO_OK;
}

/* interp_tetrahedron (ndgrid.F90:464-629), non-degenerate branch: the
 * barycentric coefficients c of xi in the simplex X (n x (n+1), column-major).
 * Returns 0 (inside or outside decided: *out set) or OKO_EUNSUP when
 * |det| <= tol (the degenerate branch). */
static int interp_tetrahedron(int n, const double* X, const double* xi, int* out, double* c) {
  const int m = n + 1;
  double A[5][5], b[5], xc[4];
  for (int d = 0; d < n; ++d) {
    double s = 0.0;
    for (int k = 0; k < m; ++k) s = s + X[d + n * k];
    xc[d] = s / m;
  }
  for (int k = 0; k < m; ++k) A[0][k] = 1.0;
  for (int d = 0; d < n; ++d)
    for (int k = 0; k < m; ++k) A[d + 1][k] = X[d + n * k] - xc[d];
  b[0] = 1.0;
  for (int d = 0; d < n; ++d) b[d + 1] = xi[d] - xc[d];
  /* LU with partial pivoting (first largest |pivot|), then substitution */
  double det = 1.0;
  for (int j = 0; j < m; ++j) {
    int p = j;
    double best = fabs(A[j][j]);
    for (int i = j + 1; i < m; ++i)
      if (fabs(A[i][j]) > best) {
        best = fabs(A[i][j]);
        p = i;
      }
    if (p != j) {
      for (int k = 0; k < m; ++k) {
        const double t = A[j][k];
        A[j][k] = A[p][k];
        A[p][k] = t;
      }
      const double t = b[j];
      b[j] = b[p];
      b[p] = t;
      det = -det;
    }
    const double piv = A[j][j];
    det = det * piv;
    if (piv == 0.0) break;
    for (int i = j + 1; i < m; ++i) {
      const double l = A[i][j] / piv;
      for (int k = j + 1; k < m; ++k) A[i][k] = A[i][k] - l * A[j][k];
      b[i] = b[i] - l * b[j];
    }
  }
  if (!(fabs(det) > OKO_TOL)) return OKO_EUNSUP;
  for (int i = m - 1; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < m; ++k) s = s - A[i][k] * c[k];
    c[i] = s / A[i][i];
  }
  int o = 0;
  for (int k = 0; k < m; ++k)
    if (!(0.0 - OKO_TOL <= c[k] && c[k] <= 1.0 + OKO_TOL)) o = 1;
  *out = o;
  return OKO_OK;
}

/* interp_cube (ndgrid.F90:636-665): the first tetrahedron containing xi;
 * c (2^n corner weights) when inside.  px: n x 2^n corner coordinates. */
static int interp_cube(int n, const double* tc, const double* px, const double* xi, int* out,
                       double* c) {
  const int twon = 1 << n, m = n + 1, nbth = factorial(n) * (1 << (n - 1));
  double X[4 * 5], coeff[5];
  for (int l = 0; l < nbth; ++l) {
    const double* t = tc + (size_t)twon * m * l;
    for (int d = 0; d < n; ++d)
      for (int k = 0; k < m; ++k) {
        double s = 0.0;
        for (int j = 0; j < twon; ++j) s = s + px[d + n * j] * t[j + twon * k];
        X[d + n * k] = s;
      }
    int o = 1;
    const int rc = interp_tetrahedron(n, X, xi, &o, coeff);
    if (rc) return rc;
    if (!o) {
      if (c)
        for (int j = 0; j < twon; ++j) {
          double s = 0.0;
          for (int k = 0; k < m; ++k) s = s + t[j + twon * k] * coeff[k];
          c[j] = s;
        }
      *out = 0;
      return OKO_OK;
    }
  }
  *out = 1;
  return OKO_OK;
}

/* ---------------------------------------------------------------------- */
/* The grid: n dims, shape g[4], node coordinates coord[d][lin] (lin =     */
/* i0 + g0 (i1 + g1 (i2 + g2 i3)), 0-based), mask (1 = masked) or NULL.    */
/* ---------------------------------------------------------------------- */
typedef struct {
  int n;
  int64_t g[4];
  const double* coord[4];
  const uint8_t* mask;
  double* tc;
} Grid;

static int64_t lin_index(const Grid* G, const int64_t* ind) {
  int64_t l = 0, off = 1;
  for (int d = 0; d < G->n; ++d) {
    l += ind[d] * off;
    off *= G->g[d];
  }
  return l;
}

/* Corner points of the cell with lower corner ind (InCube,
 * ndgrid_inc.F90:304-367: index min(ind+1, g-1) along set bits). */
static void corners(const Grid* G, const int64_t* ind, double* px, int64_t* cidx) {
  const int n = G->n, twon = 1 << n;
  for (int j = 0; j < twon; ++j) {
    int64_t p[4];
    for (int k = 0; k < n; ++k) {
      p[k] = ind[k];
      if ((j >> k) & 1) p[k] = ind[k] + 1 < G->g[k] - 1 ? ind[k] + 1 : G->g[k] - 1;
    }
    const int64_t l = lin_index(G, p);
    if (cidx) cidx[j] = l;
    for (int d = 0; d < n; ++d) px[d + n * j] = G->coord[d][l];
  }
}

/* The search tree of locate_databox (ndgrid_inc.F90:534-990): the box
 * [imin, imax] of node indices, bounded by the smallest and largest
 * coordinates of its nodes (search_boundarybox); a box with a side above 1
 * is split at im = (imin + imax) / 2 into the sub-boxes m = 1 .. 2^n (bit d
 * of m-1 set: [imin, im] along d, else [im, imax]) that have a side >= 1;
 * a box with every side <= 1 is a cell.  The sub-boxes cover their parent's
 * nodes, so its bounds are those of its sub-boxes. */
typedef struct {
  int64_t imin[4], imax[4];
  double xmin[4], xmax[4];
  int64_t child0;
  int nchild;
} Box;

typedef struct {
  Box* b;
  int64_t n, cap;
} Tree;

static int64_t tree_add(Tree* T) {
  if (T->n == T->cap) {
    int64_t c = T->cap ? 2 * T->cap : 1024;
    Box* nb = (Box*)realloc(T->b, sizeof(Box) * (size_t)c);
    if (!nb) return -1;
    T->b = nb;
    T->cap = c;
  }
  return T->n++;
}

static int tree_build(const Grid* G, Tree* T, int64_t id) {
  const int n = G->n;
  Box* B = &T->b[id];
  int split = 0;
  for (int d = 0; d < n; ++d)
    if (B->imax[d] - B->imin[d] > 1) split = 1;
  if (!split) {
    B->nchild = 0;
    B->child0 = -1;
    double xmin[4], xmax[4];
    for (int d = 0; d < n; ++d) {
      xmin[d] = HUGE_VAL;
      xmax[d] = -HUGE_VAL;
    }
    int64_t cnt[4], tot = 1;
    for (int d = 0; d < n; ++d) {
      cnt[d] = B->imax[d] - B->imin[d] + 1;
      tot *= cnt[d];
    }
    for (int64_t m = 0; m < tot; ++m) {
      int64_t p[4], r = m;
      for (int d = 0; d < n; ++d) {
        p[d] = B->imin[d] + r % cnt[d];
        r /= cnt[d];
      }
      const int64_t l = lin_index(G, p);
      for (int d = 0; d < n; ++d) {
        const double x = G->coord[d][l];
        if (x < xmin[d]) xmin[d] = x;
        if (x > xmax[d]) xmax[d] = x;
      }
    }
    memcpy(B->xmin, xmin, sizeof xmin);
    memcpy(B->xmax, xmax, sizeof xmax);
    return OKO_OK;
  }
  int64_t imin[4], imax[4], im[4];
  memcpy(imin, B->imin, sizeof imin);
  memcpy(imax, B->imax, sizeof imax);
  for (int d = 0; d < n; ++d) im[d] = (imin[d] + imax[d]) / 2;
  int64_t kids[16];
  int nk = 0;
  for (int mm = 0; mm < (1 << n); ++mm) {
    int64_t smin[4], smax[4];
    int keep = 0;
    for (int d = 0; d < n; ++d) {
      if ((mm >> d) & 1) {
        smin[d] = imin[d];
        smax[d] = im[d];
      } else {
        smin[d] = im[d];
        smax[d] = imax[d];
      }
      if (smax[d] - smin[d] >= 1) keep = 1;
    }
    if (!keep) continue;
    const int64_t k = tree_add(T);
    if (k < 0) return OKO_ENOMEM;
    memcpy(T->b[k].imin, smin, sizeof smin);
    memcpy(T->b[k].imax, smax, sizeof smax);
    kids[nk++] = k;
  }
  /* children are contiguous (added in order above) */
  T->b[id].child0 = kids[0];
  T->b[id].nchild = nk;
  for (int q = 0; q < nk; ++q) {
    const int rc = tree_build(G, T, kids[q]);
    if (rc) return rc;
  }
  B = &T->b[id];
  for (int d = 0; d < n; ++d) {
    double lo = HUGE_VAL, hi = -HUGE_VAL;
    for (int q = 0; q < nk; ++q) {
      const Box* C = &T->b[B->child0 + q];
      if (C->xmin[d] < lo) lo = C->xmin[d];
      if (C->xmax[d] > hi) hi = C->xmax[d];
    }
    B->xmin[d] = lo;
    B->xmax[d] = hi;
  }
  return OKO_OK;
}

/* locate_databox: depth first in sub-box order; the first cell whose
 * tetrahedra contain xi (InCube).  Returns 1 and its lower corner. */
static int locate(const Grid* G, const Tree* T, int64_t id, const double* xi, int64_t* ind,
                  int* err) {
  const int n = G->n;
  const Box* B = &T->b[id];
  for (int d = 0; d < n; ++d)
    if (xi[d] < B->xmin[d] || xi[d] > B->xmax[d]) return 0;
  if (B->nchild == 0) {
    double px[4 * 16];
    corners(G, B->imin, px, 0);
    int out = 1;
    const int rc = interp_cube(n, G->tc, px, xi, &out, 0);
    if (rc) {
      *err = rc;
      return 0;
    }
    if (out) return 0;
    for (int d = 0; d < n; ++d) ind[d] = B->imin[d];
    return 1;
  }
  for (int q = 0; q < B->nchild; ++q) {
    if (locate(G, T, B->child0 + q, xi, ind, err)) return 1;
    if (*err) return 0;
  }
  return 0;
}

static int tree_init(const Grid* G, Tree* T) {
  T->b = 0;
  T->n = T->cap = 0;
  const int64_t r = tree_add(T);
  if (r < 0) return OKO_ENOMEM;
  for (int d = 0; d < 4; ++d) {
    T->b[r].imin[d] = 0;
    T->b[r].imax[d] = d < G->n ? G->g[d] - 1 : 0;
  }
  return tree_build(G, T, r);
}

/* cinterp (ndgrid.F90:1183-1257) at xi: nbp = 2^n corners (0 if out of the
 * grid or next to a masked node), their linear node indices and weights. */
static int cinterp(const Grid* G, const Tree* T, const double* xi, int64_t* cidx, double* coeff,
                   int* nbp) {
  const int twon = 1 << G->n;
  int64_t ind[4];
  *nbp = 0;
  int err = 0;
  if (!locate(G, T, 0, xi, ind, &err)) return err;
  double px[4 * 16];
  corners(G, ind, px, cidx);
  if (G->mask)
    for (int j = 0; j < twon; ++j)
      if (G->mask[cidx[j]]) return OKO_OK;
  int out = 1;
  const int rc = interp_cube(G->n, G->tc, px, xi, &out, coeff);
  if (rc) return rc;
  *nbp = twon;
  return OKO_OK;
}

/* genObservationOper + packSparseMatrix + the zone permutation of the state
 * columns (assimilation.F90:2471-2656, 4132-4208, 2404-2409) for one model
 * variable per observation:
 *   nvar grids: ndim[v], shape[4 v + d], coord[4 v + d] (node coordinates),
 *     mask[v] (1 = land, may be NULL), state[v] (node -> 0-based column of H
 *     in the permuted state, -1 = not in the state; may be NULL = the node
 *     index itself);
 *   obs_var[l] (-1: unknown variable), ox/oy/oz/ot (may be NULL past ndim).
 * Writes row_ptr[m+1], and up to cap entries col/val (2^n per observation
 * inside its grid, corner order of cinterp) and valid[l] (0: OAK excludes the
 * observation, validobs of loadObservationOper).  Returns OKO_EUNSUP for a
 * degenerate cell, OKO_EARG for a bad argument or cap too small. */
int oko_obs_operator(int nvar, const int* ndim, const int64_t* shape, const double* const* coord,
                     const uint8_t* const* mask, const int64_t* const* state, int64_t m,
                     const int32_t* obs_var, const double* ox, const double* oy, const double* oz,
                     const double* ot, int64_t cap, int64_t* row_ptr, int32_t* col, double* val,
                     uint8_t* valid) {
  if (nvar < 0 || m < 0) return OKO_EARG;
  Grid* G = (Grid*)calloc((size_t)(nvar > 0 ? nvar : 1), sizeof(Grid));
  Tree* T = (Tree*)calloc((size_t)(nvar > 0 ? nvar : 1), sizeof(Tree));
  if (!G || !T) { free(G); free(T); return OKO_ENOMEM; }
  int rc = OKO_OK;
  for (int v = 0; v < nvar && !rc; ++v) {
    const int n = ndim[v];
    if (n < 1 || n > 4) { rc = OKO_EARG; break; }
    G[v].n = n;
    for (int d = 0; d < 4; ++d) G[v].g[d] = d < n ? shape[4 * v + d] : 1;
    for (int d = 0; d < n; ++d) {
      if (G[v].g[d] < 2) rc = OKO_EUNSUP;
      G[v].coord[d] = coord[4 * v + d];
    }
    G[v].mask = mask ? mask[v] : 0;
    const int nb = factorial(n) * (1 << (n - 1));
    G[v].tc = (double*)malloc(sizeof(double) * (size_t)(1 << n) * (n + 1) * nb);
    if (!G[v].tc) { rc = OKO_ENOMEM; break; }
    oko_tetrahedra(n, G[v].tc);
    if (!rc) rc = tree_init(&G[v], &T[v]);
  }
  const double* oc[4] = {ox, oy, oz, ot};
  int64_t nz = 0;
  row_ptr[0] = 0;
  for (int64_t l = 0; l < m && !rc; ++l) {
    const int v = obs_var[l];
    valid[l] = 0;
    if (v >= 0 && v < nvar) {
      const Grid* g = &G[v];
      double xi[4];
      for (int d = 0; d < g->n; ++d) xi[d] = oc[d][l];
      int64_t cidx[16];
      double c[16];
      int nbp = 0;
      rc = cinterp(g, &T[v], xi, cidx, c, &nbp);
      if (rc) break;
      if (nbp > 0) {
        int ok = 1;
        for (int j = 0; j < nbp; ++j) {
          const int64_t s = state && state[v] ? state[v][cidx[j]] : cidx[j];
          if (s < 0) { ok = 0; continue; }
          if (nz >= cap) { rc = OKO_EARG; break; }
          col[nz] = (int32_t)s;
          val[nz] = c[j];
          ++nz;
          /* validobs = val2 of the observation's last entry (4171) */
          ok = 1;
        }
        valid[l] = (uint8_t)ok;
      }
    }
    row_ptr[l + 1] = nz;
  }
  for (int v = 0; v < nvar; ++v) {
    free(G[v].tc);
    free(T[v].b);
  }
  free(G);
  free(T);
  return rc;
}

/* One point through cinterp (for the tests): corner node indices and
 * weights; returns nbp (0 = out) or -rc. */
int oko_cinterp(int n, const int64_t* shape, const double* const* coord, const uint8_t* mask,
                const double* xi, int64_t* cidx, double* coeff) {
  Grid G;
  memset(&G, 0, sizeof G);
  G.n = n;
  for (int d = 0; d < 4; ++d) G.g[d] = d < n ? shape[d] : 1;
  for (int d = 0; d < n; ++d) G.coord[d] = coord[d];
  G.mask = mask;
  const int nb = factorial(n) * (1 << (n - 1));
  G.tc = (double*)malloc(sizeof(double) * (size_t)(1 << n) * (n + 1) * nb);
  if (!G.tc) return -OKO_ENOMEM;
  oko_tetrahedra(n, G.tc);
  Tree T;
  int rc = tree_init(&G, &T);
  int nbp = 0;
  if (!rc) rc = cinterp(&G, &T, xi, cidx, coeff, &nbp);
  free(G.tc);
  free(T.b);
  return rc ? -rc : nbp;
}
