"""Pins the oracle's restatement of OAK's observation-operator generation
(oracle/oak_oracle_interp.c: split, interp_tetrahedron, interp_cube, the
databox search and cinterp of ndgrid.F90 / ndgrid_inc.F90, and
genObservationOper + packSparseMatrix of assimilation.F90) against
test/test_ndgrid.F90's known answers and against what the mathematics fixes:
linear fields are interpolated exactly, the tetrahedra partition the cube,
the weights are a convex combination, and the search order of the databox
tree picks the upper of two cells that share a face."""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as O


def _simplex_volume(V):
    n = V.shape[0]
    return abs(np.linalg.det(V[:, 1:] - V[:, :1])) / math.factorial(n)


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_tetrahedra_partition_the_cube(n):
    """n! 2^(n-1) simplices (init_basegrid, ndgrid.F90:886), each vertex a
    convex combination of the 2^n corners, the last vertex the centre of the
    cube (n >= 2; for n = 1 the segment itself); their volumes add up to the
    unit cube's."""
    tc = O.tetrahedra(n)
    nb = math.factorial(n) * 2 ** (n - 1)
    assert tc.shape == (2 ** n, n + 1, nb)
    assert np.all(tc >= 0) and np.allclose(tc.sum(axis=0), 1.0, atol=0, rtol=1e-15)
    corners = np.array([[(j >> d) & 1 for j in range(2 ** n)] for d in range(n)], float)
    vols = []
    for l in range(nb):
        V = corners @ tc[:, :, l]
        if n >= 2:
            np.testing.assert_allclose(V[:, -1], 0.5, rtol=1e-15)
        vols.append(_simplex_volume(V))
    assert min(vols) > 0
    assert abs(sum(vols) - 1.0) < 1e-13
    # the first vertices of each simplex are corners of the cube, chained
    # along edges (corner j, then a neighbour) as split builds them
    first = [tuple(tc[:, 0, l]) for l in range(nb)]
    assert all(sorted(f).count(1.0) == 1 for f in first)


def _interp(coords, f, xi, mask=None):
    r = O.cinterp(coords, xi, mask)
    if r is None:
        return None
    cidx, c = r
    return float(np.sum(c * f.ravel(order="F")[cidx])), cidx, c


def test_ndgrid_2d_known_answer():
    """test/test_ndgrid.F90:49-102: x = i + 1, y = j + 2 on 10 x 20 nodes,
    f = 2x + 4y; the middle of the domain is inside and f is reproduced."""
    m, n = 10, 20
    i, j = np.meshgrid(np.arange(1, m + 1), np.arange(1, n + 1), indexing="ij")
    x = i + 1.0
    y = j + 2.0
    f = 2 * x + 4 * y
    xi = np.array([x.sum() / (n * m), y.sum() / (n * m)])
    fi, cidx, c = _interp([x, y], f, xi)
    assert abs(fi - (2 * xi[0] + 4 * xi[1])) < 1e-6
    assert len(cidx) == 4 and abs(c.sum() - 1.0) < 1e-15


@pytest.mark.parametrize("shape", [(10,), (10, 20), (2, 2, 2), (2, 2, 2, 2), (5, 4, 3), (3, 4, 2, 3)])
def test_linear_fields_exact_on_curvilinear_grids(shape):
    """test_ndgrid_nd (test/test_ndgrid.F90:9-21) with a linear field, here on
    smoothly distorted (curvilinear) grids and random points inside: each
    simplex interpolates a linear function exactly, so every point gives
    f(xi) to rounding; weights are >= -tol and add up to 1."""
    g = np.random.default_rng(len(shape) * 10 + sum(shape))
    n = len(shape)
    idx = np.meshgrid(*[np.arange(s, dtype=float) for s in shape], indexing="ij")
    # distortion small enough to keep every cell convex
    coords = []
    for d in range(n):
        c = 1.5 * idx[d] + 0.7 * d
        for e in range(n):
            if e != d:
                c = c + 0.08 * np.sin(1.3 * idx[e] + d)
        coords.append(c)
    a = g.uniform(-2, 2, n)
    f = sum(a[d] * coords[d] for d in range(n)) + 0.5
    hits = 0
    for _ in range(60):
        # a random point inside a random cell: a convex combination of its corners
        lo = [int(g.integers(0, s - 1)) for s in shape]
        w = g.dirichlet(np.ones(2 ** n))
        xi = np.zeros(n)
        for jj, bits in enumerate(itertools.product([0, 1], repeat=n)):
            node = tuple(lo[d] + bits[::-1][d] for d in range(n))
            for d in range(n):
                xi[d] += w[jj] * coords[d][node]
        r = _interp(coords, f, xi)
        assert r is not None, xi
        fi, cidx, c = r
        assert abs(fi - (a @ xi + 0.5)) < 1e-11 * max(1, abs(fi))
        assert np.all(c >= -1e-8) and abs(c.sum() - 1.0) < 1e-13
        hits += 1
    assert hits == 60


def test_outside_and_masked_points():
    m, n = 6, 5
    i, j = np.meshgrid(np.arange(m, dtype=float), np.arange(n, dtype=float), indexing="ij")
    x, y = 2.0 * i, 3.0 * j
    assert O.cinterp([x, y], [-0.1, 4.0]) is None          # left of the grid
    assert O.cinterp([x, y], [4.0, 12.0 + 1e-6]) is None    # above it
    assert O.cinterp([x, y], [10.0, 12.0]) is not None      # the last node
    mask = np.zeros((m, n), bool)
    mask[2, 1] = True
    # cells touching node (2, 1): out; the others in
    assert O.cinterp([x, y], [3.0, 1.5], mask) is None
    assert O.cinterp([x, y], [5.0, 4.5], mask) is None
    assert O.cinterp([x, y], [7.0, 7.5], mask) is not None


def test_search_order_prefers_the_upper_cell():
    """A point on an interior grid line is inside both cells that share it;
    the databox search (ndgrid_inc.F90:921-990) visits the upper sub-box
    first at every split, so the cell found has its lower corner on the line
    (node indices of the first corner)."""
    m, n = 9, 7
    i, j = np.meshgrid(np.arange(m, dtype=float), np.arange(n, dtype=float), indexing="ij")
    x, y = 1.0 * i, 1.0 * j
    for k in range(1, m - 1):
        cidx, c = O.cinterp([x, y], [float(k), 2.4])
        assert cidx[0] % m == k, (k, cidx)
    for k in range(1, n - 1):
        cidx, c = O.cinterp([x, y], [3.3, float(k)])
        assert cidx[0] // m == k, (k, cidx)
    # on a node: the cell above and to the right
    cidx, c = O.cinterp([x, y], [4.0, 3.0])
    assert (cidx[0] % m, cidx[0] // m) == (4, 3)
    assert abs(c[0] - 1.0) < 1e-15 and np.all(np.abs(c[1:]) < 1e-15)


def test_obs_operator_assembly():
    """genObservationOper + packSparseMatrix (assimilation.F90:2471-2656,
    4132-4208): 2^n entries per observation inside its grid, in corner
    order, mapped to state columns; unknown variables and points outside the
    grid get no entry and are not valid (validobs); a node outside the state
    drops its entry, and validity follows the observation's last entry."""
    m1, n1 = 5, 4
    i, j = np.meshgrid(np.arange(m1, dtype=float), np.arange(n1, dtype=float), indexing="ij")
    g1 = dict(coords=[i, j])
    ii = np.arange(6, dtype=float)
    state2 = 100 + np.arange(6)
    state2[5] = -1
    g2 = dict(coords=[ii], state=state2)
    obs_var = np.array([0, 0, 1, -1, 1, 1, 0])
    ox = np.array([1.5, 9.0, 2.25, 0.0, 4.5, 0.5, 0.0])
    oy = np.array([2.5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    rp, col, val, valid = O.obs_operator([g1, g2], obs_var, [ox, oy])
    assert list(np.diff(rp)) == [4, 0, 2, 0, 1, 2, 4]
    assert list(valid) == [True, False, True, False, False, True, True]
    # observation 0: cell (1, 2) of grid 0, corners in cinterp order
    np.testing.assert_array_equal(col[:4], [1 + 5 * 2, 2 + 5 * 2, 1 + 5 * 3, 2 + 5 * 3])
    np.testing.assert_allclose(val[:4], [0.25, 0.25, 0.25, 0.25], rtol=1e-15)
    # observation 2 on grid 1 (1-D): nodes 2, 3 -> state 102, 103
    np.testing.assert_array_equal(col[4:6], [102, 103])
    np.testing.assert_allclose(val[4:6], [0.75, 0.25], rtol=1e-15)
    # observation 4: cell (4, 5), node 5 not in the state: one entry, not valid
    np.testing.assert_array_equal(col[6:7], [104])
