"""Observation operator from coordinates on the GPU (oakg_obs_operator:
genObservationOper + packSparseMatrix with OAK's databox search and
hyper-tetrahedral weights) against the oracle's restatement: indices, weights
and validity bit for bit, on curvilinear grids in 1 to 4 dimensions with land
masks, points on grid lines and nodes, points outside and unknown variables;
then the analysis with the generated operator.  Needs an MI355X."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.kfcheck import rel_err

pytestmark = pytest.mark.gpu


def _curvilinear(shape, seed, amp=0.08):
    g = np.random.default_rng(seed)
    n = len(shape)
    idx = np.meshgrid(*[np.arange(s, dtype=float) for s in shape], indexing="ij")
    coords = []
    for d in range(n):
        c = 1.5 * idx[d] + 0.7 * d
        for e in range(n):
            if e != d:
                c = c + amp * np.sin(1.3 * idx[e] + d + seed)
        coords.append(c)
    return coords


def _points(coords, m, g, on_lines=0.2, outside=0.1):
    """Random points inside cells, a share exactly on grid lines / nodes,
    and some outside the grid."""
    n = len(coords)
    shape = coords[0].shape
    pts = np.zeros((n, m))
    for l in range(m):
        u = g.uniform()
        lo = [int(g.integers(0, s - 1)) for s in shape]
        if u < on_lines:
            node = tuple(lo[d] + int(g.integers(0, 2)) for d in range(n))
            pts[:, l] = [coords[d][node] for d in range(n)]
            if g.uniform() < 0.5:  # on an edge: halfway to a neighbour along one dim
                dd = int(g.integers(0, n))
                nb = list(node)
                nb[dd] = min(nb[dd] + 1, shape[dd] - 1)
                pts[:, l] = 0.5 * (pts[:, l] + np.array([coords[d][tuple(nb)] for d in range(n)]))
        elif u < on_lines + outside:
            pts[:, l] = [coords[d].min() - 1.0 - g.uniform() for d in range(n)]
        else:
            w = g.dirichlet(np.ones(2 ** n))
            for j in range(2 ** n):
                node = tuple(lo[d] + ((j >> d) & 1) for d in range(n))
                for d in range(n):
                    pts[d, l] += w[j] * coords[d][node]
    return pts


@pytest.mark.parametrize("shape", [(17,), (23, 19), (9, 8, 7), (5, 4, 4, 3)])
def test_obs_operator_bit_exact_vs_oracle(shape):
    from oak_amd import assimilation as A
    g = np.random.default_rng(len(shape))
    coords = _curvilinear(shape, seed=len(shape))
    mask = g.uniform(size=shape) < 0.04
    nodes = int(np.prod(shape))
    # the state holds the sea nodes, in a permuted order (zone layout)
    perm = g.permutation(nodes)
    state = np.where(mask.ravel(order="F"), -1, perm).reshape(shape, order="F")
    m = 400
    pts = _points(coords, m, g)
    obs_var = np.zeros(m, np.int32)
    obs_var[::37] = -1  # unknown variable
    grids = [dict(coords=coords, mask=mask, state=state)]
    rp, col, val, valid = A.obs_operator(grids, obs_var, list(pts))
    rpo, colo, valo, valido = O.obs_operator(grids, obs_var, list(pts))
    np.testing.assert_array_equal(rp, rpo)
    np.testing.assert_array_equal(col, colo)
    np.testing.assert_array_equal(val, valo)  # bit for bit
    np.testing.assert_array_equal(valid, valido)
    rows = np.diff(rp)
    assert (rows == 2 ** len(shape)).sum() > m // 2 and (rows == 0).sum() > 20


def test_obs_operator_two_variables_and_analysis():
    """Two model variables (a 2-D surface field and a 3-D one) in one
    state; the observations of both through the generated H, then the
    ensemble analysis with that H against the oracle's with its own H."""
    from oak_amd import assimilation as A
    g = np.random.default_rng(7)
    nx, ny, nz = 14, 11, 4
    c2 = _curvilinear((nx, ny), seed=3, amp=0.05)
    z3 = np.broadcast_to(np.array([0.0, 10.0, 25.0, 50.0]), (nx, ny, nz))
    c3 = [np.repeat(c2[0][:, :, None], nz, 2), np.repeat(c2[1][:, :, None], nz, 2), z3]
    n2, n3 = nx * ny, nx * ny * nz
    state2 = np.arange(n2).reshape((nx, ny), order="F")
    state3 = (n2 + np.arange(n3)).reshape((nx, ny, nz), order="F")
    grids = [dict(coords=c2, state=state2), dict(coords=c3, state=state3)]
    m2, m3 = 60, 90
    p2 = _points(c2, m2, g, on_lines=0.1, outside=0.05)
    p3 = _points(c3, m3, g, on_lines=0.1, outside=0.05)
    ox = np.concatenate([p2[0], p3[0]]); oy = np.concatenate([p2[1], p3[1]])
    oz = np.concatenate([np.zeros(m2), p3[2]])
    obs_var = np.concatenate([np.zeros(m2, np.int32), np.ones(m3, np.int32)])
    rp, col, val, valid = A.obs_operator(grids, obs_var, [ox, oy, oz])
    rpo, colo, valo, valido = O.obs_operator(grids, obs_var, [ox, oy, oz])
    np.testing.assert_array_equal(col, colo)
    np.testing.assert_array_equal(val, valo)
    np.testing.assert_array_equal(valid, valido)
    # analysis: one zone per horizontal point (the SSH-like variable and its column)
    n = n2 + n3
    N, m = 10, m2 + m3
    E = np.asfortranarray(g.standard_normal((n, N)))
    Hi = np.repeat(np.arange(m), np.diff(rp)).astype(np.int32)
    yo = g.standard_normal(m)
    r2 = np.full(m, 0.3)
    # one zone holding the whole state, all observations (validobs excluded)
    ex = (~valid).astype(np.uint8)
    geo = A.Geometry(np.array([0, n]), np.zeros((1, 4)), np.array([1e9]), np.array([1e9]),
                     ox, oy)
    Ea = A.ensemble_analysis(geo, (rp, col, val), np.zeros(m), yo, r2, E, excluded=ex)
    Eo, _, _ = O.ensemble_analysis(dict(zone_ptr=np.array([0, n]), zpos=np.zeros((1, 4)),
                                        L=np.array([1e9]), cutoff=np.array([1e9]), metric=0,
                                        loctype=1, ox=ox, oy=oy),
                                   (Hi, colo, valo), np.zeros(m), yo, r2, E, dex=1.0 - ex)
    xf = E.mean(axis=1)
    assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < 1e-10


def test_obs_operator_degenerate_dimension_unsupported():
    from oak_amd import OakError
    from oak_amd import assimilation as A
    x = np.arange(5.0)[:, None] * np.ones((1, 1))
    y = np.zeros((5, 1))
    with pytest.raises(OakError) as e:
        A.obs_operator([dict(coords=[x, y])], np.zeros(3, np.int32),
                       [np.array([1.0, 2.0, 3.0]), np.zeros(3)])
    assert e.value.code == 5
