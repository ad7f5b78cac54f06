"""GPU parity cases beyond the basic grid: observations that cover only part
of the zone domain (bins outside the observations' bounding box), the
vertical/time localisation and the spherical metrics compared set for set
with OAK's selectObservations, OAK's per-element maxCorrection, a config-1
miniature with thousands of local observations per zone at N = 64, and the
likelihood adjoint at N = 32 (the kernel instantiation config 3 runs).
Needs an MI355X."""
import numpy as np
import pytest

from oak_amd.errors import ensemble_errors
from oracle import oracle as O
from tests.kfcheck import rel_err

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL_MIXED = 1e-4


@pytest.fixture(scope="module", autouse=True)
def _oracle_lapack():
    O.load_lapack()
    yield
    O.set_backend(0)


def _geo_dict(P, **kw):
    d = dict(zone_ptr=P.zone_ptr, zpos=P.zpos, L=P.corr_len, cutoff=P.max_len, metric=P.metric,
             loctype=P.loctype, ox=P.ox, oy=P.oy)
    d.update(kw)
    return d


def _geometry(P, **kw):
    from oak_amd.assimilation import Geometry
    a = dict(metric=P.metric, loctype=P.loctype)
    a.update(kw)
    return Geometry(P.zone_ptr, P.zpos, P.corr_len, P.max_len, P.ox, P.oy, **a)


def _check_sets(sp, si, sw, zpos, L, cut, metric, loctype, ox, oy, oz=None, ot=None):
    """Each zone's GPU list equals OAK's relevant set (index order), weights
    to rounding; returns the number of non-empty and empty zones.  The
    great-circle distance R acos(...) of the spherical metrics amplifies the
    rounding of sin/cos (device vs host libm) near d = 0, to ~1e-12 of d for
    100 km: its weights are compared at 1e-10."""
    rtol = 1e-10 if (metric != 0 and loctype == 1) else 1e-15
    full = empty = 0
    for z in range(len(sp) - 1):
        w, rel, cnt = O.select(metric, loctype, zpos[z], L[z], cut[z], ox, oy, oz, ot)
        idx = np.nonzero(rel)[0]
        assert sp[z + 1] - sp[z] == cnt, z
        np.testing.assert_array_equal(si[sp[z]:sp[z + 1]], idx)
        np.testing.assert_allclose(sw[sp[z]:sp[z + 1]], w[idx], rtol=rtol, atol=0)
        full += cnt > 0
        empty += cnt == 0
    return full, empty


def test_selection_observations_in_part_of_the_domain():
    """Observations only in the middle third of the grid: zones left, right,
    above and below the observations' bounding box, several cutoffs away, take
    the binned path's empty range; the others match OAK's sets exactly, and
    the analysis matches the oracle."""
    from oak_amd import assimilation as A
    from oak_amd import synth
    P = synth.config0(N=12, seed_cfg=41)
    g = np.random.default_rng(41)
    keep = (P.ox > 24) & (P.ox < 40) & (P.oy > 28) & (P.oy < 36)
    assert keep.sum() > 10
    idx = np.nonzero(keep)[0]
    remap = -np.ones(P.m, dtype=np.int64); remap[idx] = np.arange(len(idx))
    hk = keep[P.Hi]
    P.Hi = remap[P.Hi[hk]].astype(np.int32); P.Hj = P.Hj[hk]; P.Hs = P.Hs[hk]
    P.ox, P.oy, P.yo, P.rmse2, P.Hshift = (a[idx] for a in (P.ox, P.oy, P.yo, P.rmse2, P.Hshift))
    P.m = len(idx)
    # smaller cutoffs than the grid: most zones see no observation
    P.max_len = np.full(P.nzones, 6.0) + g.uniform(0, 2, P.nzones)
    P.corr_len = P.max_len / 3.0
    sp, si, sw = A.select_observations(_geometry(P), P.yo, P.rmse2)
    full, empty = _check_sets(sp, si, sw, P.zpos, P.corr_len, P.max_len, 0, 1, P.ox, P.oy)
    assert full > 100 and empty > 2000, (full, empty)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    Ea = A.ensemble_analysis(_geometry(P), H, P.Hshift, P.yo, P.rmse2, P.E)
    Eo, _, cnt = O.ensemble_analysis(_geo_dict(P), (P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2,
                                     P.E)
    xf = P.E.mean(axis=1)
    assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < TOL64
    np.testing.assert_array_equal(np.diff(sp), cnt)


def _sphere_problem(seed, n_lon=18, n_lat=11, m=150):
    g = np.random.default_rng(seed)
    lon, lat = np.meshgrid(np.linspace(-40, 40, n_lon), np.linspace(-70, 75, n_lat),
                           indexing="ij")
    nz = n_lon * n_lat
    zpos = np.zeros((nz, 4))
    zpos[:, 0] = lon.ravel(order="F"); zpos[:, 1] = lat.ravel(order="F")
    zpos[:, 2] = g.uniform(0, 500, nz); zpos[:, 3] = g.uniform(0, 30, nz)
    ox = g.uniform(-45, 45, m); oy = g.uniform(-80, 80, m)
    # a few observations exactly at zone positions and on the cutoff
    ox[:5] = zpos[:5, 0]; oy[:5] = zpos[:5, 1]
    oz = g.uniform(0, 500, m); ot = g.uniform(0, 30, m)
    return g, zpos, ox, oy, oz, ot


@pytest.mark.parametrize("metric,loctype", [(1, 1), (2, 1), (0, 2), (0, 3), (1, 2), (2, 3)])
def test_selection_metrics_and_loctypes_set_for_set(metric, loctype):
    """selectObservations (assimilation.F90:3635-3771) for the spherical
    metrics (great circle and its approximation, metres) and for vertical
    (|dz|) and time (|dt|) localisation: the GPU's relevant sets equal OAK's
    zone by zone, weights to rounding, with the cutoff hit exactly by some
    observations."""
    from oak_amd import assimilation as A
    g, zpos, ox, oy, oz, ot = _sphere_problem(10 * metric + loctype)
    nz, m = len(zpos), len(ox)
    if loctype == 1:
        L = np.full(nz, 900e3) if metric else np.full(nz, 8.0)
        cut = 2.0 * L
    elif loctype == 2:
        L = np.full(nz, 60.0); cut = np.full(nz, 120.0)
        # an observation exactly at the cutoff of zone 0 (inclusive test;
        # values whose difference is exact in binary)
        zpos[0, 2] = 100.0
        oz[7] = 220.0
    else:
        L = np.full(nz, 3.0); cut = np.full(nz, 6.0)
        zpos[0, 3] = 10.0
        ot[7] = 4.0
    geo = A.Geometry(np.arange(nz + 1), zpos, L, cut, ox, oy, oz, ot, metric=metric,
                     loctype=loctype)
    yo = g.standard_normal(m); r2 = np.full(m, 0.3)
    sp, si, sw = A.select_observations(geo, yo, r2)
    full, empty = _check_sets(sp, si, sw, zpos, L, cut, metric, loctype, ox, oy, oz, ot)
    assert full > 20
    if loctype in (2, 3):
        assert 7 in si[sp[0]:sp[1]]


@pytest.mark.parametrize("loctype", [2, 3])
def test_vertical_time_localisation_analysis_vs_oracle(loctype):
    """The analysis with loctype 2 / 3 through the ensemble path."""
    from oak_amd import assimilation as A
    g, zpos, ox, oy, oz, ot = _sphere_problem(70 + loctype, m=90)
    nz, m, N = len(zpos), len(ox), 14
    E = np.asfortranarray(g.standard_normal((nz, N)))
    Hi = np.arange(m, dtype=np.int32); Hj = g.integers(0, nz, m).astype(np.int32)
    Hs = np.ones(m)
    yo = g.standard_normal(m); r2 = g.uniform(0.2, 0.6, m)
    L, cut = (np.full(nz, 60.0), np.full(nz, 120.0)) if loctype == 2 else \
        (np.full(nz, 3.0), np.full(nz, 6.0))
    geo = A.Geometry(np.arange(nz + 1), zpos, L, cut, ox, oy, oz, ot, metric=0, loctype=loctype)
    H = A.coo_to_csr(m, Hi, Hj, Hs)
    Ea = A.ensemble_analysis(geo, H, np.zeros(m), yo, r2, E)
    Eo, _, _ = O.ensemble_analysis(dict(zone_ptr=np.arange(nz + 1), zpos=zpos, L=L, cutoff=cut,
                                        metric=0, loctype=loctype, ox=ox, oy=oy, oz=oz, ot=ot),
                                   (Hi, Hj, Hs), np.zeros(m), yo, r2, E)
    xf = E.mean(axis=1)
    assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < TOL64


@pytest.fixture(scope="module")
def two_var():
    """Ocean columns with two clamps: T (levels) at 0.05, the other
    variables and SSH at 0.3 (OAK's Model.maxCorrection per variable,
    assimilation.F90:462-476)."""
    from oak_amd import synth
    P = synth.column_ocean(10, 8, 3, 24, 80, 5, seed_cfg=620, hlen=3, vlen=2, L=2.0,
                           cutoff=5.0)
    s = 4 * 3 + 1
    per_zone = np.full(s, 0.3); per_zone[:3] = 0.05
    mc = np.tile(per_zone, P.nzones)
    return P, mc


@pytest.mark.parametrize("N", [24, 128])
def test_per_element_max_correction_vs_oracle(two_var, N):
    from oak_amd import assimilation as A
    from oak_amd import synth
    P, mc = two_var
    if N != P.N:
        P = synth.column_ocean(10, 8, 3, N, 80, 5, seed_cfg=620, hlen=3, vlen=2, L=2.0,
                               cutoff=5.0)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    Eo, xo, _ = O.ensemble_analysis(_geo_dict(P), (P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2,
                                    P.E, inflation=1.1, max_correction=mc)
    xf = P.E.mean(axis=1)
    clamped = np.abs(np.abs(xo - xf) - mc) < 1e-12
    assert clamped[mc == 0.05].sum() > 5 and clamped[mc == 0.3].sum() > 0
    Ea = A.ensemble_analysis(_geometry(P), H, P.Hshift, P.yo, P.rmse2, P.E, inflation=1.1,
                             max_correction=mc)
    assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < TOL64, N
    E32 = np.asfortranarray(P.E, dtype=np.float32)
    Ea32 = A.ensemble_analysis(_geometry(P), H, P.Hshift, P.yo, P.rmse2, E32, inflation=1.1,
                               max_correction=mc)
    e = ensemble_errors(Ea32, Eo, P.E)
    assert e["mean_err"] < TOL_MIXED and e["spread_err"] < TOL_MIXED, e


def test_per_element_max_correction_device_and_errors(two_var):
    """Device pointers give the host call's bits; a negative or NaN clamp is
    rejected with OAKG_EARG before anything runs."""
    import torch
    from oak_amd import OakError
    from oak_amd import assimilation as A
    P, mc = two_var
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    Ea = A.ensemble_analysis(_geometry(P), H, P.Hshift, P.yo, P.rmse2, P.E, max_correction=mc)
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(dev, dt)
    gd = A.Geometry(t(P.zone_ptr, torch.int64), t(P.zpos), t(P.corr_len), t(P.max_len), t(P.ox),
                    t(P.oy), metric=0, loctype=1)
    Hd = (t(H[0], torch.int64), t(H[1], torch.int32), t(H[2]))
    Ed = A.ensemble_analysis(gd, Hd, t(P.Hshift), t(P.yo), t(P.rmse2), t(P.E.T).t(),
                             max_correction=t(mc))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(Ed.cpu().numpy(), Ea)
    for bad in (-1.0, np.nan):
        m2 = mc.copy(); m2[5] = bad
        with pytest.raises(OakError) as e:
            A.ensemble_analysis(_geometry(P), H, P.Hshift, P.yo, P.rmse2, P.E, max_correction=m2)
        assert e.value.code == 1


@pytest.fixture(scope="module")
def cfg1_mini():
    """Config 1 in miniature: ocean columns (40 x 40, 2 levels, 9 elements
    per zone), N = 64, 6000 SSH observations along tracks and 40 profiles,
    L = 6, cutoff 18 cells: 700-4700 local observations per zone (the
    Gram-dominated regime of config 1, where p_z ~ 3000)."""
    from oak_amd import synth
    return synth.column_ocean(40, 40, 2, 64, 6000, 40, seed_cfg=401, hlen=8, vlen=2, L=6.0,
                              cutoff=18.0)


def test_config1_miniature_vs_oracle(cfg1_mini):
    from oak_amd import assimilation as A
    P = cfg1_mini
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    Eo, _, cnt = O.ensemble_analysis(_geo_dict(P), (P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2,
                                     P.E)
    assert np.median(cnt) > 2000 and cnt.max() > 4000
    Ea = A.ensemble_analysis(_geometry(P), H, P.Hshift, P.yo, P.rmse2, P.E)
    xf = P.E.mean(axis=1)
    assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < TOL64
    E32 = np.asfortranarray(P.E, dtype=np.float32)
    Ea32 = A.ensemble_analysis(_geometry(P), H, P.Hshift, P.yo, P.rmse2, E32)
    e = ensemble_errors(Ea32, Eo, P.E)
    assert e["mean_err"] < TOL_MIXED and e["spread_err"] < TOL_MIXED, e


@pytest.fixture(scope="module")
def nll32():
    """Config 3's ensemble size N = 32 (the NP = 32 kernels of the
    likelihood), on a 24 x 24 grid with 150 observations."""
    from oak_amd import synth
    from oak_amd.nll import problem_from_ensemble
    P = synth.config3(nx=24, N=32, m=150, seed_cfg=34)
    return P, problem_from_ensemble(P)


def test_nll_n32_vs_reference_and_differences(nll32):
    import torch
    from oak_amd.nll import evaluate
    from tests.nll_ref import nll_reference, selection
    P, prob = nll32
    nz = P.nzones
    ls = 0.5 * np.log(P.rmse2) + 0.2
    L, a = 4.0, 1.2
    J, gth, gLz, ga = evaluate(prob, np.exp(2 * ls), np.full(nz, L), a)
    sel = selection(P, prob.HSf)
    lsr = torch.as_tensor(ls).clone().requires_grad_(True)
    Lr = torch.tensor(L, dtype=torch.float64, requires_grad=True)
    ar = torch.tensor(a, dtype=torch.float64, requires_grad=True)
    Jr = nll_reference(sel, prob.HSf, prob.Hxf, prob.yo, lsr, Lr, ar)
    Jr.backward()
    assert abs(J - float(Jr)) <= 1e-9 * abs(float(Jr))
    gthr = lsr.grad.numpy()
    assert np.abs(gth - gthr).max() <= 1e-8 * np.abs(gthr).max()
    assert abs(gLz.sum() - float(Lr.grad)) <= 1e-8 * max(1.0, abs(float(Lr.grad)))
    assert abs(ga - float(ar.grad)) <= 1e-8 * max(1.0, abs(float(ar.grad)))
    h = 1e-6
    f = lambda ls_, L_, a_: evaluate(prob, np.exp(2 * ls_), np.full(nz, L_), a_)[0]
    tol = 1e-6 * max(1.0, np.abs(gth).max())
    for l in (0, 3, 77, 149):
        e = np.zeros_like(ls); e[l] = h
        fd = (f(ls + e, L, a) - f(ls - e, L, a)) / (2 * h)
        assert abs(fd - gth[l]) <= tol, (l, fd, gth[l])
    fdL = (f(ls, L + h, a) - f(ls, L - h, a)) / (2 * h)
    fda = (f(ls, L, a + h) - f(ls, L, a - h)) / (2 * h)
    assert abs(fdL - gLz.sum()) <= 1e-6 * max(1.0, abs(fdL))
    assert abs(fda - ga) <= 1e-6 * max(1.0, abs(fda))
