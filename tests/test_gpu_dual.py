"""Zones with few local observations solved in observation space
(OAKG_TUNE_DUAL, oak_amd/csrc/dual.hip): the p x p matrix S~ S~^T instead of
the N x N matrix I + S~^T S~.  The analysis is a matrix function of M, so
it is the same to rounding: against the oracle at the FP64 bound (1e-10 of
the largest increment / anomaly) and the mixed bound (1e-4), on workloads
whose zones have counts on both sides of the threshold; the amplitudes of
localise_obs = .false.; and bit-identical under the zone order and graph
replay.  Needs an MI355X."""
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


def _geo(P):
    from oak_amd.assimilation import Geometry
    return Geometry(P.zone_ptr, P.zpos, P.corr_len, P.max_len, P.ox, P.oy, metric=0, loctype=1)


def _problem(N, seed):
    from oak_amd import synth
    # sparse tracks: local counts from 0 to about 90 at cutoff 5
    return synth.column_ocean(22, 16, 2, N, 160, 3, seed_cfg=seed, hlen=3, vlen=2, L=2.0,
                              cutoff=5.0)


def _oracle(P, infl=1.0, maxc=np.inf):
    geo = dict(zone_ptr=P.zone_ptr, zpos=P.zpos, L=P.corr_len, cutoff=P.max_len, metric=0,
               loctype=1, ox=P.ox, oy=P.oy)
    return O.ensemble_analysis(geo, (P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2, P.E,
                               inflation=infl, max_correction=maxc)


def _ctx(**tune):
    from oak_amd._lib import Context
    c = Context(0)
    c.set_tuning(**tune)
    return c


@pytest.mark.parametrize("dual", [16, 32, 64])
@pytest.mark.parametrize("N", [72, 100, 128])
def test_dual_fp64_vs_oracle(N, dual):
    from oak_amd import assimilation as A
    P = _problem(N, 300 + N)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    ctx = _ctx(dual=dual)
    for infl, maxc in ((1.0, np.inf), (1.07, 0.05)):
        Eo, _, cnt = _oracle(P, infl, maxc)
        # zones on both sides of the threshold, and empty ones
        assert ((cnt >= 1) & (cnt < dual)).sum() >= 5 and (cnt >= dual).sum() >= 5
        assert (cnt == 0).sum() >= 1
        Ea = A.ensemble_analysis(_geo(P), H, P.Hshift, P.yo, P.rmse2, P.E, inflation=infl,
                                 max_correction=maxc, ctx=ctx)
        xf = P.E.mean(axis=1)
        assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < TOL64, (N, dual, infl)
    assert ctx.tuning()["dual"] == dual
    ctx.close()


@pytest.mark.parametrize("N", [100, 128])
def test_dual_mixed_vs_oracle(N):
    from oak_amd import assimilation as A
    P = _problem(N, 400 + N)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    E32 = np.asfortranarray(P.E, dtype=np.float32)
    ctx = _ctx(dual=64)
    for infl, maxc in ((1.0, np.inf), (1.07, 0.05)):
        Eo, _, _ = _oracle(P, infl, maxc)
        Ea = A.ensemble_analysis(_geo(P), H, P.Hshift, P.yo, P.rmse2, E32, inflation=infl,
                                 max_correction=maxc, ctx=ctx)
        e = ensemble_errors(Ea, Eo, P.E)
        assert e["mean_err"] < TOL_MIXED and e["spread_err"] < TOL_MIXED, (N, infl, e)
    ctx.close()


def test_dual_locanalysis_with_amplitudes():
    """locAnalysis from a CSR selection, with localise_obs = .false. (all m
    observations in every zone that has one relevant, rrsqrt.F90:373-385, so
    every non-empty zone is dual when m < 64) and with localisation: xa, Sa and
    the amplitudes as the ensemble-space analysis gives them."""
    from oak_amd.rrsqrt import loc_analysis_csr
    g = np.random.default_rng(17)
    N = 110
    sizes = [1, 17, 64, 130, 300, 5]
    zone_ptr = np.concatenate([[0], np.cumsum(sizes)])
    n, m = int(zone_ptr[-1]), 50
    Sf = np.asfortranarray(g.standard_normal((n, N)) / np.sqrt(N - 1))
    HSf = np.asfortranarray(g.standard_normal((m, N)) / np.sqrt(N - 1))
    xf = g.standard_normal(n); Hxf = g.standard_normal(m)
    yo = Hxf + 0.3 * g.standard_normal(m)
    sig2 = g.uniform(0.05, 0.5, m)
    for localise_obs in (True, False):
        counts = [40, 7, 0, 50, 33, 1] if localise_obs else [m, m, 0, m, m, m]
        sp, si, sw = [0], [], []
        for cnt in counts:
            idx = np.sort(g.choice(m, cnt, replace=False))
            si.extend(idx.tolist()); sw.extend(g.uniform(0.05, 1.0, cnt).tolist())
            sp.append(sp[-1] + cnt)
        ref_ctx, ctx = _ctx(), _ctx(dual=64)
        xa0, Sa0, am0 = loc_analysis_csr(zone_ptr, sp, si, sw, xf, Hxf, yo, Sf, HSf, sig2,
                                         amplitudes=True, localise_obs=localise_obs, ctx=ref_ctx)
        xa1, Sa1, am1 = loc_analysis_csr(zone_ptr, sp, si, sw, xf, Hxf, yo, Sf, HSf, sig2,
                                         amplitudes=True, localise_obs=localise_obs, ctx=ctx)
        xo, So, amo = O.loc_analysis_csr(zone_ptr, sp, si, sw, xf, Hxf, yo, sig2, Sf, HSf,
                                         localise_obs=localise_obs, want_ampl=True)
        assert rel_err(xa1 - xf, xo - xf) < TOL64 and rel_err(Sa1, So) < TOL64
        assert rel_err(xa1 - xf, xa0 - xf) < TOL64 and rel_err(Sa1, Sa0) < TOL64
        if not localise_obs:
            assert rel_err(am1, am0) < TOL64 and rel_err(am1, amo) < TOL64
        ref_ctx.close(); ctx.close()


def test_dual_excluded_observations_and_single_observation_zones():
    """Excluded observations (rho = 0: zero rows of S~) inside dual zones, zones
    whose observations are all excluded (G = 0, mu = 1: a = 0, T = Omega) and
    zones with one observation, against the oracle with the same exclusions."""
    from oak_amd import assimilation as A
    P = _problem(128, 700)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    g = np.random.default_rng(3)
    ex = (g.uniform(size=P.m) < 0.3).astype(np.uint8)
    geo = dict(zone_ptr=P.zone_ptr, zpos=P.zpos, L=P.corr_len, cutoff=P.max_len, metric=0,
               loctype=1, ox=P.ox, oy=P.oy)
    # the oracle takes OAK's mask D (1 = used, 0 = excluded, assimilation.F90:3086-3092)
    Eo, _, cnt = O.ensemble_analysis(geo, (P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2, P.E,
                                     dex=1.0 - ex)
    assert ((cnt >= 1) & (cnt < 64)).sum() >= 20
    ctx = _ctx(dual=64)
    Ea = A.ensemble_analysis(_geo(P), H, P.Hshift, P.yo, P.rmse2, P.E, excluded=ex, ctx=ctx)
    xf = P.E.mean(axis=1)
    assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < TOL64
    # every observation of the problem excluded: each zone keeps its forecast
    # mean, and its anomalies are the forecast's rotated by Omega
    exa = np.ones(P.m, np.uint8)
    Eo2, _, _ = O.ensemble_analysis(geo, (P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2, P.E,
                                    dex=1.0 - exa)
    Ea2 = A.ensemble_analysis(_geo(P), H, P.Hshift, P.yo, P.rmse2, P.E, excluded=exa, ctx=ctx)
    assert rel_err(Ea2 - xf[:, None], Eo2 - xf[:, None]) < TOL64
    assert np.abs(Ea2.mean(axis=1) - xf).max() < 1e-12
    ctx.close()


@pytest.mark.parametrize("split", [2, 3, 4])
@pytest.mark.parametrize("N", [100, 128])
def test_three_launch_reduction_vs_oracle(N, split):
    """OAKG_TUNE_TRIDIAG_SPLIT = 2 (steps 32..63 of the Householder reduction
    on 6 x 6 tiles), 3 (steps 0..63 on the lower triangle only, the upper
    part of A v from column sums) and 4 (steps 64..125 too, by the NP = 64
    kernel on the trailing block): the partial sums are grouped
    differently; against the oracle, with and without the observation-space
    bucket."""
    from oak_amd import assimilation as A
    P = _problem(N, 600 + N)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    Eo, _, _ = _oracle(P, 1.05, 0.1)
    xf = P.E.mean(axis=1)
    for dual in (0, 64):
        ctx = _ctx(tri_split=split, dual=dual, order=1)
        Ea = A.ensemble_analysis(_geo(P), H, P.Hshift, P.yo, P.rmse2, P.E, inflation=1.05,
                                 max_correction=0.1, ctx=ctx)
        assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < TOL64, (N, dual)
        ctx.close()


def test_dual_same_bits_under_order_and_graph():
    from oak_amd import assimilation as A
    import torch
    P = _problem(128, 500)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    ref_ctx = _ctx(dual=64)
    ref = A.ensemble_analysis(_geo(P), H, P.Hshift, P.yo, P.rmse2, P.E, ctx=ref_ctx)
    ref_ctx.close()
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(dev, dt)
    g = A.Geometry(t(P.zone_ptr, torch.int64), t(P.zpos), t(P.corr_len), t(P.max_len), t(P.ox),
                   t(P.oy))
    rp, co, va = H
    Hd = (t(rp, torch.int64), t(co, torch.int32), t(va))
    E = t(P.E.T).t()
    out = torch.empty((E.shape[1], E.shape[0]), dtype=torch.float64, device=dev).t()
    ctx = _ctx(dual=64, order=1, graph=1, tri_split=1)
    for _ in range(3):
        A.ensemble_analysis(g, Hd, t(P.Hshift), t(P.yo), t(P.rmse2), E, ctx=ctx, out=out)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), ref)
    assert ctx.tuning()["graph_replays"] == 2
    ctx.close()
