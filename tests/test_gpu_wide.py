"""Ensembles of 129..256 members (NP = 256, oak_amd/csrc/wide256.hip): the
Gram on 16 waves of 32 x 32 blocks, the Householder reduction on 8 x 8
register tiles of the lower triangle, Q in four column blocks, QL with each
row's rotations in registers, the transform in 64-row blocks and the update
in 32-column blocks.  Against the
oracle at the FP64 bound (1e-10 of the largest increment / anomaly) and the
mixed bound (1e-4, at N = 200 and 256), on column-ocean workloads whose zones have from 0 to a
few hundred observations (0 to about 500) (so p < N, p ~ N and p > N all occur), with
inflation and a binding clamp; locAnalysis on a CSR selection with zones of
1 to 300 elements; the same bits under the zone order and graph replay.
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


def _geo(P):
    from oak_amd.assimilation import Geometry
    return Geometry(P.zone_ptr, P.zpos, P.corr_len, P.max_len, P.ox, P.oy, metric=0, loctype=1)


def _problem(N, seed, n_ssh=1500):
    from oak_amd import synth
    # 320 columns, tracks of 1500 SSH values: local counts from 0 to about 500
    return synth.column_ocean(20, 16, 2, N, n_ssh, 4, seed_cfg=seed, hlen=3, vlen=2, L=1.0,
                              cutoff=3.0)


def _oracle(P, infl=1.0, maxc=np.inf):
    geo = dict(zone_ptr=P.zone_ptr, zpos=P.zpos, L=P.corr_len, cutoff=P.max_len, metric=0,
               loctype=1, ox=P.ox, oy=P.oy)
    return O.ensemble_analysis(geo, (P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2, P.E,
                               inflation=infl, max_correction=maxc)


def _ctx(**tune):
    from oak_amd._lib import Context
    c = Context(0)
    if tune:
        c.set_tuning(**tune)
    return c


@pytest.mark.parametrize("N", [129, 256])
def test_wide_fp64_vs_oracle(N):
    from oak_amd import assimilation as A
    P = _problem(N, 900 + N)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    ctx = _ctx()
    xf = P.E.mean(axis=1)
    for infl, maxc in ((1.0, np.inf), (1.07, 0.05)):
        Eo, _, cnt = _oracle(P, infl, maxc)
        # zones with no, few (p < N) and many (p > N) observations
        assert (cnt == 0).sum() >= 1 and ((cnt >= 1) & (cnt < N)).sum() >= 5
        assert (cnt > N).sum() >= 5
        Ea = A.ensemble_analysis(_geo(P), H, P.Hshift, P.yo, P.rmse2, P.E, inflation=infl,
                                 max_correction=maxc, ctx=ctx)
        assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < TOL64, (N, infl)
    ctx.close()


@pytest.mark.parametrize("N", [200, 256])
def test_wide_mixed_vs_oracle(N):
    from oak_amd import assimilation as A
    P = _problem(N, 1000 + N)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    E32 = np.asfortranarray(P.E, dtype=np.float32)
    ctx = _ctx()
    for infl, maxc in ((1.0, np.inf), (1.07, 0.05)):
        Eo, _, _ = _oracle(P, infl, maxc)
        Ea = A.ensemble_analysis(_geo(P), H, P.Hshift, P.yo, P.rmse2, E32, inflation=infl,
                                 max_correction=maxc, ctx=ctx)
        e = ensemble_errors(Ea, Eo, P.E)
        assert e["mean_err"] < TOL_MIXED and e["spread_err"] < TOL_MIXED, (N, infl, e)
    ctx.close()


@pytest.mark.parametrize("N", [131, 256])
def test_wide_locanalysis_csr_vs_oracle(N):
    """locAnalysis from a CSR selection (rrsqrt.F90:258-426): xa, Sa and, with
    localise_obs = .false., the amplitudes."""
    from oak_amd.rrsqrt import loc_analysis_csr
    g = np.random.default_rng(N)
    sizes = [1, 17, 64, 130, 300, 5, 41]
    zone_ptr = np.concatenate([[0], np.cumsum(sizes)])
    n, m = int(zone_ptr[-1]), 420
    Sf = np.asfortranarray(g.standard_normal((n, N)) / np.sqrt(N - 1))
    HSf = np.asfortranarray(g.standard_normal((m, N)) / np.sqrt(N - 1))
    xf = g.standard_normal(n)
    Hxf = g.standard_normal(m)
    yo = Hxf + 0.3 * g.standard_normal(m)
    sig2 = g.uniform(0.05, 0.5, m)
    for localise_obs in (True, False):
        counts = [40, 7, 0, 300, 410, 1, 180] if localise_obs else [m] * 7
        sp, si, sw = [0], [], []
        for cnt in counts:
            idx = np.sort(g.choice(m, cnt, replace=False))
            si.extend(idx.tolist())
            sw.extend(g.uniform(0.05, 1.0, cnt).tolist())
            sp.append(sp[-1] + cnt)
        ctx = _ctx()
        xa, Sa, am = loc_analysis_csr(zone_ptr, sp, si, sw, xf, Hxf, yo, Sf, HSf, sig2,
                                      amplitudes=True, localise_obs=localise_obs, ctx=ctx)
        xo, So, amo = O.loc_analysis_csr(zone_ptr, sp, si, sw, xf, Hxf, yo, sig2, Sf, HSf,
                                         localise_obs=localise_obs, want_ampl=True)
        assert rel_err(xa - xf, xo - xf) < TOL64 and rel_err(Sa, So) < TOL64, (N, localise_obs)
        if not localise_obs:
            assert rel_err(am, amo) < TOL64
        ctx.close()


def test_wide_same_bits_under_order_and_graph():
    from oak_amd import assimilation as A
    import torch
    P = _problem(200, 1300, n_ssh=800)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    ref_ctx = _ctx()
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
    ctx = _ctx(order=1, graph=1, ql_zones=8)
    for _ in range(3):
        A.ensemble_analysis(g, Hd, t(P.Hshift), t(P.yo), t(P.rmse2), E, ctx=ctx, out=out)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), ref)
    assert ctx.tuning()["graph_replays"] == 2
    ctx.close()


def test_wide_likelihood_unsupported():
    """The likelihood's adjoint kernels stop at N = 128: OAKG_EUNSUP above."""
    from oak_amd import synth
    from oak_amd._lib import OAKG_EUNSUP, OakError
    from oak_amd.nll import NLLProblem, evaluate, problem_from_ensemble
    P = synth.config3(nx=16, N=12, m=40, seed_cfg=3)
    prob = problem_from_ensemble(P)
    g = np.random.default_rng(5)
    wide = NLLProblem(prob.geo, prob.Hxf, prob.yo,
                      np.asfortranarray(g.standard_normal((prob.m, 140))), 140)
    with pytest.raises(OakError) as ei:
        evaluate(wide, P.rmse2, np.full(P.nzones, 5.0), 1.0)
    assert ei.value.code == OAKG_EUNSUP
