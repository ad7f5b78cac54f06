"""Mixed-precision path (FP32 state, FP32 MFMA contractions, FP64 eigensolve)
against the FP64 oracle, through oakg_ensemble_analysis_f32.  Needs an MI355X.

Bound (stated for every output): analysis mean error relative to the largest
mean increment, and analysis spread error relative to the largest analysed
anomaly, both <= 1e-4 (oak_amd.errors.ensemble_errors)."""
import numpy as np
import pytest

from oak_amd.errors import ensemble_errors
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL_MIXED = 1e-4


@pytest.fixture(scope="module", autouse=True)
def _oracle_lapack():
    O.load_lapack()
    yield
    O.set_backend(0)


def _geo(P):
    from oak_amd.assimilation import Geometry
    return Geometry(P.zone_ptr, P.zpos, P.corr_len, P.max_len, P.ox, P.oy, metric=0, loctype=1)


def _oracle(P, infl=1.0, maxc=np.inf):
    geo = dict(zone_ptr=P.zone_ptr, zpos=P.zpos, L=P.corr_len, cutoff=P.max_len, metric=0,
               loctype=1, ox=P.ox, oy=P.oy)
    Eo, _, _ = O.ensemble_analysis(geo, (P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2, P.E,
                                   inflation=infl, max_correction=maxc)
    return Eo


def test_mixed_cfg0_vs_oracle():
    from oak_amd import synth
    from oak_amd import assimilation as A
    P = synth.config0()
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    E32 = np.asfortranarray(P.E, dtype=np.float32)
    Ea = A.ensemble_analysis(_geo(P), H, P.Hshift, P.yo, P.rmse2, E32)
    assert Ea.dtype == np.float32
    e = ensemble_errors(Ea, _oracle(P), P.E)
    assert e["mean_err"] < TOL_MIXED and e["spread_err"] < TOL_MIXED, e


@pytest.mark.parametrize("N", [5, 24, 64, 100, 128])
def test_mixed_sizes_vs_oracle(N):
    from oak_amd import synth
    from oak_amd import assimilation as A
    P = synth.column_ocean(12, 9, 3, N, 60, 4, seed_cfg=200 + N, hlen=3, vlen=2, L=2.0,
                           cutoff=5.0)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    E32 = np.asfortranarray(P.E, dtype=np.float32)
    for infl, maxc in ((1.0, np.inf), (1.07, 0.05)):
        Ea = A.ensemble_analysis(_geo(P), H, P.Hshift, P.yo, P.rmse2, E32, inflation=infl,
                                 max_correction=maxc)
        e = ensemble_errors(Ea, _oracle(P, infl, maxc), P.E)
        assert e["mean_err"] < TOL_MIXED and e["spread_err"] < TOL_MIXED, (N, infl, e)


def test_mixed_device_matches_host():
    import torch
    from oak_amd import synth
    from oak_amd import assimilation as A
    from oak_amd.assimilation import Geometry
    P = synth.config0()
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    E32 = np.asfortranarray(P.E, dtype=np.float32)
    Eh = A.ensemble_analysis(_geo(P), H, P.Hshift, P.yo, P.rmse2, E32)
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(dev, dt)
    gd = Geometry(t(P.zone_ptr, torch.int64), t(P.zpos), t(P.corr_len), t(P.max_len), t(P.ox),
                  t(P.oy), metric=0, loctype=1)
    Ed = t(P.E.T, torch.float32).t()
    Hd = (t(H[0], torch.int64), t(H[1], torch.int32), t(H[2]))
    Ea = A.ensemble_analysis(gd, Hd, t(P.Hshift, torch.float32), t(P.yo), t(P.rmse2), Ed)
    torch.cuda.synchronize()
    assert Ea.dtype == torch.float32
    np.testing.assert_array_equal(Ea.cpu().numpy(), Eh)


def test_mixed_empty_zones_keep_forecast():
    from oak_amd import synth
    from oak_amd import assimilation as A
    P = synth.config0()
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    E32 = np.asfortranarray(P.E, dtype=np.float32)
    cut = np.full_like(P.max_len, 1e-9)  # no observation within reach of any zone
    from oak_amd.assimilation import Geometry
    g = Geometry(P.zone_ptr, P.zpos, P.corr_len, cut, P.ox, P.oy, metric=0, loctype=1)
    Ea = A.ensemble_analysis(g, H, P.Hshift, P.yo, P.rmse2, E32)
    np.testing.assert_allclose(Ea, E32, rtol=0, atol=2e-6 * np.abs(E32).max())


@pytest.mark.parametrize("N", [100, 128])
def test_mixed_deep_columns_vs_oracle(N):
    """161-element zones: the FP32 update runs two strip rounds per wave."""
    from oak_amd import synth
    from oak_amd import assimilation as A
    P = synth.column_ocean(4, 3, 40, N, 40, 3, seed_cfg=320 + N, hlen=3, vlen=4, L=2.0,
                           cutoff=5.0)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    E32 = np.asfortranarray(P.E, dtype=np.float32)
    for infl, maxc in ((1.0, np.inf), (1.05, 0.1)):
        Ea = A.ensemble_analysis(_geo(P), H, P.Hshift, P.yo, P.rmse2, E32, inflation=infl,
                                 max_correction=maxc)
        e = ensemble_errors(Ea, _oracle(P, infl, maxc), P.E)
        assert e["mean_err"] < TOL_MIXED and e["spread_err"] < TOL_MIXED, (N, infl, e)


def test_mixed_empty_zones_keep_forecast_n128():
    """N = 128, no observation within reach: T = I in the fused transform."""
    from oak_amd import synth
    from oak_amd import assimilation as A
    from oak_amd.assimilation import Geometry
    P = synth.column_ocean(6, 5, 40, 128, 40, 3, seed_cfg=333, hlen=3, vlen=4, L=2.0, cutoff=5.0)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    E32 = np.asfortranarray(P.E, dtype=np.float32)
    cut = np.full_like(P.max_len, 1e-9)
    g = Geometry(P.zone_ptr, P.zpos, P.corr_len, cut, P.ox, P.oy, metric=0, loctype=1)
    Ea = A.ensemble_analysis(g, H, P.Hshift, P.yo, P.rmse2, E32)
    np.testing.assert_allclose(Ea, E32, rtol=0, atol=2e-6 * np.abs(E32).max())
