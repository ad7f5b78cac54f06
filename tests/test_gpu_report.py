"""Innovation statistics of report (assimilation.F90:3584-3628) on the GPU,
through the C ABI (oakg_report, oakg_ensemble_report[_f32]), against the CPU
oracle and, at a large size, through properties that do not depend on it.
Needs an MI355X."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

# FP64 sums in a different order than the oracle's: relative bound per value
TOL64 = 1e-12


@pytest.fixture(scope="module")
def A():
    import oak_amd.assimilation as A
    return A


def _case(seed, m, N, frac_ex=0.2):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=m)
    HS = np.asfortranarray(rng.normal(size=(m, N)))
    sig2 = rng.uniform(0.1, 3.0, size=m)
    ex = (rng.random(m) < frac_ex).astype(np.uint8)
    return d, HS, sig2, ex


def _oracle_ranges(d, HS, sig2, ex, cut):
    dex = None if ex is None else 1.0 - ex.astype(float)
    return np.array([O.report(d[a:b], HS[a:b], sig2[a:b], None if dex is None else dex[a:b])
                     for a, b in zip(cut[:-1], cut[1:])])


def _close(got, exp, tol):
    got = np.asarray(got)
    assert got.shape == exp.shape
    np.testing.assert_array_equal(np.isnan(got), np.isnan(exp))
    ok = ~np.isnan(exp)
    np.testing.assert_allclose(got[ok], exp[ok], rtol=tol, atol=0)


@pytest.mark.parametrize("N", [1, 7, 16, 33, 64, 100, 128, 200, 256])
def test_report_vs_oracle(A, N):
    m = 5000
    d, HS, sig2, ex = _case(10 + N, m, N)
    # ranges: one observation, a full piece, an empty range, a ragged tail
    cut = [0, 1, 1025, 1025, 3100, 5000]
    got = A.report(d, HS, sig2, ex, ranges=cut)
    _close(got, _oracle_ranges(d, HS, sig2, ex, cut), TOL64)
    assert np.all(np.isnan(got[2, 1:])) and got[2, 0] == 0
    # whole set, and nothing excluded (plain DiagCovar)
    _close(A.report(d, HS, sig2, ex), _oracle_ranges(d, HS, sig2, ex, [0, m]), TOL64)
    _close(A.report(d, HS, sig2), _oracle_ranges(d, HS, sig2, None, [0, m]), TOL64)


def test_report_edge_cases(A):
    # no observations: ingrid 0, NaN statistics
    out = A.report(np.zeros(0), np.zeros((0, 5), order="F"), np.zeros(0))
    assert out[0, 0] == 0 and np.all(np.isnan(out[0, 1:]))
    # every observation excluded
    d, HS, sig2, _ = _case(3, 50, 4)
    out = A.report(d, HS, sig2, np.ones(50, np.uint8))
    assert out[0, 0] == 0 and np.all(np.isnan(out[0, 1:]))
    # non-finite innovation of an excluded observation propagates into
    # sum(d R^-1 d) as in OAK (0 * Inf), not into the masked sums
    d, HS, sig2, ex = _case(4, 40, 3, 0.0)
    ex[7] = 1
    d[7] = np.inf
    got = A.report(d, HS, sig2, ex)
    exp = _oracle_ranges(d, HS, sig2, ex, [0, 40])
    _close(got, exp, TOL64)
    assert np.isfinite(got[0, 1]) and np.isnan(got[0, 3])


def test_report_rejects_bad_arguments(A):
    from oak_amd._lib import OakError
    d, HS, sig2, ex = _case(5, 100, 8)
    bad = sig2.copy()
    bad[3] = 0.0
    with pytest.raises(OakError):
        A.report(d, HS, bad, ex)
    with pytest.raises(OakError):
        A.report(d, HS, sig2, ex, ranges=[0, 60, 50])
    with pytest.raises(OakError):
        A.report(d, HS, sig2, ex, ranges=[0, 101])
    with pytest.raises(OakError):
        A.report(d, np.zeros((100, 257), order="F"), sig2, ex)  # N above OAKG_MAX_N


def test_report_device_pointers_and_leading_dimension(A):
    import torch
    m, N = 3000, 100
    d, HS, sig2, ex = _case(6, m, N)
    host = A.report(d, HS, sig2, ex, ranges=[0, 700, 3000])
    dev = torch.device("cuda:0")
    big = torch.zeros((N, m + 9), dtype=torch.float64, device=dev)
    big[:, :m] = torch.from_numpy(np.ascontiguousarray(HS.T)).to(dev)
    HSd = big.t()[:m]  # column-major view with ld = m + 9
    got = A.report(torch.from_numpy(d).to(dev), HSd, torch.from_numpy(sig2).to(dev),
                   torch.from_numpy(ex).to(dev), ranges=[0, 700, 3000])
    torch.cuda.synchronize()
    # same kernels, same piece layout: the same bits
    np.testing.assert_array_equal(got.cpu().numpy(), host)
    again = A.report(d, HS, sig2, ex, ranges=[0, 700, 3000])
    np.testing.assert_array_equal(again, host)


def test_report_large_checksum_of_checksums(A):
    """m = 400 000, N = 128 on device tensors: the statistics of eight ranges
    add up to those of the whole (ingrid * rms^2, ingrid * bias,
    ingrid * rms_invsqrtR^2), and the whole agrees with a torch FP64
    evaluation of the definitions."""
    import torch
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(7)
    m, N = 400_000, 128
    d = torch.randn(m, dtype=torch.float64, device=dev, generator=g)
    HS = torch.randn((N, m), dtype=torch.float64, device=dev, generator=g).t()
    sig2 = torch.rand(m, dtype=torch.float64, device=dev, generator=g) * 2 + 0.1
    ex = (torch.rand(m, device=dev, generator=g) < 0.1).to(torch.uint8)
    cut = np.linspace(0, m, 9).astype(np.int64)
    cut[3] += 123
    parts = A.report(d, HS, sig2, ex, ranges=cut).cpu().numpy()
    tot = A.report(d, HS, sig2, ex).cpu().numpy()[0]
    gi = parts[:, 0]
    assert gi.sum() == tot[0]
    np.testing.assert_allclose((gi * parts[:, 1] ** 2).sum(), tot[0] * tot[1] ** 2, rtol=1e-11)
    np.testing.assert_allclose((gi * parts[:, 2]).sum(), tot[0] * tot[2], rtol=1e-9)
    np.testing.assert_allclose((gi * parts[:, 3] ** 2).sum(), tot[0] * tot[3] ** 2, rtol=1e-11)
    valid = ex == 0
    r = torch.where(valid, d / sig2, torch.zeros_like(d))
    ig = valid.sum().item()
    exp = [ig, (d[valid] ** 2).sum().item() / ig, d[valid].sum().item() / ig,
           (d * r).sum().item() / ig, ((HS.t() @ r) ** 2).sum().item() / ig]
    np.testing.assert_allclose(tot[[1, 3, 4]] ** 2, np.array(exp)[[1, 3, 4]], rtol=1e-10)
    np.testing.assert_allclose(tot[2], exp[2], rtol=1e-8)


def _ens_problem(N):
    from oak_amd import synth
    P = synth.column_ocean(10, 8, 4, N, 300, 6, seed_cfg=700 + N, hlen=3, vlen=2, L=2.0,
                           cutoff=6.0)
    rng = np.random.default_rng(N)
    ex = (rng.random(P.m) < 0.15).astype(np.uint8)
    return P, ex


@pytest.mark.parametrize("N", [16, 64, 100, 128])
def test_ensemble_report_vs_oracle(A, N):
    P, ex = _ens_problem(N)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    cut = [0, 300, P.m]  # SSH observations, then the profiles
    got = A.ensemble_report(H, P.Hshift, P.yo, P.rmse2, P.E, ex, ranges=cut)
    exp = O.ensemble_report((P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2, P.E,
                            1.0 - ex.astype(float), ranges=cut)
    _close(got, exp, 1e-11)


@pytest.mark.parametrize("N", [16, 128])
def test_ensemble_report_f32(A, N):
    P, ex = _ens_problem(N)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    got = A.ensemble_report(H, P.Hshift, P.yo, P.rmse2, P.E.astype(np.float32), ex)
    exp = O.ensemble_report((P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2,
                            P.E.astype(np.float32).astype(np.float64), 1.0 - ex.astype(float))
    # FP32 observation anomalies (rounded once), FP64 sums
    _close(got, exp, 1e-5)


def test_forecast_and_analysis_reports(A):
    """Assim's two reports (assimilation.F90:3389-3391) around one analysis
    on config 0: both agree with the oracle, and the analysis fits the
    observations better than the forecast."""
    from oak_amd import synth
    P = synth.config0()
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    geo = A.Geometry(P.zone_ptr, P.zpos, P.corr_len, P.max_len, P.ox, P.oy)
    Ea = A.ensemble_analysis(geo, H, P.Hshift, P.yo, P.rmse2, P.E)
    rf = A.ensemble_report(H, P.Hshift, P.yo, P.rmse2, P.E)
    ra = A.ensemble_report(H, P.Hshift, P.yo, P.rmse2, Ea)
    coo = (P.Hi, P.Hj, P.Hs)
    _close(rf, O.ensemble_report(coo, P.Hshift, P.yo, P.rmse2, P.E), 1e-11)
    _close(ra, O.ensemble_report(coo, P.Hshift, P.yo, P.rmse2, Ea), 1e-11)
    assert ra[0, 1] < rf[0, 1] and ra[0, 3] < rf[0, 3]


def test_ensemble_report_device_tensors_and_ranges(A):
    """FP64 and FP32 device ensembles with three observation ranges: the
    same statistics as the host call (FP64: the same bits) and the oracle."""
    import torch
    P, ex = _ens_problem(100)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    cut = [0, 150, 300, P.m]
    host = A.ensemble_report(H, P.Hshift, P.yo, P.rmse2, P.E, ex, ranges=cut)
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(dev, dt)
    Hd = (t(H[0], torch.int64), t(H[1], torch.int32), t(H[2]))
    Ed = t(P.E.T).t()
    got = A.ensemble_report(Hd, t(P.Hshift), t(P.yo), t(P.rmse2), Ed, t(ex, torch.uint8),
                            ranges=cut)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(got.cpu().numpy(), host)
    exp = O.ensemble_report((P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2, P.E,
                            1.0 - ex.astype(float), ranges=cut)
    _close(host, exp, 1e-11)
    E32 = t(P.E.astype(np.float32).T, torch.float32).t()
    got32 = A.ensemble_report(Hd, t(P.Hshift), t(P.yo), t(P.rmse2), E32, t(ex, torch.uint8),
                              ranges=cut).cpu().numpy()
    exp32 = O.ensemble_report((P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2,
                              P.E.astype(np.float32).astype(np.float64), 1.0 - ex.astype(float),
                              ranges=cut)
    _close(got32, exp32, 1e-5)
