"""Parity of the HIP path (through the C ABI) with the CPU oracle and with
OAK's known answers.  Needs an MI355X."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.kfcheck import global_kf, local_kf_rows, rel_err, rrsqrt_inputs

pytestmark = pytest.mark.gpu

# FP64 parity bound of the plain path, relative to the largest magnitude of
# the compared quantity (increments xa - xf, anomalies Sa, ensembles Ea).
TOL64 = 1e-10


@pytest.fixture(scope="module")
def oak():
    import oak_amd.assimilation as A
    import oak_amd.rrsqrt as R
    return R, A


def _gc_select(n, m):
    xmod = np.arange(n) / (n - 1.0)
    xobs = np.arange(m) / (m - 1.0)

    def sel(i):
        c = np.array([O.locfun(abs(xmod[i] - xo) / 0.21) for xo in xobs])
        return c, c != 0
    return sel


def test_global_analysis_textbook_kf(oak):
    """test/test_rrsqrt.F90:27-77 on the GPU."""
    R, _ = oak
    d = rrsqrt_inputs()
    Hxf = d["H"] @ d["xf"]; HSf = d["H"] @ d["Sf"]
    xa, Sa = R.analysis(d["xf"], Hxf, d["y"], d["Sf"], HSf, R.DiagCovar(d["sig2"]))
    xa_chk, Pa_chk = global_kf(d["xf"], d["Sf"], d["H"], d["y"], np.diag(d["sig2"]))
    assert rel_err(xa, xa_chk) < 1e-10
    assert rel_err(Sa @ Sa.T, Pa_chk) < 1e-10
    assert np.max(np.abs(Sa.sum(axis=1))) < 1e-12


def test_loc_analysis_select_all(oak):
    """test/test_rrsqrt.F90:142-159: one zone, then one zone per element."""
    R, _ = oak
    d = rrsqrt_inputs()
    Hxf = d["H"] @ d["xf"]; HSf = d["H"] @ d["Sf"]
    xa_chk, Pa_chk = global_kf(d["xf"], d["Sf"], d["H"], d["y"], np.diag(d["sig2"]))
    allsel = lambda i: (np.ones(d["m"]), np.ones(d["m"], bool))
    xa, Sa = R.loc_analysis([d["n"]], allsel, d["xf"], Hxf, d["y"], d["Sf"], HSf,
                            R.DiagCovar(d["sig2"]))
    assert rel_err(xa, xa_chk) < 1e-10 and rel_err(Sa @ Sa.T, Pa_chk) < 1e-10
    xa, Sa = R.loc_analysis([1] * d["n"], allsel, d["xf"], Hxf, d["y"], d["Sf"], HSf,
                            R.DiagCovar(d["sig2"]))
    assert rel_err(xa, xa_chk) < 1e-10
    assert rel_err(np.sum(Sa * Sa, axis=1), np.diag(Pa_chk)) < 1e-10


@pytest.mark.parametrize("localise_obs", [False, True])
def test_loc_analysis_gaspari_cohn(oak, localise_obs):
    """test/test_rrsqrt.F90:163-236 (w^2 taper) on the GPU, and against the oracle."""
    R, _ = oak
    d = rrsqrt_inputs()
    n, m = d["n"], d["m"]
    Hxf = d["H"] @ d["xf"]; HSf = d["H"] @ d["Sf"]
    sel = _gc_select(n, m)
    xa, Sa, ampl = R.loc_analysis([1] * n, sel, d["xf"], Hxf, d["y"], d["Sf"], HSf,
                                  R.DiagCovar(d["sig2"]), amplitudes=True,
                                  localise_obs=localise_obs)
    for i in range(n):
        w, rel = sel(i)
        idx = np.nonzero(rel if localise_obs else np.ones(m, bool))[0]
        xa_i, Pa_i = local_kf_rows(d["xf"], d["Sf"], d["H"], d["y"], d["sig2"], [i], idx, w[idx])
        assert abs(xa[i] - xa_i[0]) < 1e-10 * max(1, abs(xa_i[0]))
        assert abs(Sa[i] @ Sa[i] - Pa_i[0, 0]) < 1e-10 * max(1, Pa_i[0, 0])
    # oracle, same selection
    from oak_amd.rrsqrt import selection_csr
    sp, si, sw = selection_csr(np.arange(n + 1), sel, m, localise_obs)
    xo, So, ao = O.loc_analysis_csr(np.arange(n + 1), sp, si, sw, d["xf"], Hxf, d["y"], d["sig2"],
                                    d["Sf"], HSf, localise_obs=localise_obs, want_ampl=True)
    assert rel_err(xa - d["xf"], xo - d["xf"]) < TOL64
    assert rel_err(Sa, So) < TOL64
    if localise_obs:
        assert not ampl.any()
    else:
        assert rel_err(ampl, ao) < TOL64


def test_empty_zones_exact_identity(oak):
    R, _ = oak
    d = rrsqrt_inputs()
    Hxf = d["H"] @ d["xf"]; HSf = d["H"] @ d["Sf"]
    none = lambda i: (np.zeros(d["m"]), np.zeros(d["m"], bool))
    xa, Sa = R.loc_analysis([3, 7], none, d["xf"], Hxf, d["y"], d["Sf"], HSf, R.DiagCovar(d["sig2"]))
    np.testing.assert_array_equal(xa, d["xf"])
    np.testing.assert_array_equal(Sa, d["Sf"])


def test_excluded_observations(oak):
    R, _ = oak
    from oak_amd.rrsqrt import loc_analysis_csr
    d = rrsqrt_inputs()
    Hxf = d["H"] @ d["xf"]; HSf = d["H"] @ d["Sf"]
    ex = np.array([0, 1, 0, 0, 1], np.uint8)
    keep = ex == 0
    m = d["m"]
    xa1, Sa1 = loc_analysis_csr([0, d["n"]], [0, m], np.arange(m), np.ones(m), d["xf"], Hxf, d["y"],
                                d["Sf"], HSf, d["sig2"], excluded=ex)
    xa2, Sa2 = R.analysis(d["xf"], Hxf[keep], d["y"][keep], d["Sf"], HSf[keep],
                          R.DiagCovar(d["sig2"][keep]))
    assert rel_err(xa1 - d["xf"], xa2 - d["xf"]) < TOL64 and rel_err(Sa1, Sa2) < TOL64


def _geometry(P):
    from oak_amd.assimilation import Geometry
    return Geometry(P.zone_ptr, P.zpos, P.corr_len, P.max_len, P.ox, P.oy, metric=P.metric,
                    loctype=P.loctype)


def _anomalies(P):
    N = P.N
    xf = P.E.mean(axis=1)
    Sf = (P.E - xf[:, None]) / np.sqrt(N - 1)
    HE = np.zeros((P.m, N))
    for k in range(N):
        HE[:, k] = O.obsoper(P.Hi, P.Hj, P.Hs, P.m, P.E[:, k])
    Hxf = HE.mean(axis=1)
    HSf = (HE - Hxf[:, None]) / np.sqrt(N - 1)
    return xf, Sf, Hxf, HSf


@pytest.fixture(scope="module")
def cfg0():
    from oak_amd import synth
    return synth.config0()


def test_selection_cfg0_bit_exact(oak, cfg0):
    """GPU selectObservations = OAK's (same obs sets, weights to rounding)."""
    _, A = oak
    P = cfg0
    g = _geometry(P)
    sp, si, sw = A.select_observations(g, P.yo, P.rmse2)
    for z in range(0, P.nzones, 7):
        w, rel, cnt = O.select(0, 1, P.zpos[z], P.corr_len[z], P.max_len[z], P.ox, P.oy)
        idx = np.nonzero(rel)[0]
        assert sp[z + 1] - sp[z] == cnt
        np.testing.assert_array_equal(si[sp[z]:sp[z + 1]], idx)
        np.testing.assert_allclose(sw[sp[z]:sp[z + 1]], w[idx], rtol=4e-16, atol=0)


def test_loc_analysis_cfg0_vs_oracle(oak, cfg0):
    _, A = oak
    P = cfg0
    xf, Sf, Hxf, HSf = _anomalies(P)
    g = _geometry(P)
    xa, Sa = A.loc_analysis(g, xf, Hxf, P.yo, Sf, HSf, P.rmse2)
    geo = dict(zone_ptr=P.zone_ptr, zpos=P.zpos, L=P.corr_len, cutoff=P.max_len, metric=0,
               loctype=1, ox=P.ox, oy=P.oy)
    xo, So, cnt = O.loc_analysis_geo(geo, xf, Hxf, P.yo, P.rmse2, Sf, HSf)
    assert rel_err(xa - xf, xo - xf) < TOL64
    assert rel_err(Sa, So) < TOL64


def test_ensemble_analysis_cfg0_vs_oracle(oak, cfg0):
    _, A = oak
    P = cfg0
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    Ea = A.ensemble_analysis(_geometry(P), H, P.Hshift, P.yo, P.rmse2, P.E)
    geo = dict(zone_ptr=P.zone_ptr, zpos=P.zpos, L=P.corr_len, cutoff=P.max_len, metric=0,
               loctype=1, ox=P.ox, oy=P.oy)
    Eo, xo, cnt = O.ensemble_analysis(geo, (P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2, P.E)
    xf = P.E.mean(axis=1)
    assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < TOL64


@pytest.fixture(scope="module", autouse=True)
def _oracle_lapack():
    O.load_lapack()
    yield
    O.set_backend(0)


@pytest.fixture(scope="module")
def jacobi_ctx():
    from oak_amd._lib import OAKG_EIG_JACOBI, Context
    c = Context(0)
    c.set_eigensolver(OAKG_EIG_JACOBI)
    yield c
    c.close()


@pytest.mark.parametrize("eig", ["ql", "jacobi"])
@pytest.mark.parametrize("N", [5, 12, 24, 40, 64, 100, 128])
def test_ensemble_sizes_vs_oracle(oak, N, eig, request):
    from oak_amd import synth
    _, A = oak
    ctx = request.getfixturevalue("jacobi_ctx") if eig == "jacobi" else None
    P = synth.column_ocean(12, 9, 3, N, 60, 4, seed_cfg=100 + N, hlen=3, vlen=2, L=2.0,
                           cutoff=5.0)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    for infl, maxc in ((1.0, np.inf), (1.07, 0.05)):
        Ea = A.ensemble_analysis(_geometry(P), H, P.Hshift, P.yo, P.rmse2, P.E, inflation=infl,
                                 max_correction=maxc, ctx=ctx)
        geo = dict(zone_ptr=P.zone_ptr, zpos=P.zpos, L=P.corr_len, cutoff=P.max_len, metric=0,
                   loctype=1, ox=P.ox, oy=P.oy)
        Eo, _, _ = O.ensemble_analysis(geo, (P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2, P.E,
                                       inflation=infl, max_correction=maxc)
        xf = P.E.mean(axis=1)
        assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < TOL64, (N, infl, eig)


@pytest.mark.parametrize("N", [100, 128])
def test_large_zones_csr_vs_oracle(oak, N):
    """Zones of 1, 17, 64 (no observations), 130 and 300 elements: the update
    runs over several 16-row strips per wave and reloads strips past 128 rows."""
    from oak_amd.rrsqrt import loc_analysis_csr
    g = np.random.default_rng(7 + N)
    sizes = [1, 17, 64, 130, 300]
    zone_ptr = np.concatenate([[0], np.cumsum(sizes)])
    n, m = int(zone_ptr[-1]), 400
    Sf = np.asfortranarray(g.standard_normal((n, N)) / np.sqrt(N - 1))
    HSf = np.asfortranarray(g.standard_normal((m, N)) / np.sqrt(N - 1))
    xf = g.standard_normal(n); Hxf = g.standard_normal(m)
    yo = Hxf + 0.3 * g.standard_normal(m)
    sig2 = g.uniform(0.05, 0.5, m)
    for localise_obs in (True, False):
        # localise_obs = False needs all or none of the observations per zone
        counts = [40, 150, 0, 300, 90] if localise_obs else [m, m, 0, m, m]
        sp, si, sw = [0], [], []
        for cnt in counts:
            idx = np.sort(g.choice(m, cnt, replace=False))
            si.extend(idx.tolist()); sw.extend(g.uniform(0.05, 1.0, cnt).tolist())
            sp.append(sp[-1] + cnt)
        xa, Sa, ampl = loc_analysis_csr(zone_ptr, sp, si, sw, xf, Hxf, yo, Sf, HSf, sig2,
                                        amplitudes=True, localise_obs=localise_obs)
        xo, So, ao = O.loc_analysis_csr(zone_ptr, sp, si, sw, xf, Hxf, yo, sig2, Sf, HSf,
                                        localise_obs=localise_obs, want_ampl=True)
        assert rel_err(xa - xf, xo - xf) < TOL64, (N, localise_obs)
        assert rel_err(Sa, So) < TOL64, (N, localise_obs)
        np.testing.assert_array_equal(Sa[17 + 1:17 + 1 + 64], Sf[17 + 1:17 + 1 + 64])
        np.testing.assert_array_equal(xa[17 + 1:17 + 1 + 64], xf[17 + 1:17 + 1 + 64])


@pytest.mark.parametrize("N", [100, 128])
def test_ensemble_deep_columns_vs_oracle(oak, N):
    """Ensemble mode with 161-element zones (40 levels): two strip rounds per wave."""
    from oak_amd import synth
    _, A = oak
    P = synth.column_ocean(4, 3, 40, N, 40, 3, seed_cfg=300 + N, hlen=3, vlen=4, L=2.0,
                           cutoff=5.0)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    for infl, maxc in ((1.0, np.inf), (1.05, 0.1)):
        Ea = A.ensemble_analysis(_geometry(P), H, P.Hshift, P.yo, P.rmse2, P.E, inflation=infl,
                                 max_correction=maxc)
        geo = dict(zone_ptr=P.zone_ptr, zpos=P.zpos, L=P.corr_len, cutoff=P.max_len, metric=0,
                   loctype=1, ox=P.ox, oy=P.oy)
        Eo, _, _ = O.ensemble_analysis(geo, (P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2, P.E,
                                       inflation=infl, max_correction=maxc)
        xf = P.E.mean(axis=1)
        assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < TOL64, (N, infl)


@pytest.mark.parametrize("N", [16, 128])
def test_chunk_pipeline_matches_in_order(oak, N):
    """Without phase timing the zone chunks run on several streams; with it
    they run in order on one stream; with a small workspace each stream runs
    many small chunks in turn.  Every zone is computed alone, so the three
    schedules (and their different chunk sizes) give the same bits."""
    from oak_amd import synth
    from oak_amd._lib import Context
    _, A = oak
    P = synth.column_ocean(64, 64, 1, N, 3000, 200, seed_cfg=500 + N, hlen=4, vlen=1, L=3.0,
                           cutoff=9.0)
    assert P.nzones >= 4096  # the pipeline runs from 4096 zones on
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    out = []
    for timing, ws in ((False, None), (True, None), (False, 16 << 20)):
        # a 16 MiB workspace: chunks of 64-400 zones, each stream running many
        if ws:
            os.environ["OAKG_WORKSPACE_BYTES"] = str(ws)
        try:
            c = Context(0)
        finally:
            os.environ.pop("OAKG_WORKSPACE_BYTES", None)
        c.set_timing(timing)
        out.append(A.ensemble_analysis(_geometry(P), H, P.Hshift, P.yo, P.rmse2, P.E, ctx=c))
        c.close()
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])
    assert not np.array_equal(out[0], P.E)


@pytest.mark.parametrize("metric", [1, 2])
def test_spherical_metrics_vs_oracle(oak, metric):
    _, A = oak
    from oak_amd.assimilation import Geometry
    g0 = np.random.default_rng(metric)
    nx, ny, N, m = 20, 12, 10, 80
    n = nx * ny
    lon, lat = np.meshgrid(np.linspace(-30, 30, nx), np.linspace(20, 60, ny), indexing="ij")
    zpos = np.zeros((n, 4)); zpos[:, 0] = lon.ravel(order="F"); zpos[:, 1] = lat.ravel(order="F")
    ox = g0.uniform(-30, 30, m); oy = g0.uniform(20, 60, m)
    E = np.asfortranarray(g0.standard_normal((n, N)))
    Hi = np.arange(m, dtype=np.int32)
    Hj = g0.integers(0, n, m).astype(np.int32)
    Hs = np.ones(m)
    yo = g0.standard_normal(m); r2 = np.full(m, 0.5)
    L = np.full(n, 800e3); cut = np.full(n, 1500e3)
    geo = Geometry(np.arange(n + 1), zpos, L, cut, ox, oy, metric=metric, loctype=1)
    H = A.coo_to_csr(m, Hi, Hj, Hs)
    Ea = A.ensemble_analysis(geo, H, np.zeros(m), yo, r2, E)
    Eo, _, _ = O.ensemble_analysis(dict(zone_ptr=np.arange(n + 1), zpos=zpos, L=L, cutoff=cut,
                                        metric=metric, loctype=1, ox=ox, oy=oy),
                                   (Hi, Hj, Hs), np.zeros(m), yo, r2, E)
    xf = E.mean(axis=1)
    assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < 1e-9


def test_device_pointers_match_host(oak, cfg0):
    import torch
    _, A = oak
    P = cfg0
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    g = _geometry(P)
    Ea_h = A.ensemble_analysis(g, H, P.Hshift, P.yo, P.rmse2, P.E)
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(dev, dt)
    from oak_amd.assimilation import Geometry
    gd = Geometry(t(P.zone_ptr, torch.int64), t(P.zpos), t(P.corr_len), t(P.max_len), t(P.ox),
                  t(P.oy), metric=0, loctype=1)
    Ed = t(P.E.T).t()
    Hd = (t(H[0], torch.int64), t(H[1], torch.int32), t(H[2]))
    Ea_d = A.ensemble_analysis(gd, Hd, t(P.Hshift), t(P.yo), t(P.rmse2), Ed)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(Ea_d.cpu().numpy(), Ea_h)


def test_host_leading_dimensions(oak, cfg0):
    """Host ensembles with leading dimension n + 3 through the C ABI (staged
    by a strided copy, written back column by column) give the bits of the
    dense call."""
    import ctypes as C
    from oak_amd import _lib
    from oak_amd.rrsqrt import default_context
    _, A = oak
    P = cfg0
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    g = _geometry(P)
    Ea = A.ensemble_analysis(g, H, P.Hshift, P.yo, P.rmse2, P.E)
    n, N = P.E.shape
    ld = n + 3
    Ef = np.full((ld, N), np.nan, order="F")
    Ef[:n] = P.E
    Eo = np.full((ld, N), -7.0, order="F")
    gh = g.host_arrays()
    yo = np.ascontiguousarray(P.yo, np.float64)
    r2 = np.ascontiguousarray(P.rmse2, np.float64)
    z, o = gh.structs(yo, r2)
    rp, co, va = (np.ascontiguousarray(H[0], np.int64), np.ascontiguousarray(H[1], np.int32),
                  np.ascontiguousarray(H[2], np.float64))
    h = _lib.Csr(P.m, n, len(va), rp.ctypes.data, co.ctypes.data, va.ctypes.data)
    hs = np.ascontiguousarray(P.Hshift, np.float64)
    ctx = default_context()
    rc = _lib.load().oakg_ensemble_analysis(ctx.handle, C.byref(z), C.byref(o), C.byref(h),
                                            hs.ctypes.data, N, Ef.ctypes.data, ld,
                                            Eo.ctypes.data, ld, 1.0, float("inf"), None, 0)
    ctx.check(rc)
    np.testing.assert_array_equal(Eo[:n], Ea)
    assert np.all(Eo[n:] == -7.0)  # rows beyond n untouched


def test_deterministic(oak, cfg0):
    _, A = oak
    P = cfg0
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    a = A.ensemble_analysis(_geometry(P), H, P.Hshift, P.yo, P.rmse2, P.E)
    b = A.ensemble_analysis(_geometry(P), H, P.Hshift, P.yo, P.rmse2, P.E)
    np.testing.assert_array_equal(a, b)


def test_error_paths(oak):
    R, _ = oak
    from oak_amd import NaNError, OakError
    d = rrsqrt_inputs()
    Hxf = d["H"] @ d["xf"]; HSf = d["H"] @ d["Sf"]
    bad = d["sig2"].copy(); bad[2] = 0.0
    with pytest.raises(OakError) as e:
        R.analysis(d["xf"], Hxf, d["y"], d["Sf"], HSf, R.DiagCovar(bad))
    assert e.value.code == 1
    y = d["y"].copy(); y[0] = np.nan
    with pytest.raises(NaNError):
        R.analysis(d["xf"], Hxf, y, d["Sf"], HSf, R.DiagCovar(d["sig2"]))


def test_golden_fixture(oak):
    """The committed small ensemble case (tests/golden/ens_small.npz):
    the analysed ensemble and both reports against the stored oracle
    outputs."""
    _, A = oak
    d = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "ens_small.npz")))
    H = A.coo_to_csr(len(d["yo"]), d["Hi"], d["Hj"], d["Hs"])
    geo = A.Geometry(d["zone_ptr"], d["zpos"], d["L"], d["cutoff"], d["ox"], d["oy"])
    Ea = A.ensemble_analysis(geo, H, d["Hshift"], d["yo"], d["rmse2"], d["E"],
                             excluded=d["excluded"], inflation=float(d["inflation"]),
                             max_correction=float(d["max_correction"]))
    xf = d["E"].mean(axis=1)[:, None]
    assert rel_err(Ea - xf, d["Ea"] - xf) < TOL64
    rf = A.ensemble_report(H, d["Hshift"], d["yo"], d["rmse2"], d["E"], d["excluded"])
    ra = A.ensemble_report(H, d["Hshift"], d["yo"], d["rmse2"], Ea, d["excluded"])
    np.testing.assert_allclose(rf[0], d["report_forecast"], rtol=1e-11)
    np.testing.assert_allclose(ra[0], d["report_analysis"], rtol=1e-9)
