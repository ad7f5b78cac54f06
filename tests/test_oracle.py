"""Pins the CPU oracle (oracle/oak_oracle.c) against OAK's own known answers
(test/test_rrsqrt.F90, test/test_covariance.F90) and against closed forms that
do not share the oracle's formulas."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.kfcheck import global_kf, local_kf_rows, rel_err, rrsqrt_inputs

BACKENDS = [0]
if O.load_lapack():
    BACKENDS.append(1)
O.set_backend(0)


@pytest.fixture(params=BACKENDS, ids=lambda b: ["jacobi", "lapack"][b])
def backend(request):
    O.set_backend(request.param)
    yield request.param
    O.set_backend(0)


# --- covariance.F90:645-667, known answers test/test_covariance.F90:612-617 ---
def test_locfun_known_answers():
    assert abs(O.locfun(0.0) - 1.0) < 1e-7
    assert abs(O.locfun(0.4) - 0.783573333333333) < 1e-6
    assert abs(O.locfun(1.5) - 0.0164930555555556) < 1e-6
    assert abs(O.locfun(2.5) - 0.0) < 1e-7


# --- assimilation.F90:3635-3672 ---
def test_distance_cartesian_closed_form():
    assert O.distance(0, (1.0, 2.0), (4.0, 6.0)) == 5.0
    assert O.distance(0, (0.0, 0.0), (0.0, 0.0)) == 0.0


def test_distance_spherical_closed_forms():
    R = 6378137.0
    # quarter of a meridian and a quarter of the equator
    assert abs(O.distance(1, (0.0, 0.0), (0.0, 90.0)) - R * math.pi / 2) < 1e-6
    assert abs(O.distance(1, (10.0, 0.0), (100.0, 0.0)) - R * math.pi / 2) < 1e-6
    # antipodes: acos argument clamped to -1
    assert abs(O.distance(1, (0.0, 0.0), (180.0, 0.0)) - R * math.pi) < 1e-6
    # generic points against the haversine formula (a different closed form)
    rng = np.random.default_rng(1)
    for _ in range(50):
        lon0, lon1 = rng.uniform(-180, 180, 2)
        lat0, lat1 = rng.uniform(-80, 80, 2)
        p0, p1 = np.radians([lon0, lat0]), np.radians([lon1, lat1])
        h = np.sin((p1[1] - p0[1]) / 2) ** 2 + np.cos(p0[1]) * np.cos(p1[1]) * np.sin((p1[0] - p0[0]) / 2) ** 2
        d = 2 * R * np.arcsin(np.sqrt(h))
        assert abs(O.distance(1, (lon0, lat0), (lon1, lat1)) - d) < 1e-3


def test_distance_spherical_approx():
    R = 6378137.0
    c = math.pi * R / 180.0
    # along the equator the approximation is exact in longitude
    assert abs(O.distance(2, (0.0, 0.0), (3.0, 0.0)) - 3 * c) < 1e-6
    # along a meridian: pure latitude difference
    assert abs(O.distance(2, (5.0, 10.0), (5.0, 14.0)) - 4 * c) < 1e-6
    # at 60N a degree of longitude is half as long
    assert abs(O.distance(2, (0.0, 60.0), (2.0, 60.0)) - c) < 1e-6


def test_select_inclusive_cutoff_and_weights():
    # zone at (10, 10); obs exactly at the cutoff (15) must be relevant
    ox = np.array([10.0, 10.0, 25.0, 10.0, 22.0])
    oy = np.array([10.0, 25.0, 10.0, 25.0000001, 19.0])
    w, rel, cnt = O.select(0, 1, (10.0, 10.0), 5.0, 15.0, ox, oy)
    d = np.hypot(ox - 10.0, oy - 10.0)
    assert list(rel) == [True, True, True, False, True]
    assert cnt == 4
    np.testing.assert_allclose(w, np.exp(-(d / 5.0) ** 2), rtol=1e-15)


def test_select_no_relevant_leaves_distances():
    ox = np.array([100.0, 200.0]); oy = np.array([0.0, 0.0])
    w, rel, cnt = O.select(0, 1, (0.0, 0.0), 5.0, 15.0, ox, oy)
    assert cnt == 0 and not rel.any()
    np.testing.assert_array_equal(w, [100.0, 200.0])


def test_select_vertical_and_time():
    oz = np.array([0.0, 3.0, 7.0]); ot = np.array([1.0, 5.0, 9.0])
    ox = np.zeros(3); oy = np.zeros(3)
    w, rel, cnt = O.select(0, 2, (0, 0, 4.0, 0), 2.0, 3.0, ox, oy, oz, ot)
    assert list(rel) == [False, True, True] and cnt == 2
    np.testing.assert_allclose(w, np.exp(-(np.abs(oz - 4.0) / 2.0) ** 2))
    w, rel, cnt = O.select(0, 3, (0, 0, 0, 5.0), 2.0, 4.0, ox, oy, oz, ot)
    assert list(rel) == [True, True, True] and cnt == 3


# --- matoper_inc.F90:939-1005 (dsyev 'V','L') ---
@pytest.mark.parametrize("n", [1, 2, 5, 16, 33, 64])
def test_symeig_against_numpy(backend, n):
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n + 3, n))
    A = G.T @ G
    # only the lower triangle may be read: poison the strict upper one
    Au = A.copy(order="F")
    Au[np.triu_indices(n, 1)] = 1e300
    E, V = O.symeig(Au)
    En, _ = np.linalg.eigh(A)
    assert np.all(np.diff(E) >= 0)
    np.testing.assert_allclose(E, En, rtol=1e-12, atol=1e-12 * En.max())
    assert rel_err(A @ V, V * E) < 1e-12
    assert rel_err(V.T @ V, np.eye(n)) < 1e-13


def test_symeig_degenerate(backend):
    # rank-deficient PSD matrix with a repeated zero eigenvalue (p < N case)
    rng = np.random.default_rng(7)
    G = rng.standard_normal((3, 10))
    A = G.T @ G
    E, V = O.symeig(A)
    assert np.sum(np.abs(E) < 1e-12 * E.max()) == 7
    assert rel_err(A @ V, V * E) < 1e-12


# --- matoper.F90:509-535, rrsqrt.F90:737-756 ---
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_perp_space_and_rotation(sign):
    rng = np.random.default_rng(3)
    n = 9
    w = rng.standard_normal(n); w[-1] = sign * abs(w[-1]); w /= np.linalg.norm(w)
    P = O.perp_space(w)
    assert rel_err(P.T @ P, np.eye(n - 1)) < 1e-14
    assert np.max(np.abs(P.T @ w)) < 1e-14
    v = rng.standard_normal(n); v /= np.linalg.norm(v)
    Om = O.rotate_vector(w, v)
    assert np.max(np.abs(Om @ w - v)) < 1e-14
    assert rel_err(Om.T @ Om, np.eye(n)) < 1e-14
    # RotateVector(w, w) is the identity ("generally the identity matrix")
    assert rel_err(O.rotate_vector(w, w), np.eye(n)) < 1e-14


@pytest.mark.parametrize("case", ["random+", "random-", "+e_N", "-e_N", "+0 last"])
def test_perp_space_is_a_householder_reflector(case):
    """perpSpace(u) (matoper.F90:509-535) is the first N-1 columns of the
    reflector I - 2 z z^T / z^T z with z = u + sign(u_N) e_N, Fortran's
    sign(1., u_N) = +1 for u_N >= +0 (written here from the reflector, not
    from the closed form of the oracle).  Any other orthonormal complement
    (a QR-based basis, the opposite sign convention) fails this, and it is
    what fixes the individual analysed members."""
    rng = np.random.default_rng(17)
    n = 11
    if case.startswith("random"):
        u = rng.standard_normal(n)
        u[-1] = abs(u[-1]) * (1 if case.endswith("+") else -1)
        u /= np.linalg.norm(u)
    elif case == "+0 last":
        u = rng.standard_normal(n); u[-1] = 0.0; u /= np.linalg.norm(u)
    else:
        u = np.zeros(n); u[-1] = 1.0 if case[0] == "+" else -1.0
    s = 1.0 if u[-1] >= 0 else -1.0
    z = u.copy(); z[-1] += s
    Hh = np.eye(n) - 2.0 * np.outer(z, z) / (z @ z)
    P = O.perp_space(u)
    assert np.max(np.abs(P - Hh[:, : n - 1])) < 1e-15
    if case.endswith("e_N"):
        assert np.max(np.abs(P - np.eye(n)[:, : n - 1])) == 0.0


def test_rotate_vector_members_from_the_reflectors():
    """RotateVector(w, v) = v w^T + P(v) P(w)^T (rrsqrt.F90:737-756) with
    P(.) the reflector columns above, applied to an ensemble: the analysed
    members depend on this particular basis, so Omega is pinned member by
    member (not only by Omega w = v and orthogonality)."""
    rng = np.random.default_rng(19)
    n = 7

    def refl(u):
        s = 1.0 if u[-1] >= 0 else -1.0
        z = u.copy(); z[-1] += s
        return (np.eye(n) - 2.0 * np.outer(z, z) / (z @ z))[:, : n - 1]
    for _ in range(5):
        w = rng.standard_normal(n); w /= np.linalg.norm(w)
        v = rng.standard_normal(n); v /= np.linalg.norm(v)
        Om = O.rotate_vector(w, v)
        ref = np.outer(v, w) + refl(v) @ refl(w).T
        assert np.max(np.abs(Om - ref)) < 1e-14


# --- rrsqrt.F90:100-190 against the textbook KF (test/test_rrsqrt.F90:27-77) ---
def test_analysis_global_kf(backend):
    d = rrsqrt_inputs()
    Hxf = d["H"] @ d["xf"]
    HSf = d["H"] @ d["Sf"]
    dx, Sa, _ = O.analysis_increment(Hxf, d["y"], d["Sf"], HSf, d["sig2"])
    xa_chk, Pa_chk = global_kf(d["xf"], d["Sf"], d["H"], d["y"], np.diag(d["sig2"]))
    assert rel_err(d["xf"] + dx, xa_chk) < 1e-10
    assert rel_err(Sa @ Sa.T, Pa_chk) < 1e-10
    # rotation keeps the anomalies centred: Sa . 1 = 0
    assert np.max(np.abs(Sa.sum(axis=1))) < 1e-12


def _csr_all(nz, m, w=None):
    sel_ptr = np.arange(nz + 1, dtype=np.int64) * m
    sel_idx = np.tile(np.arange(m, dtype=np.int32), nz)
    sel_w = np.ones(nz * m) if w is None else np.concatenate(w)
    return sel_ptr, sel_idx, sel_w


def test_loc_analysis_select_all_equals_global(backend):
    """test/test_rrsqrt.F90:142-159: one zone, then one zone per element."""
    d = rrsqrt_inputs()
    Hxf = d["H"] @ d["xf"]; HSf = d["H"] @ d["Sf"]
    xa_chk, Pa_chk = global_kf(d["xf"], d["Sf"], d["H"], d["y"], np.diag(d["sig2"]))
    for zone_ptr in (np.array([0, d["n"]]), np.arange(d["n"] + 1)):
        nz = len(zone_ptr) - 1
        sp, si, sw = _csr_all(nz, d["m"])
        xa, Sa = O.loc_analysis_csr(zone_ptr, sp, si, sw, d["xf"], Hxf, d["y"], d["sig2"], d["Sf"], HSf)
        assert rel_err(xa, xa_chk) < 1e-10
        if nz == 1:
            assert rel_err(Sa @ Sa.T, Pa_chk) < 1e-10
        else:
            # per-element zones: the diagonal of Pa is reproduced exactly
            assert rel_err(np.sum(Sa * Sa, axis=1), np.diag(Pa_chk)) < 1e-10


def _gc_selection(n, m, localise_obs):
    """The callback of test/test_rrsqrt.F90:254-271 (Gaspari-Cohn, radius 0.21)."""
    xmod = np.arange(n) / (n - 1.0)
    xobs = np.arange(m) / (m - 1.0)
    sp = [0]; si = []; sw = []; sets = []
    for i in range(n):
        c = np.array([O.locfun(abs(xmod[i] - xo) / 0.21) for xo in xobs])
        rel = c != 0
        if not localise_obs:
            rel = np.ones(m, bool)
        idx = np.nonzero(rel)[0]
        si.extend(idx); sw.extend(c[idx]); sp.append(len(si)); sets.append((idx, c[idx]))
    return (np.array(sp, dtype=np.int64), np.array(si, dtype=np.int32), np.array(sw)), sets


@pytest.mark.parametrize("localise_obs", [False, True])
def test_loc_analysis_gaspari_cohn_dense_kf(backend, localise_obs):
    """test/test_rrsqrt.F90:163-236: pins the w^2 taper (R_loc^-1 = R^-1 o w w^T)."""
    d = rrsqrt_inputs()
    n, m = d["n"], d["m"]
    Hxf = d["H"] @ d["xf"]; HSf = d["H"] @ d["Sf"]
    (sp, si, sw), sets = _gc_selection(n, m, localise_obs)
    zone_ptr = np.arange(n + 1)
    xa, Sa = O.loc_analysis_csr(zone_ptr, sp, si, sw, d["xf"], Hxf, d["y"], d["sig2"], d["Sf"],
                                HSf, localise_obs=localise_obs)
    for i, (idx, w) in enumerate(sets):
        xa_i, Pa_i = local_kf_rows(d["xf"], d["Sf"], d["H"], d["y"], d["sig2"], [i], idx, w)
        assert abs(xa[i] - xa_i[0]) < 1e-10 * max(1.0, abs(xa_i[0]))
        assert abs(np.sum(Sa[i] ** 2) - Pa_i[0, 0]) < 1e-10 * max(1.0, Pa_i[0, 0])


def test_loc_analysis_amplitudes_only_without_localisation():
    d = rrsqrt_inputs()
    Hxf = d["H"] @ d["xf"]; HSf = d["H"] @ d["Sf"]
    (sp, si, sw), _ = _gc_selection(d["n"], d["m"], False)
    xa, Sa, ampl = O.loc_analysis_csr(np.arange(d["n"] + 1), sp, si, sw, d["xf"], Hxf, d["y"],
                                      d["sig2"], d["Sf"], HSf, localise_obs=False, want_ampl=True)
    # xa - xf = Sf . ampl per zone (rrsqrt.F90:151)
    for i in range(d["n"]):
        assert abs(xa[i] - d["xf"][i] - d["Sf"][i] @ ampl[:, i]) < 1e-12
    (sp, si, sw), _ = _gc_selection(d["n"], d["m"], True)
    _, _, ampl = O.loc_analysis_csr(np.arange(d["n"] + 1), sp, si, sw, d["xf"], Hxf, d["y"],
                                    d["sig2"], d["Sf"], HSf, localise_obs=True, want_ampl=True)
    assert not ampl.any()


def test_empty_zones_are_identity():
    d = rrsqrt_inputs()
    Hxf = d["H"] @ d["xf"]; HSf = d["H"] @ d["Sf"]
    zone_ptr = np.array([0, 4, 10])
    sp = np.array([0, 0, 0]); si = np.zeros(0, np.int32); sw = np.zeros(0)
    xa, Sa = O.loc_analysis_csr(zone_ptr, sp, si, sw, d["xf"], Hxf, d["y"], d["sig2"], d["Sf"], HSf)
    np.testing.assert_array_equal(xa, d["xf"])
    np.testing.assert_array_equal(Sa, d["Sf"])


def test_excluded_observations_drop_out():
    """assimilation.F90:3086-3092: excluded obs get R^-1 = 0."""
    d = rrsqrt_inputs()
    Hxf = d["H"] @ d["xf"]; HSf = d["H"] @ d["Sf"]
    dex = np.array([1.0, 0.0, 1.0, 1.0, 0.0])
    keep = dex > 0
    dx1, Sa1, _ = O.analysis_increment(Hxf, d["y"], d["Sf"], HSf, d["sig2"], dex=dex)
    dx2, Sa2, _ = O.analysis_increment(Hxf[keep], d["y"][keep], d["Sf"], HSf[keep], d["sig2"][keep])
    assert rel_err(dx1, dx2) < 1e-12 and rel_err(Sa1 @ Sa1.T, Sa2 @ Sa2.T) < 1e-12


# --- matoper_inc.F90:220-242 ---
def test_obsoper_dense():
    rng = np.random.default_rng(5)
    m, n, nnz = 7, 20, 30
    Hi = rng.integers(0, m, nnz); Hj = rng.integers(0, n, nnz); Hs = rng.standard_normal(nnz)
    x = rng.standard_normal(n)
    Hd = np.zeros((m, n))
    np.add.at(Hd, (Hi, Hj), Hs)
    np.testing.assert_allclose(O.obsoper(Hi, Hj, Hs, m, x), Hd @ x, rtol=1e-13, atol=1e-13)


# --- assimilation.F90:578-641, 408-412 ---
def test_partition_stable_with_gaps():
    part = np.array([5, 2, 5, 9, 2, 2, 9])
    zs, zi, izi = O.partition(part)
    assert list(zs) == [3, 2, 2]
    assert list(zi) == [1, 4, 5, 0, 2, 3, 6]
    assert all(zi[izi[i]] == i for i in range(len(part)))


def test_ensemble_analysis_geometry_against_dense_kf(backend):
    """Ensemble path end to end on a small grid, zone by zone against the dense
    KF with the tapered R (independent formulas), plus ensemble invariants."""
    rng = np.random.default_rng(11)
    nx, ny, N, m = 8, 6, 10, 14
    n = nx * ny
    E = rng.standard_normal((n, N)) + 0.3
    gx, gy = np.meshgrid(np.arange(nx, dtype=float), np.arange(ny, dtype=float), indexing="ij")
    gx = gx.ravel(order="F"); gy = gy.ravel(order="F")
    ox = rng.uniform(0, nx - 1, m); oy = rng.uniform(0, ny - 1, m)
    # bilinear H, 4 nnz per obs
    Hi, Hj, Hs = [], [], []
    for l in range(m):
        i0 = min(int(ox[l]), nx - 2); j0 = min(int(oy[l]), ny - 2)
        a = ox[l] - i0; b = oy[l] - j0
        for di, dj, c in ((0, 0, (1 - a) * (1 - b)), (1, 0, a * (1 - b)), (0, 1, (1 - a) * b), (1, 1, a * b)):
            Hi.append(l); Hj.append((i0 + di) + nx * (j0 + dj)); Hs.append(c)
    Hi = np.array(Hi); Hj = np.array(Hj); Hs = np.array(Hs)
    Hd = np.zeros((m, n)); np.add.at(Hd, (Hi, Hj), Hs)
    yo = rng.standard_normal(m); sig2 = np.full(m, 0.25); Hshift = np.zeros(m)
    geo = dict(zone_ptr=np.arange(n + 1), zpos=np.stack([gx, gy, 0 * gx, 0 * gx], 1),
               L=np.full(n, 2.0), cutoff=np.full(n, 4.0), metric=0, loctype=1, ox=ox, oy=oy)
    Ea, xa, counts = O.ensemble_analysis(geo, (Hi, Hj, Hs), Hshift, yo, sig2, E)
    xf = E.mean(axis=1); Sf = (E - xf[:, None]) / np.sqrt(N - 1)
    assert rel_err(Ea.mean(axis=1), xa) < 1e-13
    Sa = (Ea - xa[:, None]) / np.sqrt(N - 1)
    for z in range(n):
        d = np.hypot(ox - gx[z], oy - gy[z])
        idx = np.nonzero(d <= 4.0)[0]
        assert counts[z] == len(idx)
        if len(idx) == 0:
            assert abs(xa[z] - xf[z]) < 1e-14
            continue
        w = np.exp(-(d[idx] / 2.0) ** 2)
        xa_z, Pa_z = local_kf_rows(xf, Sf, Hd, yo, sig2, [z], idx, w)
        assert abs(xa[z] - xa_z[0]) < 1e-10 * max(1, abs(xa_z[0]))
        assert abs(Sa[z] @ Sa[z] - Pa_z[0, 0]) < 1e-10 * max(1, Pa_z[0, 0])


@pytest.mark.parametrize("per_element", [False, True])
def test_post_processing_inflation_and_clamp_against_dense_kf(backend, per_element):
    """Assim's post-processing (assimilation.F90:3299-3320, 3347-3357,
    462-476) against the per-zone dense Kalman filter: with inflation 1.3 and
    a maxCorrection that binds, mean(Ea) = clip(xa_KF, xf +- maxCorrection)
    and var(Ea, ddof=1) = inflation^2 Pa_KF (the clamp moves the mean only;
    inflation scales the spread only); zones without observations keep
    inflation^2 Pf.  Two variables per zone (A observed, B correlated with A)
    with their own clamps in the per-element case.  Inflating the mean,
    clamping the members instead of xa, or inflating before the transform's
    re-centring fail this."""
    rng = np.random.default_rng(23)
    nx, ny, N, m = 9, 7, 12, 16
    npt = nx * ny
    n = 2 * npt
    A0 = rng.standard_normal((npt, N))
    B0 = 0.7 * A0 + 0.5 * rng.standard_normal((npt, N))
    E = np.empty((n, N))
    E[0::2] = A0 + 0.2
    E[1::2] = B0 - 0.1
    gx, gy = np.meshgrid(np.arange(nx, dtype=float), np.arange(ny, dtype=float), indexing="ij")
    gx = gx.ravel(order="F"); gy = gy.ravel(order="F")
    # observations in the left two thirds only: the right columns have no
    # observation within the cutoff
    ox = rng.uniform(0, 4.5, m); oy = rng.uniform(0, ny - 1, m)
    Hi, Hj, Hs = [], [], []
    for l in range(m):
        i0 = min(int(ox[l]), nx - 2); j0 = min(int(oy[l]), ny - 2)
        a = ox[l] - i0; b = oy[l] - j0
        for di, dj, c in ((0, 0, (1 - a) * (1 - b)), (1, 0, a * (1 - b)), (0, 1, (1 - a) * b),
                          (1, 1, a * b)):
            Hi.append(l); Hj.append(2 * ((i0 + di) + nx * (j0 + dj))); Hs.append(c)
    Hi = np.array(Hi); Hj = np.array(Hj); Hs = np.array(Hs)
    Hd = np.zeros((m, n)); np.add.at(Hd, (Hi, Hj), Hs)
    yo = rng.standard_normal(m) * 2.0 + 1.0
    sig2 = np.full(m, 0.04); Hshift = np.zeros(m)
    infl = 1.3
    mc = np.tile([0.05, 0.2], npt) if per_element else 0.08
    geo = dict(zone_ptr=2 * np.arange(npt + 1), zpos=np.stack([gx, gy, 0 * gx, 0 * gx], 1),
               L=np.full(npt, 1.5), cutoff=np.full(npt, 3.0), metric=0, loctype=1, ox=ox, oy=oy)
    Ea, xa, counts = O.ensemble_analysis(geo, (Hi, Hj, Hs), Hshift, yo, sig2, E,
                                         inflation=infl, max_correction=mc)
    xf = E.mean(axis=1); Sf = (E - xf[:, None]) / np.sqrt(N - 1)
    mcv = np.broadcast_to(mc, (n,))
    nclamp = nempty = 0
    for z in range(npt):
        rows = [2 * z, 2 * z + 1]
        d = np.hypot(ox - gx[z], oy - gy[z])
        idx = np.nonzero(d <= 3.0)[0]
        assert counts[z] == len(idx)
        if len(idx) == 0:
            nempty += 1
            xk, Pk = xf[rows], (Sf[rows] @ Sf[rows].T)
        else:
            w = np.exp(-(d[idx] / 1.5) ** 2)
            xk, Pk = local_kf_rows(xf, Sf, Hd, yo, sig2, rows, idx, w)
        xc = np.clip(xk, xf[rows] - mcv[rows], xf[rows] + mcv[rows])
        nclamp += int(np.sum(xc != xk))
        assert np.max(np.abs(Ea[rows].mean(axis=1) - xc)) < 1e-10 * max(1, np.abs(xc).max())
        var = Ea[rows].var(axis=1, ddof=1)
        assert np.max(np.abs(var - infl ** 2 * np.diag(Pk))) < 1e-10 * max(1, np.diag(Pk).max())
        cov = np.cov(Ea[rows])
        assert abs(cov[0, 1] - infl ** 2 * Pk[0, 1]) < 1e-10 * max(1, np.abs(Pk).max())
    assert nclamp > 10 and nempty > 5, (nclamp, nempty)


# --- report, assimilation.F90:3584-3628 ---
def test_report_known_answer():
    # d = [1, 2, -2], sig2 = [1, 4, 2], third observation excluded:
    # ingrid 2, rms sqrt(5/2), bias 3/2, R^-1 d = [1, 1/2, 0],
    # d.R^-1 d = 2, HS^T R^-1 d = [1, 1/2]
    d = np.array([1.0, 2.0, -2.0])
    sig2 = np.array([1.0, 4.0, 2.0])
    dex = np.array([1.0, 1.0, 0.0])
    HS = np.array([[1.0, 0.0], [0.0, 1.0], [5.0, 5.0]])
    out = O.report(d, HS, sig2, dex)
    np.testing.assert_allclose(out, [2.0, math.sqrt(2.5), 1.5, 1.0, math.sqrt(0.625)],
                               rtol=1e-15)
    # nothing excluded: plain DiagCovar
    out = O.report(d, HS, sig2)
    r = d / sig2
    np.testing.assert_allclose(
        out, [3.0, math.sqrt(9 / 3), 1.0 / 3, math.sqrt((1 + 1 + 2) / 3),
              math.sqrt(((r[0] + 5 * r[2]) ** 2 + (r[1] + 5 * r[2]) ** 2) / 3)], rtol=1e-15)


def _report_case(seed, m=300, N=12):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=m)
    HS = rng.normal(size=(m, N))
    sig2 = rng.uniform(0.2, 3.0, size=m)
    dex = (rng.random(m) > 0.2).astype(float)
    return d, HS, sig2, dex


def test_report_scaling_of_the_innovation():
    d, HS, sig2, dex = _report_case(1)
    a = O.report(d, HS, sig2, dex)
    b = O.report(-3.0 * d, HS, sig2, dex)
    np.testing.assert_allclose(b, a * [1, 3, -3, 3, 3], rtol=1e-13)


def test_report_excluded_observation_equals_removed():
    d, HS, sig2, dex = _report_case(2)
    keep = dex != 0
    a = O.report(d, HS, sig2, dex)
    b = O.report(d[keep], HS[keep], sig2[keep])
    np.testing.assert_allclose(a, b, rtol=1e-13)


def test_report_ranges_add_up():
    # checksum of checksums: ingrid * stat^2 is additive over observation
    # ranges for rms, rms_invsqrtR; ingrid * bias as well
    d, HS, sig2, dex = _report_case(3, m=500)
    cut = [0, 123, 124, 380, 500]
    parts = np.array([O.report(d[a:b], HS[a:b], sig2[a:b], dex[a:b])
                      for a, b in zip(cut[:-1], cut[1:])])
    tot = O.report(d, HS, sig2, dex)
    g = parts[:, 0]
    assert g.sum() == tot[0]
    np.testing.assert_allclose((g * parts[:, 1] ** 2).sum(), tot[0] * tot[1] ** 2, rtol=1e-13)
    np.testing.assert_allclose((g * parts[:, 2]).sum(), tot[0] * tot[2], rtol=1e-12)
    np.testing.assert_allclose((g * parts[:, 3] ** 2).sum(), tot[0] * tot[3] ** 2, rtol=1e-13)


def test_report_all_excluded_is_nan():
    d, HS, sig2, _ = _report_case(4, m=10)
    out = O.report(d, HS, sig2, np.zeros(10))
    assert out[0] == 0 and np.all(np.isnan(out[1:]))


def test_ensemble_report_against_sample_covariance():
    rng = np.random.default_rng(5)
    n, N, m = 60, 9, 40
    E = rng.normal(size=(n, N)) + np.linspace(0, 3, n)[:, None]
    Hi = np.repeat(np.arange(m), 2)
    Hj = rng.integers(0, n, size=2 * m)
    Hs = rng.uniform(0.2, 1.0, size=2 * m)
    Hshift = rng.normal(size=m) * 0.1
    yo = rng.normal(size=m)
    sig2 = rng.uniform(0.5, 2.0, size=m)
    dex = (rng.random(m) > 0.25).astype(float)
    Hd = np.zeros((m, n))
    np.add.at(Hd, (Hi, Hj), Hs)
    HE = Hd @ E + Hshift[:, None]
    dd = yo - HE.mean(axis=1)
    cut = [0, 17, 40]
    out = O.ensemble_report((Hi, Hj, Hs), Hshift, yo, sig2, E, dex, ranges=cut)
    for k, (a, b) in enumerate(zip(cut[:-1], cut[1:])):
        v = dex[a:b] != 0
        g = v.sum()
        r = np.where(v, dd[a:b] / sig2[a:b], 0.0)
        P = np.cov(HE[a:b])  # sample covariance (ddof 1) = HS HS^T
        exp = [g, math.sqrt((dd[a:b][v] ** 2).sum() / g), dd[a:b][v].sum() / g,
               math.sqrt((dd[a:b] * r).sum() / g), math.sqrt(r @ P @ r / g)]
        np.testing.assert_allclose(out[k], exp, rtol=1e-11)


# --- the committed golden fixture (tests/golden/make_golden.py) ---
def test_oracle_reproduces_golden_fixture():
    """The oracle still gives the committed expected outputs of the small
    ensemble case (inflation, a binding maxCorrection, two excluded
    observations, forecast and analysis reports)."""
    import os
    from tests.golden.make_golden import expected
    d = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "ens_small.npz")))
    got = expected(d)
    xf = d["E"].mean(axis=1)[:, None]
    err = np.abs((got["Ea"] - xf) - (d["Ea"] - xf)).max() / np.abs(d["Ea"] - xf).max()
    assert err < 1e-10, err
    np.testing.assert_allclose(got["report_forecast"], d["report_forecast"], rtol=1e-12)
    np.testing.assert_allclose(got["report_analysis"], d["report_analysis"], rtol=1e-10)
