"""State layout and I/O on the GPU: the zone permutation of ensembles
(oakg_gather_rows / oakg_scatter_rows) bit for bit against numpy indexing,
and OAK's cycle end to end -- member files in GHER format with a land mask,
the ensemble loaded into the zone-permuted state, the observation operator
generated from the grid coordinates, the local analysis, the analysed
ensemble written back -- against the oracle run on the same files (numpy
packing and permutation, the oracle's operator and analysis).  Needs an
MI355X."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gather_scatter_bit_exact(dtype):
    from oak_amd import _lib
    from oak_amd import layout as Lay
    g = np.random.default_rng(5)
    full_len, n, N = 5000, 3700, 7
    idx = np.sort(g.choice(full_len, n, replace=False))
    g.shuffle(idx[:500])
    src = torch.as_tensor(g.standard_normal((N, full_len))).to("cuda:0", dtype).t()
    out = torch.empty((N, n), dtype=dtype, device="cuda:0").t()
    Lay._permute(True, idx, full_len, src, out, None)
    ref = src.cpu().numpy()[idx]
    assert np.array_equal(out.cpu().numpy(), ref)
    back = torch.full((N, full_len), -3.0, dtype=dtype, device="cuda:0").t()
    Lay._permute(False, idx, full_len, out, back, None)
    b = back.cpu().numpy()
    assert np.array_equal(b[idx], ref)
    rest = np.setdiff1d(np.arange(full_len), idx)
    assert np.all(b[rest] == -3.0)
    # an index outside the full vector is an argument error
    bad = idx.copy(); bad[3] = full_len
    with pytest.raises(_lib.OakError):
        Lay._permute(True, bad, full_len, src, out, None)


def _world(tmp_path, seed=11, nx=12, ny=9, nz=3, N=10):
    """Member files of T (nx, ny, nz) and SSH (nx, ny) with land, the grid
    coordinates (curvilinear lon/lat, depth levels) and observations."""
    from oak_amd import synth
    from oak_amd.ufile import uwrite
    g = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(nx, dtype=float), np.arange(ny, dtype=float), indexing="ij")
    lon = 10.0 + 0.5 * i + 0.05 * np.sin(0.7 * j)
    lat = 40.0 + 0.4 * j + 0.04 * np.cos(0.9 * i)
    depth = np.array([0.0, 20.0, 60.0])
    land2 = (i + 2 * j < 4) | ((i > 8) & (j > 6))
    landT = np.repeat(land2[:, :, None], nz, 2)
    landT[:, :, 2] |= g.uniform(size=(nx, ny)) < 0.3  # shallow columns
    fT = synth.grf(g, (nx, ny, nz), (4, 4, 2), N)
    fS = synth.grf(g, (nx, ny), (4, 4), N)
    files = []
    for k in range(N):
        t = np.where(landT, 9999.0, 10 + fT[k]); s = np.where(land2, 9999.0, 0.3 * fS[k])
        ft, fs = tmp_path / ("T_%02d.u" % k), tmp_path / ("SSH_%02d.u" % k)
        uwrite(str(ft), t); uwrite(str(fs), s)
        files.append([str(ft), str(fs)])
    uwrite(str(tmp_path / "maskT.u"), np.where(landT, 9999.0, 1.0))
    uwrite(str(tmp_path / "maskS.u"), np.where(land2, 9999.0, 1.0))
    c2 = [lon, lat]
    c3 = [np.repeat(lon[:, :, None], nz, 2), np.repeat(lat[:, :, None], nz, 2),
          np.broadcast_to(depth, (nx, ny, nz)).copy()]
    ms, mt = 50, 40
    sx = g.uniform(lon.min(), lon.max(), ms); sy = g.uniform(lat.min(), lat.max(), ms)
    tx = g.uniform(lon.min(), lon.max(), mt); ty = g.uniform(lat.min(), lat.max(), mt)
    tz = g.uniform(0, 60, mt)
    obs = dict(var=np.concatenate([np.ones(ms, np.int32), np.zeros(mt, np.int32)]),
               x=np.concatenate([sx, tx]), y=np.concatenate([sy, ty]),
               z=np.concatenate([np.zeros(ms), tz]),
               yo=np.concatenate([0.1 * g.standard_normal(ms), 10 + g.standard_normal(mt)]),
               r2=np.concatenate([np.full(ms, 0.03 ** 2), np.full(mt, 0.5 ** 2)]))
    return dict(files=files, c2=c2, c3=c3, obs=obs, nx=nx, ny=ny, nz=nz, N=N,
                land2=land2, landT=landT, lon=lon, lat=lat)


def test_cycle_files_to_files_vs_oracle(tmp_path):
    from oak_amd import assimilation as A
    from oak_amd import layout as Lay
    from oak_amd.ufile import uread
    W = _world(tmp_path)
    nx, ny, nz, N = W["nx"], W["ny"], W["nz"], W["N"]
    lay = Lay.mem_layout([uread(str(tmp_path / "maskT.u")), uread(str(tmp_path / "maskS.u"))])
    # column zones: label = horizontal position of every state element
    labels = np.concatenate([np.nonzero(lay.mask[lay.start[v]:lay.start[v] + lay.size[v]])[0]
                             % (nx * ny) for v in range(2)])
    zones = Lay.zone_layout(labels)
    E = Lay.load_ensemble(W["files"], lay, zones)
    grids = [dict(coords=W["c3"], mask=W["landT"], state=Lay.state_index_map(lay, zones, 0)),
             dict(coords=W["c2"], mask=W["land2"], state=Lay.state_index_map(lay, zones, 1))]
    o = W["obs"]
    rp, col, val, valid = A.obs_operator(grids, o["var"], [o["x"], o["y"], o["z"]])
    m = len(o["var"])
    # zone positions: the first element of each zone (its column)
    zp = zones.zone_ptr
    first_full = lay.invindex[zones.zone_index[zp[:-1]]]
    ij = np.where(first_full < lay.size[0], first_full % (nx * ny), first_full - lay.size[0])
    zpos = np.zeros((len(zp) - 1, 4))
    zpos[:, 0] = W["lon"].ravel(order="F")[ij]
    zpos[:, 1] = W["lat"].ravel(order="F")[ij]
    nzn = len(zp) - 1
    L, cut = np.full(nzn, 1.2), np.full(nzn, 3.0)
    geo = A.Geometry(zp, zpos, L, cut, o["x"], o["y"])
    ex = (~valid).astype(np.uint8)
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(dev, dt)
    gd = A.Geometry(t(zp, torch.int64), t(zpos), t(L), t(cut), t(o["x"]), t(o["y"]))
    Hd = (t(rp, torch.int64), t(col, torch.int32), t(val))
    Ea = A.ensemble_analysis(gd, Hd, t(np.zeros(m)), t(o["yo"]), t(o["r2"]), E,
                             excluded=t(ex, torch.uint8))
    outs = [[str(tmp_path / ("aT_%02d.u" % k)), str(tmp_path / ("aS_%02d.u" % k))]
            for k in range(N)]
    Lay.save_ensemble(Ea, outs, lay, zones)
    # the oracle on the same files, packed and permuted with numpy
    full = np.stack([np.concatenate([uread(f)[0].ravel(order="F") for f in mem])
                     for mem in W["files"]], 1)
    Eh = full[lay.invindex[zones.zone_index]]
    rpo, colo, valo, valido = O.obs_operator(grids, o["var"], [o["x"], o["y"], o["z"]])
    assert np.array_equal(colo, col) and np.array_equal(valo, val)
    Hi = np.repeat(np.arange(m), np.diff(rpo)).astype(np.int32)
    Eo, _, cnt = O.ensemble_analysis(dict(zone_ptr=zp, zpos=zpos, L=L, cutoff=cut, metric=0,
                                          loctype=1, ox=o["x"], oy=o["y"]),
                                     (Hi, colo, valo), np.zeros(m), o["yo"], o["r2"], Eh,
                                     dex=valido.astype(float))
    assert (cnt > 0).sum() > 10
    back = np.stack([np.concatenate([uread(f)[0].ravel(order="F") for f in mem])
                     for mem in outs], 1)
    sea = lay.mask
    xf = Eh.mean(axis=1)
    ref = np.full_like(full, 9999.0)
    ref[lay.invindex[zones.zone_index]] = Eo
    d = np.abs(back[sea] - ref[sea]).max() / np.abs(Eo - xf[:, None]).max()
    assert d < 1e-10, d
    assert np.all(back[~sea] == 9999.0)
    assert not np.allclose(back[sea], full[sea])
