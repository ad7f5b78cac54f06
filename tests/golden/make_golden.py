"""Writes tests/golden/ens_small.npz: a small ensemble analysis (12 x 10
grid, one zone per grid point, N = 8, 30 observations with two excluded,
bilinear H, Gaussian taper L = 2.5, cutoff 6) and its expected outputs from
the oracle (oracle/oak_oracle.c): the analysed ensemble of Assim's ensemble
path (inflation 1.1, maxCorrection 0.2, which binds) and the forecast and analysis
report statistics.  Run from the repository root:

    python3 tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402


def inputs():
    rng = np.random.default_rng(20261202)
    nx, ny, N, m = 12, 10, 8, 30
    n = nx * ny
    gi, gj = np.meshgrid(np.arange(1, nx + 1.0), np.arange(1, ny + 1.0), indexing="ij")
    base = np.sin(gi / 3.0) + np.cos(gj / 4.0)
    E = np.asfortranarray(base.ravel(order="F")[:, None] + 0.4 * rng.standard_normal((n, N)))
    ox = rng.uniform(1, nx, m)
    oy = rng.uniform(1, ny, m)
    i0 = np.minimum(np.floor(ox).astype(int), nx - 1)
    j0 = np.minimum(np.floor(oy).astype(int), ny - 1)
    fx, fy = ox - i0, oy - j0
    Hi, Hj, Hs = [], [], []
    for l in range(m):
        for di, dj, wgt in ((0, 0, (1 - fx[l]) * (1 - fy[l])), (1, 0, fx[l] * (1 - fy[l])),
                            (0, 1, (1 - fx[l]) * fy[l]), (1, 1, fx[l] * fy[l])):
            Hi.append(l)
            Hj.append((i0[l] + di - 1) + nx * (j0[l] + dj - 1))
            Hs.append(wgt)
    zpos = np.zeros((n, 4))
    zpos[:, 0] = gi.ravel(order="F")
    zpos[:, 1] = gj.ravel(order="F")
    yo = np.sin(ox / 3.0) + np.cos(oy / 4.0) + 0.1 * rng.standard_normal(m)
    rmse2 = rng.uniform(0.005, 0.02, m)
    excluded = np.zeros(m, np.uint8)
    excluded[[4, 17]] = 1
    return dict(E=E, Hi=np.array(Hi, np.int32), Hj=np.array(Hj, np.int32), Hs=np.array(Hs),
                Hshift=np.zeros(m), yo=yo, rmse2=rmse2, excluded=excluded, ox=ox, oy=oy,
                zone_ptr=np.arange(n + 1, dtype=np.int64), zpos=zpos, L=np.full(n, 2.5),
                cutoff=np.full(n, 6.0), inflation=np.float64(1.1), max_correction=np.float64(0.2))


def expected(d):
    O.load_lapack()
    geo = dict(zone_ptr=d["zone_ptr"], zpos=d["zpos"], L=d["L"], cutoff=d["cutoff"], metric=0,
               loctype=1, ox=d["ox"], oy=d["oy"])
    dex = 1.0 - d["excluded"].astype(float)
    coo = (d["Hi"], d["Hj"], d["Hs"])
    Ea, xa, _ = O.ensemble_analysis(geo, coo, d["Hshift"], d["yo"], d["rmse2"], d["E"], dex=dex,
                                    inflation=float(d["inflation"]),
                                    max_correction=float(d["max_correction"]))
    rf = O.ensemble_report(coo, d["Hshift"], d["yo"], d["rmse2"], d["E"], dex)
    ra = O.ensemble_report(coo, d["Hshift"], d["yo"], d["rmse2"], Ea, dex)
    return dict(Ea=Ea, report_forecast=rf[0], report_analysis=ra[0])


if __name__ == "__main__":
    d = inputs()
    d.update(expected(d))
    np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ens_small.npz"), **d)
    print("wrote ens_small.npz:", {k: np.shape(v) for k, v in d.items()})
