"""Bin range of the binned Cartesian selection (oak_amd/csrc/select_cells.hpp),
compiled for the host with g++: zones outside the observations' bounding box
(left, right, above, below, far away, NaN) get either an empty range or bins
inside the grid, and for zones anywhere every observation within the cutoff
lies in a bin of the range (brute force over random geometries)."""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <stdio.h>
#include <stdlib.h>
#include "select_cells.hpp"
int main() {
  oakg::GridInfo g;
  double zx, zy, c;
  while (scanf("%lf %lf %lf %d %d %lf %lf %lf", &g.x0, &g.y0, &g.h, &g.gx, &g.gy, &zx, &zy, &c) == 8) {
    int cx0, cx1, cy0, cy1;
    bool ok = oakg::zone_cells(g, zx, zy, c, cx0, cx1, cy0, cy1);
    printf("%d %d %d %d %d\n", ok ? 1 : 0, cx0, cx1, cy0, cy1);
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def cells(tmp_path_factory):
    d = tmp_path_factory.mktemp("cells")
    src = d / "cells.cpp"
    src.write_text(DRIVER)
    exe = d / "cells"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "oak_amd", "csrc"),
                           str(src), "-o", str(exe)])

    def run(rows):
        inp = "\n".join(" ".join(repr(float(v)) if i not in (3, 4) else str(int(v))
                                 for i, v in enumerate(r)) for r in rows) + "\n"
        out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True)
        return [tuple(int(t) for t in l.split()) for l in out.stdout.splitlines()]
    return run


def _valid(res, gx, gy):
    ok, cx0, cx1, cy0, cy1 = res
    if not ok:
        # no bin is read: the loops over [cy0, cy1] and [cx0, cx1] are empty
        return cy0 > cy1 or cx0 > cx1
    return 0 <= cx0 <= cx1 < gx and 0 <= cy0 <= cy1 < gy


def test_zones_outside_the_box(cells):
    """Observations in [10, 20] x [5, 9]; bins of side 1 (gx = 11, gy = 5)."""
    x0, y0, h, gx, gy = 10.0, 5.0, 1.0, 11, 5
    zones = [(0.0, 0.5, 1.0),      # left of the box, below it: the case that read cell_ptr[-8]
             (0.0, 7.0, 1.0),      # left
             (40.0, 7.0, 1.0),     # right
             (15.0, -30.0, 1.0),   # below
             (15.0, 60.0, 1.0),    # above
             (1e300, 7.0, 1.0),    # far away (int conversion would overflow)
             (-1e300, -1e300, 5.0),
             (float("nan"), 7.0, 1.0),
             (15.0, 7.0, float("inf")),  # unbounded cutoff: the whole grid
             (8.5, 7.0, 1.0),      # left, disc reaches x = 9.5 < 10: misses
             (9.5, 7.0, 1.0),      # disc reaches x = 10.5: first column
             (21.5, 7.0, 1.0),     # right of the box, disc reaches back into the last bin [20, 21)
             (15.0, 7.0, 1.0)]     # inside
    res = cells([(x0, y0, h, gx, gy, zx, zy, c) for zx, zy, c in zones])
    for z, r in zip(zones, res):
        assert _valid(r, gx, gy), (z, r)
    assert res[0][0] == 0 and res[1][0] == 0 and res[2][0] == 0
    assert res[3][0] == 0 and res[4][0] == 0 and res[5][0] == 0 and res[6][0] == 0
    assert res[7][0] == 0
    assert res[8][:5] == (1, 0, gx - 1, 0, gy - 1)
    assert res[9][0] == 0
    assert res[10][0] == 1 and res[10][1] == 0
    assert res[12] == (1, 3, 6, 0, 3)  # margins of 1e-7 bin at the exact bin edges


def test_every_observation_within_the_cutoff_is_in_range(cells):
    """Random observation clouds, bin sides and zones inside and around the
    box: each observation at distance <= c from the zone lies in a bin of the
    range (bins assigned as k_obs_cell assigns them)."""
    g = np.random.default_rng(5)
    rows, cases = [], []
    for _ in range(60):
        m = int(g.integers(1, 200))
        ox = g.uniform(-50, 50, m) * g.uniform(0.1, 3)
        oy = g.uniform(-20, 80, m) * g.uniform(0.1, 3)
        c = float(g.uniform(0.5, 30))
        x0, y0 = ox.min(), oy.min()
        ex, ey = ox.max() - x0, oy.max() - y0
        h = max(c, max(ex, ey) / 2048.0)
        gx, gy = int(math.floor(ex / h)) + 1, int(math.floor(ey / h)) + 1
        for _ in range(20):
            zx = g.uniform(x0 - 3 * c - 10, x0 + ex + 3 * c + 10)
            zy = g.uniform(y0 - 3 * c - 10, y0 + ey + 3 * c + 10)
            rows.append((x0, y0, h, gx, gy, zx, zy, c))
            cases.append((ox, oy, x0, y0, h, gx, gy, zx, zy, c))
    res = cells(rows)
    hits = 0
    for (ox, oy, x0, y0, h, gx, gy, zx, zy, c), r in zip(cases, res):
        assert _valid(r, gx, gy), r
        d = np.sqrt((zx - ox) ** 2 + (zy - oy) ** 2)
        inside = d <= c
        if not inside.any():
            continue
        hits += 1
        assert r[0] == 1
        bx = np.clip(np.floor((ox[inside] - x0) / h), 0, gx - 1)
        by = np.clip(np.floor((oy[inside] - y0) / h), 0, gy - 1)
        assert np.all((bx >= r[1]) & (bx <= r[2]) & (by >= r[3]) & (by <= r[4]))
    assert hits > 100
