"""Fortran host code calling the GPU path through the iso_c_binding layer
(oak_amd/fortran): locAnalysis_gpu, with a selectObservations callback, and
report_gpu, against the oracle.  Needs an MI355X."""
import os
import subprocess

import numpy as np
import pytest

from tests.fdriver_ref import DRIVER, oracle_analysis, read

pytestmark = pytest.mark.gpu


def test_fortran_locanalysis_vs_oracle(tmp_path):
    assert os.path.exists(DRIVER), "build() builds oak_amd/fortran/oakg_fdriver"
    out = tmp_path / "fd.bin"
    r = subprocess.run([DRIVER, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    d = read(out)
    assert d["info"] == 0 and d["info2"] == 0 and d["info3"] == 0
    for loc, kx, ks in ((True, "xa", "Sa"), (False, "xa2", "Sa2"), (True, "xa3", "Sa3")):
        xa, Sa = oracle_analysis(d, localise_obs=loc)
        dx = np.abs(d[kx] - xa).max() / np.abs(xa - d["xf"]).max()
        ds = np.abs(d[ks] - Sa).max() / np.abs(Sa).max()
        assert dx < 1e-10 and ds < 1e-10, (loc, dx, ds)
    # the last zone has no observation within the cutoff: untouched
    last = slice(d["n"] - d["zsize"], d["n"])
    assert np.array_equal(d["xa"][last], d["xf"][last])
    assert np.array_equal(d["Sa2"][last], d["Sf"][last])
    # the callback called with the zone number instead of the first state
    # index (the mistake this pins) gives a different analysis
    xz, _ = oracle_analysis(d, zone_number=True)
    assert np.abs(d["xa"] - xz).max() > 1e-3 * np.abs(xz - d["xf"]).max()
    # geometry mode (selection on the GPU) = callback mode: the same sets,
    # weights equal up to the last bit of exp (device vs host libm)
    dx = np.abs(d["xa3"] - d["xa"]).max() / np.abs(d["xa"] - d["xf"]).max()
    ds = np.abs(d["Sa3"] - d["Sa"]).max() / np.abs(d["Sa"]).max()
    assert dx < 1e-13 and ds < 1e-13, (dx, ds)


def test_fortran_report_vs_oracle(tmp_path):
    """report_gpu's log lines (assimilation.F90:3606-3620 formats) for the
    driver's forecast, last observation excluded."""
    from oracle import oracle as O
    out = tmp_path / "fd.bin"
    r = subprocess.run([DRIVER, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    d = read(out)
    vals = {}
    for line in r.stdout.splitlines():
        if line.startswith("forecast."):
            key, v = line[len("forecast."):].split()
            vals[key] = float(v)
    assert vals["mjd"] == 55000.5
    dex = np.ones(d["m"])
    dex[-1] = 0
    exp = O.report(d["yo"] - d["Hxf"], d["HSf"], d["Rd"], dex)
    got = [vals["rms_yo-Hx"], vals["bias_yo-Hx"], vals["rms_invsqrtR_yo-Hx"],
           vals["projection_yo-Hx_into_HSf"]]
    # E15.7: seven significant digits
    np.testing.assert_allclose(got, exp[1:], rtol=1e-6)
