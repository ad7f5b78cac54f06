"""Fortran host code calling the GPU path through the iso_c_binding layer
(oak_amd/fortran): locAnalysis_gpu, with a selectObservations callback, against
the oracle.  Needs an MI355X."""
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
    assert d["info"] == 0
    xa, Sa = oracle_analysis(d)
    dx = np.abs(d["xa"] - xa).max() / np.abs(xa - d["xf"]).max()
    ds = np.abs(d["Sa"] - Sa).max() / np.abs(Sa).max()
    assert dx < 1e-10 and ds < 1e-10, (dx, ds)
