"""The Fortran bindings (oak_amd/fortran, iso_c_binding) link against
liboakg.so and report the library's error codes: without a GPU the driver's
locAnalysis_gpu returns OAKG_EHIP."""
import os
import subprocess

import pytest

from tests.fdriver_ref import DRIVER, read


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (build())")
def test_fortran_driver_without_gpu(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible: covered by tests/test_gpu_fortran.py")
    out = tmp_path / "fd.bin"
    r = subprocess.run([DRIVER, str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3, r.stdout + r.stderr
    assert "info = 3" in r.stdout
    d = read(out)
    assert d["info"] == 3 and d["info2"] == 3 and d["info3"] == 3
    assert d["n"] == 20 and d["N"] == 6 and d["m"] == 8
