"""CPU checks of the C ABI library: it loads, exports every function that
include/oakg.h declares, and its host-only helpers (COO -> CSR, partition)
agree with the oracle.  No GPU work is started here."""
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^(?:int|int64_t|void|const char\s*\*)\s*(oakg_[a-z0-9_]+)\s*\(",
                                   txt, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    import ctypes
    from oak_amd import _lib
    L = _lib.load()
    declared = _declared()
    assert declared, "no declarations found"
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert set(_lib.EXPORTS) <= declared


def test_coo_to_csr_is_stable():
    from oak_amd.assimilation import coo_to_csr
    rows = np.array([2, 0, 2, 1, 0, 2])
    cols = np.array([5, 1, 3, 4, 0, 9])
    vals = np.arange(6.0)
    rp, co, va = coo_to_csr(4, rows, cols, vals)
    assert list(rp) == [0, 2, 3, 6, 6]
    assert list(co) == [1, 0, 4, 5, 3, 9]
    assert list(va) == [1.0, 4.0, 3.0, 0.0, 2.0, 5.0]


def test_coo_to_csr_rejects_bad_rows():
    from oak_amd import OakError
    from oak_amd.assimilation import coo_to_csr
    with pytest.raises(OakError):
        coo_to_csr(2, [0, 2], [0, 0], [1.0, 1.0])


def test_partition_matches_oracle():
    from oak_amd.assimilation import partition
    rng = np.random.default_rng(3)
    part = rng.integers(1, 40, 500) * 3  # gaps in the labels
    zs, zi, izi = partition(part)
    zs2, zi2, izi2 = O.partition(part)
    np.testing.assert_array_equal(zs, zs2)
    np.testing.assert_array_equal(zi, zi2)
    np.testing.assert_array_equal(izi, izi2)


def test_context_creation_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from oak_amd import Context, OakError
    with pytest.raises(OakError):
        Context(0)


def test_header_compiles_as_c99(tmp_path):
    """include/oakg.h is plain C: a C99 translation unit that takes the
    address of every entry point compiles without warnings."""
    import subprocess
    src = tmp_path / "use.c"
    names = sorted(_declared())
    body = "\n".join("  (void)&%s;" % n for n in names)
    src.write_text('#include "oakg.h"\nint main(void) {\n%s\n  return 0;\n}\n' % body)
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-pedantic", "-c",
                        "-I", os.path.join(ROOT, "include"), str(src), "-o",
                        str(tmp_path / "use.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_report_log_lines_use_fortran_formats():
    """write_report mirrors report's '(A,A,F15.4)' and '(A,A,E15.7)' lines
    (assimilation.F90:3606-3620)."""
    import io
    from oak_amd.assimilation import _fortran_e, write_report
    assert _fortran_e(1.5) == "  0.1500000E+01"
    assert _fortran_e(-0.000123456789) == " -0.1234568E-03"
    assert _fortran_e(9.99999996) == "  0.1000000E+02"  # rounding carries into the exponent
    assert _fortran_e(0.0) == "  0.0000000E+00"
    assert _fortran_e(float("nan")).strip() == "NaN"
    buf = io.StringIO()
    write_report(buf, "007.forecast.", 55000.5, [3.0, 1.5, -0.2, 0.3, 2e-5])
    lines = buf.getvalue().splitlines()
    assert lines[0] == "007.forecast.mjd                           55000.5000"
    assert lines[1] == "007.forecast.rms_yo-Hx                  0.1500000E+01"
    assert lines[2] == "007.forecast.bias_yo-Hx                -0.2000000E+00"
    assert lines[4] == "007.forecast.projection_yo-Hx_into_HSf  0.2000000E-04"
