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
