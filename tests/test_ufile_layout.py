"""GHER file format and the state layout on the CPU: the byte layout of a
file as ufileformat.F90 writes it (uwritc_all_kind :1321-1374: ten empty
records, the header record, value records of nbmots words), reading it back,
the implicit (degenerate) coordinate form, little-endian files; OAK's memory
layout (MemoryLayout_byfilenames, assimilation.F90:1620-1748) against a
brute-force restatement; the zone layout and the permuted indices."""
import struct

import numpy as np
import pytest

from oak_amd import layout as Lay
from oak_amd import ufile as U


def _rec(b):
    n = struct.pack(">i", len(b))
    return n + b + n


def test_byte_layout_known_answer(tmp_path):
    """A 2 x 3 field in real*4, nbmots = 4: ten empty records, the header
    (2, 3, 1, 4, 4, valex), a record of 4 values and one of the remaining 2,
    each framed by its big-endian byte length."""
    c = np.arange(1.0, 7.0).reshape((2, 3), order="F")
    p = tmp_path / "f.u"
    U.uwrite(str(p), c, valex=9999.0, prec=4, nbmots=4)
    expect = b"".join([_rec(b"")] * 10)
    expect += _rec(struct.pack(">5i", 2, 3, 1, 4, 4) + struct.pack(">f", 9999.0))
    expect += _rec(struct.pack(">4f", 1, 2, 3, 4)) + _rec(struct.pack(">2f", 5, 6))
    assert p.read_bytes() == expect
    a, vx = U.uread(str(p))
    assert vx == 9999.0 and a.shape == (2, 3) and np.array_equal(a, c)


@pytest.mark.parametrize("prec,nbmots,shape,order", [(8, -1, (7, 5, 3), ">"), (4, 10, (7, 5), ">"),
                                                     (8, 33, (64,), "<"), (8, 7, (4, 3, 2), ">")])
def test_roundtrip(tmp_path, prec, nbmots, shape, order):
    g = np.random.default_rng(1)
    c = g.standard_normal(shape)
    c.ravel()[::5] = 9999.0
    p = tmp_path / "r.u"
    U.uwrite(str(p), c, valex=9999.0, prec=prec, nbmots=nbmots, byteorder=order)
    a, vx = U.uread(str(p))
    ref = c.astype(np.float32).astype(np.float64) if prec == 4 else c
    assert a.shape == c.shape and np.array_equal(a, ref)
    shp, v, ip, nb, degen = U.uinquire(str(p))
    assert ip == prec and not degen and v == 9999.0
    assert shp == tuple(list(shape) + [1] * (3 - len(shape)))


def test_degenerate_implicit_coordinates(tmp_path):
    """C(i, j, k) = c0 + i dci + j dcj + k dck (1-based), negative shape in
    the header (ufileformat.F90:1195-1200, 1607-1634)."""
    p = tmp_path / "x.u"
    U.uwrite_degenerate(str(p), (5, 4), 10.0, 0.5, -2.0, 0.0)
    a, _ = U.uread(str(p))
    i, j = np.meshgrid(np.arange(1, 6), np.arange(1, 5), indexing="ij")
    assert a.shape == (5, 4) and np.array_equal(a, 10.0 + 0.5 * i - 2.0 * j)


def test_bad_files(tmp_path):
    p = tmp_path / "bad.u"
    p.write_bytes(b"\x00" * 37)
    with pytest.raises(U.UFileError):
        U.uread(str(p))


def test_mem_layout_against_brute_force():
    g = np.random.default_rng(2)
    masks = [np.where(g.uniform(size=(6, 5, 3)) < 0.2, 9999.0, 1.0),
             np.where(g.uniform(size=(6, 5)) < 0.3, 9999.0, 0.0)]
    lay = Lay.mem_layout([(m, 9999.0) for m in masks])
    assert list(lay.start) == [0, 90] and list(lay.size) == [90, 30]
    # brute force: walk the variables, number the sea nodes in order
    sea, inv, k = [], [], 0
    for v, m in enumerate(masks):
        for t, x in enumerate(m.ravel(order="F")):
            if x != 9999.0:
                sea.append(k); inv.append(lay.start[v] + t); k += 1
            else:
                sea.append(-1)
    assert np.array_equal(lay.sea_index, sea) and np.array_equal(lay.invindex, inv)
    assert list(lay.size_sea) == [int((masks[0] != 9999).sum()), int((masks[1] != 9999).sum())]
    assert lay.start_sea[1] == lay.size_sea[0] and lay.effsize == sum(lay.size_sea)


def test_zone_layout_and_indices():
    """Column zones: label = horizontal position; permuted rows are zone
    after zone, each zone in state order (stable); gather_index and
    state_index_map agree with the permutation."""
    g = np.random.default_rng(3)
    nx, ny, nz = 5, 4, 3
    masks = [g.uniform(size=(nx, ny, nz)) > 0.2, g.uniform(size=(nx, ny)) > 0.1]
    lay = Lay.mem_layout(masks)
    col = []
    for v, m in enumerate(masks):
        idx = np.nonzero(m.ravel(order="F"))[0]
        ij = idx % (nx * ny)
        col.append(ij)
    labels = np.concatenate(col)
    zones = Lay.zone_layout(labels)
    zp = zones.zone_ptr
    for z in range(len(zones.zone_size)):
        rows = zones.zone_index[zp[z]:zp[z + 1]]
        assert np.all(np.diff(rows) > 0) and np.all(labels[rows] == labels[rows[0]])
    gi = Lay.gather_index(lay, zones)
    assert np.array_equal(gi, lay.invindex[zones.zone_index])
    for v in range(2):
        smap = Lay.state_index_map(lay, zones, v).ravel(order="F")
        for t, s in enumerate(smap):
            full = lay.start[v] + t
            if lay.mask[full]:
                assert gi[s] == full
            else:
                assert s == -1
