"""GHER binary files (OAK's ``.u`` format), read and written as OAK's
ufileformat does (ufileformat.F90:1164-1246 ureadc_all_kind, 1252-1315
read_stream, 1321-1374 uwritc_all_kind).

A file is a sequence of Fortran unformatted sequential records, each framed
by its byte length as a 4-byte integer before and after (big-endian, the
byte order GHER files are exchanged in; little-endian files are recognised
by their framing):

  * 10 empty records (KBLANC, left to the user);
  * the header: imax, jmax, kmax, iprec, nbmots (int32) and valex (real*4);
  * the values in records of nbmots words (iprec = 4: real*4, 8: real*8),
    the last record holding the remaining imax*jmax*kmax mod nbmots words.

A "degenerate" file (any of imax, jmax, kmax negative) holds four values
c0, dci, dcj, dck of a field defined implicitly by
C(i, j, k) = c0 + i dci + j dcj + k dck (1-based i, j, k), the form OAK uses
for regular coordinates; ``uread`` returns it expanded to |imax| x |jmax| x
|kmax| (uload_degenerated, ufileformat.F90:1607-1634).

Arrays are returned in Fortran order (i fastest), shape (imax, jmax, kmax)
with trailing dimensions of 1 dropped, in float64 (as OAK reads them into
its default real in a PRECISION=double build); valex marks missing values
(land)."""
import struct

import numpy as np

KBLANC = 10


class UFileError(ValueError):
    pass


def _records(raw):
    """Splits the bytes into Fortran records; returns (records, byte order)."""
    for order in (">", "<"):
        recs, pos, ok = [], 0, True
        while pos < len(raw):
            if pos + 4 > len(raw):
                ok = False
                break
            (n,) = struct.unpack(order + "i", raw[pos:pos + 4])
            if n < 0 or pos + 8 + n > len(raw):
                ok = False
                break
            (n2,) = struct.unpack(order + "i", raw[pos + 4 + n:pos + 8 + n])
            if n2 != n:
                ok = False
                break
            recs.append(raw[pos + 4:pos + 4 + n])
            pos += 8 + n
        if ok and len(recs) > KBLANC:
            return recs, order
    raise UFileError("not a GHER file (Fortran record framing not recognised)")


def uinquire(path):
    """Header of a GHER file: (shape, valex, iprec, nbmots, degenerate)."""
    recs, order = _records(open(path, "rb").read())
    imax, jmax, kmax, iprec, nbmots = struct.unpack(order + "5i", recs[KBLANC][:20])
    (valex,) = struct.unpack(order + "f", recs[KBLANC][20:24])
    degen = imax < 0 or jmax < 0 or kmax < 0
    shape = tuple(abs(v) for v in (imax, jmax, kmax))
    return shape, float(valex), iprec, nbmots, degen


def uread(path):
    """Returns (array, valex): the field in Fortran order, float64."""
    recs, order = _records(open(path, "rb").read())
    imax, jmax, kmax, iprec, nbmots = struct.unpack(order + "5i", recs[KBLANC][:20])
    (valex,) = struct.unpack(order + "f", recs[KBLANC][20:24])
    if iprec not in (4, 8):
        raise UFileError("iprec must be 4 or 8, not %d" % iprec)
    dt = np.dtype(order + ("f4" if iprec == 4 else "f8"))
    data = np.concatenate([np.frombuffer(r, dtype=dt) for r in recs[KBLANC + 1:]]) \
        if len(recs) > KBLANC + 1 else np.zeros(0, dt)
    degen = imax < 0 or jmax < 0 or kmax < 0
    shape = (abs(imax), abs(jmax), abs(kmax))
    if degen:
        if data.size < 4:
            raise UFileError("degenerate file with fewer than four values")
        c0, di, dj, dk = (float(v) for v in data[:4])
        i, j, k = np.meshgrid(np.arange(1, shape[0] + 1), np.arange(1, shape[1] + 1),
                              np.arange(1, shape[2] + 1), indexing="ij")
        out = c0 + i * di + j * dj + k * dk
    else:
        n = shape[0] * shape[1] * shape[2]
        nl, ir = divmod(n, nbmots) if nbmots > 0 else (0, n)
        if data.size != n:
            raise UFileError("%d values for a %s field" % (data.size, shape))
        sizes = [len(r) // dt.itemsize for r in recs[KBLANC + 1:]]
        if any(s != nbmots for s in sizes[:nl]) or (len(sizes) != nl + 1):
            raise UFileError("records do not follow nbmots = %d" % nbmots)
        out = data.astype(np.float64).reshape(shape, order="F")
    while out.ndim > 1 and out.shape[-1] == 1:
        out = out.reshape(out.shape[:-1], order="F")
    return out, float(valex)


def _record(order, payload):
    n = struct.pack(order + "i", len(payload))
    return n + payload + n


def uwrite(path, c, valex=9999.0, prec=8, nbmots=-1, byteorder=">"):
    """Writes the field c (up to 3 dimensions, Fortran order) as OAK's
    uwritc does: nbmots = -1 puts all values in one record."""
    c = np.asarray(c, dtype=np.float64)
    if c.ndim > 3:
        raise UFileError("at most 3 dimensions")
    shape = list(c.shape) + [1] * (3 - c.ndim)
    n = int(np.prod(shape))
    nbm = n if nbmots == -1 else int(nbmots)
    dt = np.dtype(byteorder + ("f4" if prec == 4 else "f8"))
    flat = c.ravel(order="F").astype(dt)
    out = [_record(byteorder, b"") for _ in range(KBLANC)]
    out.append(_record(byteorder, struct.pack(byteorder + "5i", shape[0], shape[1], shape[2],
                                              prec, nbm) +
                       struct.pack(byteorder + "f", valex)))
    nl, ir = divmod(n, nbm) if nbm > 0 else (0, n)
    for k in range(nl):
        out.append(_record(byteorder, flat[k * nbm:(k + 1) * nbm].tobytes()))
    out.append(_record(byteorder, flat[nl * nbm:].tobytes()))
    with open(path, "wb") as f:
        f.write(b"".join(out))


def uwrite_degenerate(path, shape, c0, dci, dcj, dck, valex=9999.0, prec=8, byteorder=">"):
    """The implicit form C(i, j, k) = c0 + i dci + j dcj + k dck (negative
    dimensions in the header, four values)."""
    shape = list(shape) + [1] * (3 - len(shape))
    dt = np.dtype(byteorder + ("f4" if prec == 4 else "f8"))
    out = [_record(byteorder, b"") for _ in range(KBLANC)]
    out.append(_record(byteorder, struct.pack(byteorder + "5i", -shape[0], -shape[1], -shape[2],
                                              prec, 4) + struct.pack(byteorder + "f", valex)))
    out.append(_record(byteorder, np.array([c0, dci, dcj, dck], dtype=dt).tobytes()))
    with open(path, "wb") as f:
        f.write(b"".join(out))
