"""Reads the file written by oak_amd/fortran/oakg_fdriver and recomputes its
analyses with the oracle: the same 1-D geometry and Gaussian taper as the
driver's selectObservations callback, which OAK calls with the state index of
each zone's first element (rrsqrt.F90:358-368), for localise_obs = .true. and
.false. (rrsqrt.F90:373-385: all observations once one is relevant)."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oak_amd", "fortran", "oakg_fdriver")


def read(path):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:56], dtype=np.int64)
    info, nzone, zsize, N, m, info2, info3 = (int(v) for v in hdr)
    n = nzone * zsize
    v = np.frombuffer(raw[56:], dtype=np.float64)
    sizes = [("xf", n), ("Hxf", m), ("yo", m), ("Sf", n * N), ("HSf", m * N), ("Rd", m),
             ("xa", n), ("Sa", n * N), ("xa2", n), ("Sa2", n * N), ("xa3", n), ("Sa3", n * N)]
    out, o = {}, 0
    for k, s in sizes:
        out[k] = v[o:o + s]
        o += s
    for k, r in (("Sf", n), ("HSf", m), ("Sa", n), ("Sa2", n), ("Sa3", n)):
        out[k] = out[k].reshape((r, N), order="F")
    out.update(info=info, info2=info2, info3=info3, nzone=nzone, zsize=zsize, N=N, m=m, n=n)
    return out


def select_gauss(ind, m):
    """The driver's callback for the 1-based state index ind."""
    d = np.abs(0.5 * ind + 3.0 - 1.25 * np.arange(1, m + 1))
    return np.exp(-(d / 2.0) ** 2), d <= 2.0


def oracle_analysis(d, localise_obs=True, zone_number=False):
    """zone_number: call the callback with the 1-based zone number instead
    (not OAK's contract; used to show that the test tells them apart)."""
    from oracle import oracle as O
    nz, m, zs = d["nzone"], d["m"], d["zsize"]
    sel_ptr, sel_idx, sel_w = [0], [], []
    for i in range(nz):
        w, rel = select_gauss(i + 1 if zone_number else i * zs + 1, m)
        if not localise_obs and rel.any():
            rel = np.ones(m, dtype=bool)
        sel_idx += list(np.nonzero(rel)[0])
        sel_w += list(w[rel])
        sel_ptr.append(len(sel_idx))
    zone_ptr = np.arange(nz + 1) * zs
    return O.loc_analysis_csr(zone_ptr, np.array(sel_ptr), np.array(sel_idx, dtype=np.int32),
                              np.array(sel_w), d["xf"], d["Hxf"], d["yo"], d["Rd"], d["Sf"],
                              d["HSf"], localise_obs=localise_obs)
