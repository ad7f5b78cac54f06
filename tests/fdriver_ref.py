"""Reads the file written by oak_amd/fortran/oakg_fdriver and recomputes its
analysis with the oracle (same 1-D geometry and Gaussian taper as the
driver's selectObservations callback)."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oak_amd", "fortran", "oakg_fdriver")


def read(path):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:40], dtype=np.int64)
    info, nzone, zsize, N, m = (int(v) for v in hdr)
    n = nzone * zsize
    v = np.frombuffer(raw[40:], dtype=np.float64)
    sizes = [("xf", n), ("Hxf", m), ("yo", m), ("Sf", n * N), ("HSf", m * N), ("Rd", m),
             ("xa", n), ("Sa", n * N)]
    out, o = {}, 0
    for k, s in sizes:
        out[k] = v[o:o + s]
        o += s
    for k, r in (("Sf", n), ("HSf", m), ("Sa", n)):
        out[k] = out[k].reshape((r, N), order="F")
    out.update(info=info, nzone=nzone, zsize=zsize, N=N, m=m, n=n)
    return out


def oracle_analysis(d):
    from oracle import oracle as O
    nz, m = d["nzone"], d["m"]
    sel_ptr, sel_idx, sel_w = [0], [], []
    for i in range(1, nz + 1):
        dist = np.abs(i - 1.25 * np.arange(1, m + 1))
        rel = dist <= 5.0
        sel_idx += list(np.nonzero(rel)[0])
        sel_w += list(np.exp(-(dist[rel] / 2.0) ** 2))
        sel_ptr.append(len(sel_idx))
    zone_ptr = np.arange(nz + 1) * d["zsize"]
    return O.loc_analysis_csr(zone_ptr, np.array(sel_ptr), np.array(sel_idx), np.array(sel_w),
                              d["xf"], d["Hxf"], d["yo"], d["Rd"], d["Sf"], d["HSf"])
