"""Host-side mirror of OAK's rrsqrt module for the local ensemble square-root
analysis; every call runs on the GPU through the C ABI (include/oakg.h).

``loc_analysis`` has the argument list of locAnalysis (rrsqrt.F90:433-457):
zone sizes, a ``select_observations`` callback, xf, Hxf, yo, Sf, HSf, R and the
optional amplitudes / localise_obs.  The callback is OAK's
``selectObservations(i, c, relevantObs)`` (rrsqrt.F90:442-448) written as a
Python function ``select_observations(i) -> (weight, relevant)`` where ``i`` is
the 0-based index of the zone's first state element (OAK passes the 1-based
``startIndex(zi)``).  It is evaluated per zone on the host, as OAK evaluates
it, and the resulting lists go to the GPU in CSR form.

``analysis`` is the global analysis (rrsqrt.F90:196-208): one zone holding the
whole state and every observation with weight 1.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import Context, Obs
from .covariance import DiagCovar

_ctx_cache = {}


def default_context(device=0, precision=_lib.OAKG_FP64):
    key = (device, precision)
    if key not in _ctx_cache:
        _ctx_cache[key] = Context(device, precision)
    return _ctx_cache[key]


def _f64(a):
    return np.asfortranarray(a, dtype=np.float64)


def _rdiag(R):
    if isinstance(R, DiagCovar):
        return R.D
    D = np.asarray(R, dtype=np.float64)
    if D.ndim == 2:
        if np.count_nonzero(D - np.diag(np.diag(D))):
            raise ValueError("only diagonal observation error covariances are supported")
        D = np.diag(D).copy()
    return np.ascontiguousarray(D)


def zone_ptr_from_sizes(zone_size):
    zs = np.asarray(zone_size, dtype=np.int64)
    zp = np.zeros(len(zs) + 1, dtype=np.int64)
    np.cumsum(zs, out=zp[1:])
    return zp


def selection_csr(zone_ptr, select_observations, m, localise_obs=True):
    """Evaluates the callback for every zone (rrsqrt.F90:368-404) and packs the
    relevant observations into CSR (sel_ptr, sel_idx, sel_w)."""
    nz = len(zone_ptr) - 1
    ptrs = np.zeros(nz + 1, dtype=np.int64)
    idx = []
    wts = []
    for z in range(nz):
        w, rel = select_observations(int(zone_ptr[z]))
        w = np.asarray(w, dtype=np.float64)
        rel = np.asarray(rel, dtype=bool)
        if rel.any():
            sel = np.nonzero(rel)[0] if localise_obs else np.arange(m)
            idx.append(sel.astype(np.int32))
            wts.append(w[sel])
            ptrs[z + 1] = ptrs[z] + len(sel)
        else:
            ptrs[z + 1] = ptrs[z]
    sel_idx = np.concatenate(idx) if idx else np.zeros(0, np.int32)
    sel_w = np.concatenate(wts) if wts else np.zeros(0)
    return ptrs, np.ascontiguousarray(sel_idx, np.int32), np.ascontiguousarray(sel_w)


def loc_analysis_csr(zone_ptr, sel_ptr, sel_idx, sel_w, xf, Hxf, yo, Sf, HSf, R,
                     amplitudes=False, localise_obs=True, excluded=None, ctx=None):
    """locAnalysis with the selection already in CSR form (host arrays)."""
    ctx = ctx or default_context()
    L = _lib.load()
    Sf = _f64(Sf); HSf = _f64(HSf)
    n, N = Sf.shape
    m = HSf.shape[0]
    xf = np.ascontiguousarray(xf, np.float64); Hxf = np.ascontiguousarray(Hxf, np.float64)
    yo = np.ascontiguousarray(yo, np.float64)
    rm = _rdiag(R)
    zone_ptr = np.ascontiguousarray(zone_ptr, np.int64)
    sel_ptr = np.ascontiguousarray(sel_ptr, np.int64)
    sel_idx = np.ascontiguousarray(sel_idx, np.int32)
    sel_w = np.ascontiguousarray(sel_w, np.float64)
    ex = None if excluded is None else np.ascontiguousarray(excluded, np.uint8)
    obs = Obs(m, None, None, None, None, yo.ctypes.data, rm.ctypes.data,
              None if ex is None else ex.ctypes.data)
    nz = len(zone_ptr) - 1
    xa = np.zeros(n)
    Sa = np.zeros((n, N), order="F")
    ampl = np.zeros((N, nz), order="F") if amplitudes else None
    rc = L.oakg_local_analysis_csr(
        ctx.handle, nz, zone_ptr.ctypes.data, sel_ptr.ctypes.data, sel_idx.ctypes.data,
        sel_w.ctypes.data, C.byref(obs), N, xf.ctypes.data, Hxf.ctypes.data, Sf.ctypes.data, n,
        HSf.ctypes.data, max(m, 1), xa.ctypes.data, Sa.ctypes.data, n,
        None if ampl is None else ampl.ctypes.data, 1 if localise_obs else 0, 0)
    ctx.check(rc)
    return (xa, Sa, ampl) if amplitudes else (xa, Sa)


def loc_analysis(zone_size, select_observations, xf, Hxf, yo, Sf, HSf, R, amplitudes=False,
                 localise_obs=True, ctx=None):
    """locAnalysis (rrsqrt.F90:433-466): returns (xa, Sa) or (xa, Sa, amplitudes).

    Zones whose callback marks no observation as relevant keep xa = xf and
    Sa = Sf (rrsqrt.F90:322-326, 371).  Raises NaNError where OAK stops on a
    NaN amplitude (rrsqrt.F90:145-149)."""
    m = len(yo)
    zone_ptr = zone_ptr_from_sizes(zone_size)
    sel_ptr, sel_idx, sel_w = selection_csr(zone_ptr, select_observations, m, localise_obs)
    return loc_analysis_csr(zone_ptr, sel_ptr, sel_idx, sel_w, xf, Hxf, yo, Sf, HSf, R,
                            amplitudes=amplitudes, localise_obs=localise_obs, ctx=ctx)


def analysis(xf, Hxf, yo, Sf, HSf, R, ctx=None):
    """Global analysis (rrsqrt.F90:196-208): returns (xa, Sa)."""
    n = len(xf)
    m = len(yo)
    zone_ptr = np.array([0, n], dtype=np.int64)
    sel_ptr = np.array([0, m], dtype=np.int64)
    return loc_analysis_csr(zone_ptr, sel_ptr, np.arange(m, dtype=np.int32), np.ones(m), xf, Hxf,
                            yo, Sf, HSf, R, ctx=ctx)
