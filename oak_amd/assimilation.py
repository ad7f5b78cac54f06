"""Host-side mirror of the parts of OAK's assimilation module on the local
analysis path; every computation runs on the GPU through the C ABI.

* ``Geometry``: the zone layout and localisation inputs that OAK keeps as
  module state (zoneSize/startIndexZones, the first element's coordinates,
  hCorrLengthToObs/hMaxCorrLengthToObs, metrictype, loctype;
  assimilation.F90:210-229, 400-451, 3683-3740) plus obsGridX/Y/Z/T.
* ``loc_analysis``: locAnalysis with OAK's own selectObservations
  (assimilation.F90:3683-3771) evaluated on the GPU.
* ``ensemble_analysis``: Assim's ensemble path for the local scheme, i.e.
  oak_assim(ntime, n, r, Ef, Ea) (assimilation.F90:4919-4927, 3094-3364).
* ``select_observations``: the selection alone, for inspection.
* ``coo_to_csr`` / ``partition``: SparseMatrix -> CSR and initPartition.

Arrays may be numpy (host; copied to and from the GPU) or torch tensors on
the GPU (used in place).  Device matrices are column-major: a torch tensor of
shape (rows, cols) with stride (1, ld).
"""
import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from ._lib import Csr, Obs, Zones
from .rrsqrt import default_context

CARTESIAN, SPHERICAL, SPHERICAL_APPROX = 0, 1, 2


def _is_dev(a):
    return a is not None and not isinstance(a, np.ndarray) and hasattr(a, "data_ptr")


@dataclass
class Geometry:
    zone_ptr: object          # [nz+1] int64
    zpos: object              # [nz, 4] float64: x, y, z, t of each zone's first element
    corr_len: object          # [nz] float64 (L)
    max_len: object           # [nz] float64 (cutoff)
    ox: object                # [m] obsGridX
    oy: object                # [m] obsGridY
    oz: Optional[object] = None
    ot: Optional[object] = None
    metric: int = CARTESIAN
    loctype: int = 1

    @property
    def nzones(self):
        return len(self.zone_ptr) - 1

    def device(self):
        return _is_dev(self.zone_ptr)

    def host_arrays(self):
        """Returns a copy with numpy arrays of the right dtypes and layouts."""
        f = lambda a, t: None if a is None else np.ascontiguousarray(a, dtype=t)
        return Geometry(f(self.zone_ptr, np.int64), f(self.zpos, np.float64),
                        f(self.corr_len, np.float64), f(self.max_len, np.float64),
                        f(self.ox, np.float64), f(self.oy, np.float64), f(self.oz, np.float64),
                        f(self.ot, np.float64), self.metric, self.loctype)

    def structs(self, yo, rmse2, excluded=None):
        p = _lib.ptr
        z = Zones(self.nzones, p(self.zone_ptr), p(self.zpos), p(self.corr_len),
                  p(self.max_len), int(self.metric), int(self.loctype))
        o = Obs(len(yo), p(self.ox), p(self.oy), p(self.oz), p(self.ot), p(yo), p(rmse2),
                p(excluded))
        return z, o


def _prep_geo(geo, dev):
    return geo if dev else geo.host_arrays()


def _host_vec(a, dtype=np.float64):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


def _check_dev_matrix(a, rows, cols, name):
    if tuple(a.shape) != (rows, cols):
        raise ValueError("%s must have shape (%d, %d)" % (name, rows, cols))
    if cols > 1 and a.stride(0) != 1:
        raise ValueError("%s must be column-major (stride (1, ld))" % name)
    return a.stride(1) if cols > 1 else rows


def loc_analysis(geo, xf, Hxf, yo, Sf, HSf, rmse2, excluded=None, localise_obs=True,
                 amplitudes=False, ctx=None):
    """locAnalysis with OAK's selectObservations on the GPU.

    Returns (xa, Sa) or (xa, Sa, amplitudes); with device inputs the outputs
    are new device tensors."""
    ctx = ctx or default_context()
    L = _lib.load()
    dev = _is_dev(Sf)
    g = _prep_geo(geo, dev)
    if dev:
        import torch
        n, N = Sf.shape
        m = HSf.shape[0]
        ldS = _check_dev_matrix(Sf, n, N, "Sf")
        ldH = _check_dev_matrix(HSf, m, N, "HSf") if m > 0 else 1
        xa = torch.empty(n, dtype=torch.float64, device=Sf.device)
        Sa = torch.empty((N, n), dtype=torch.float64, device=Sf.device).t()
        ampl = torch.zeros((N, g.nzones), dtype=torch.float64, device=Sf.device).t() if amplitudes else None
        z, o = g.structs(yo, rmse2, excluded)
        rc = L.oakg_local_analysis(ctx.handle, C.byref(z), C.byref(o), N, xf.data_ptr(),
                                   _lib.ptr(Hxf), Sf.data_ptr(), ldS, _lib.ptr(HSf), max(ldH, 1),
                                   xa.data_ptr(), Sa.data_ptr(), n,
                                   None if ampl is None else ampl.data_ptr(),
                                   1 if localise_obs else 0, _lib.OAKG_DEVICE_PTRS)
        ctx.check(rc)
        return (xa, Sa, ampl.t()) if amplitudes else (xa, Sa)
    Sf = np.asfortranarray(Sf, dtype=np.float64)
    HSf = np.asfortranarray(HSf, dtype=np.float64)
    n, N = Sf.shape
    m = HSf.shape[0]
    xf = _host_vec(xf); Hxf = _host_vec(Hxf); yo = _host_vec(yo); rmse2 = _host_vec(rmse2)
    ex = _host_vec(excluded, np.uint8)
    z, o = g.structs(yo, rmse2, ex)
    xa = np.zeros(n)
    Sa = np.zeros((n, N), order="F")
    ampl = np.zeros((N, g.nzones), order="F") if amplitudes else None
    rc = L.oakg_local_analysis(ctx.handle, C.byref(z), C.byref(o), N, xf.ctypes.data,
                               Hxf.ctypes.data, Sf.ctypes.data, n, HSf.ctypes.data, max(m, 1),
                               xa.ctypes.data, Sa.ctypes.data, n,
                               None if ampl is None else ampl.ctypes.data,
                               1 if localise_obs else 0, 0)
    ctx.check(rc)
    return (xa, Sa, ampl) if amplitudes else (xa, Sa)


def select_observations(geo, yo, rmse2, localise_obs=True, ctx=None):
    """selectObservations for every zone on the GPU (host arrays).
    Returns (sel_ptr, sel_idx, sel_w) with each zone's list sorted by
    observation index."""
    ctx = ctx or default_context()
    L = _lib.load()
    g = geo.host_arrays()
    yo = _host_vec(yo); rmse2 = _host_vec(rmse2)
    z, o = g.structs(yo, rmse2)
    nz = g.nzones
    sel_ptr = np.zeros(nz + 1, dtype=np.int64)
    total = C.c_int64(0)
    ctx.check(L.oakg_select_observations(ctx.handle, C.byref(z), C.byref(o),
                                         1 if localise_obs else 0, sel_ptr.ctypes.data, None,
                                         None, 0, C.byref(total), 0))
    cap = max(total.value, 1)
    idx = np.zeros(cap, dtype=np.int32)
    w = np.zeros(cap)
    ctx.check(L.oakg_select_observations(ctx.handle, C.byref(z), C.byref(o),
                                         1 if localise_obs else 0, sel_ptr.ctypes.data,
                                         idx.ctypes.data, w.ctypes.data, cap, C.byref(total), 0))
    idx = idx[: total.value]; w = w[: total.value]
    for zz in range(nz):
        a, b = sel_ptr[zz], sel_ptr[zz + 1]
        order = np.argsort(idx[a:b], kind="stable")
        idx[a:b] = idx[a:b][order]
        w[a:b] = w[a:b][order]
    return sel_ptr, idx, w


def coo_to_csr(m, rows, cols, vals):
    """Stable COO -> CSR by observation (SparseMatrix, matoper.F90:30-39)."""
    L = _lib.load()
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    vals = np.ascontiguousarray(vals, np.float64)
    nnz = len(vals)
    rp = np.zeros(m + 1, dtype=np.int64)
    co = np.zeros(max(nnz, 1), dtype=np.int32)
    va = np.zeros(max(nnz, 1))
    rc = L.oakg_coo_to_csr(m, nnz, rows.ctypes.data, cols.ctypes.data, vals.ctypes.data,
                           rp.ctypes.data, co.ctypes.data, va.ctypes.data)
    if rc != 0:
        raise _lib.OakError(rc, "coo_to_csr: row index out of range")
    return rp, co[:nnz], va[:nnz]


def obs_operator(grids, obs_var, positions, ctx=None):
    """genObservationOper + packSparseMatrix (assimilation.F90:2471-2656,
    4132-4208) on the GPU: the observation operator of observations at
    ``positions`` (up to 4 arrays [m]: x, y, z, t) of model variables
    ``obs_var`` [m] (-1: unknown) from the model grids' coordinates.

    grids: list of dicts, one per model variable, with 'coords' (ndim arrays
    of the grid shape, Fortran order: first index fastest, as OAK reads
    Model.gridX/Y/Z/T), optional 'mask' (True = land) and 'state' (node ->
    0-based column of H in the permuted state, -1 = not in the state; by
    default the node index).  Host (numpy) arrays.

    Returns (row_ptr, col, val, valid): H in CSR by observation (the
    hyper-tetrahedral weights of the 2^ndim corners of each observation's
    cell, cinterp's order) and validobs (False: OAK excludes the
    observation)."""
    ctx = ctx or default_context()
    L = _lib.load()
    keep = []
    gs = (_lib.Grid * max(len(grids), 1))()
    twon = 1
    for v, g in enumerate(grids):
        coords = g["coords"]
        n = len(coords)
        twon = max(twon, 2 ** n)
        gs[v].ndim = n
        for d in range(4):
            gs[v].shape[d] = coords[0].shape[d] if d < n else 1
        for d in range(n):
            a = np.ascontiguousarray(np.asarray(coords[d], np.float64).ravel(order="F"))
            keep.append(a)
            gs[v].coord[d] = a.ctypes.data
        if g.get("mask") is not None:
            a = np.ascontiguousarray(np.asarray(g["mask"], np.uint8).ravel(order="F"))
            keep.append(a)
            gs[v].masked = a.ctypes.data
        if g.get("state") is not None:
            a = np.ascontiguousarray(np.asarray(g["state"], np.int64).ravel(order="F"))
            keep.append(a)
            gs[v].state_index = a.ctypes.data
    obs_var = np.ascontiguousarray(obs_var, np.int32)
    m = len(obs_var)
    pos = [np.ascontiguousarray(p, np.float64) for p in positions]
    pp = [p.ctypes.data for p in pos] + [None] * (4 - len(pos))
    row_ptr = np.zeros(m + 1, np.int64)
    valid = np.zeros(max(m, 1), np.uint8)
    cap = max(1, m * twon)
    col = np.zeros(cap, np.int32)
    val = np.zeros(cap)
    nnz = C.c_int64(0)
    ctx.check(L.oakg_obs_operator(ctx.handle, len(grids), gs, m, obs_var.ctypes.data, pp[0],
                                  pp[1], pp[2], pp[3], row_ptr.ctypes.data, col.ctypes.data,
                                  val.ctypes.data, cap, valid.ctypes.data, C.byref(nnz), 0))
    k = nnz.value
    return row_ptr, col[:k], val[:k], valid[:m].astype(bool)


def partition(part):
    """initPartition (assimilation.F90:578-641) after the relabelling of
    :408-412.  Returns (zone_size, zone_index, inv_zone_index), 0-based."""
    L = _lib.load()
    part = np.ascontiguousarray(part, np.int64)
    n = len(part)
    zs = np.zeros(max(n, 1), np.int64); zi = np.zeros(max(n, 1), np.int64)
    izi = np.zeros(max(n, 1), np.int64)
    nz = L.oakg_partition(n, part.ctypes.data, zs.ctypes.data, zi.ctypes.data, izi.ctypes.data)
    if nz < 0:
        raise _lib.OakError(1, "partition failed")
    return zs[:nz], zi[:n], izi[:n]


def ensemble_analysis(geo, H, Hshift, yo, rmse2, Ef, excluded=None, inflation=1.0,
                      max_correction=np.inf, ctx=None, out=None):
    """oak_assim for the local scheme: Ef (n x N, zone-permuted) -> Ea.

    H = (row_ptr, col, val) in CSR by observation (see coo_to_csr).
    With torch device tensors, ``out`` may give the preallocated Ea.
    A float32 ensemble (numpy or torch) runs the mixed-precision path
    (oakg_ensemble_analysis_f32: FP32 state and MFMA contractions, FP64
    eigensolve); float64 runs the FP64 path.

    max_correction is OAK's maxCorrection (assimilation.F90:462-476,
    3308-3312): a scalar for every element, or one value per state element
    (n, zone-permuted; float64, a device tensor with device inputs)."""
    ctx = ctx or default_context()
    L = _lib.load()
    dev = _is_dev(Ef)
    g = _prep_geo(geo, dev)
    rp, co, va = H
    m = len(yo)
    n, N = Ef.shape
    maxc_elem = None
    if np.ndim(max_correction) == 0:
        maxc = float(max_correction) if np.isfinite(max_correction) else float("inf")
    else:
        maxc = float("inf")
        maxc_elem = max_correction
        if len(maxc_elem) != n:
            raise ValueError("max_correction: one value per state element (n = %d)" % n)
        if not dev:
            maxc_elem = np.ascontiguousarray(maxc_elem, dtype=np.float64)
        elif maxc_elem.dtype != __import__("torch").float64 or not maxc_elem.is_contiguous():
            raise TypeError("max_correction: a contiguous float64 device tensor")
    if dev:
        import torch
        ldE = _check_dev_matrix(Ef, n, N, "Ef")
        Ea = out if out is not None else torch.empty((N, n), dtype=Ef.dtype, device=Ef.device).t()
        ldA = _check_dev_matrix(Ea, n, N, "Ea")
        z, o = g.structs(yo, rmse2, excluded)
        h = Csr(m, n, len(va), rp.data_ptr(), co.data_ptr(), va.data_ptr())
        if Ef.dtype == torch.float64:
            fn = L.oakg_ensemble_analysis
        elif Ef.dtype == torch.float32:
            fn = L.oakg_ensemble_analysis_f32
            Hshift = Hshift.to(torch.float32).contiguous()
        else:
            raise TypeError("Ef must be float64 or float32")
        rc = fn(ctx.handle, C.byref(z), C.byref(o), C.byref(h), Hshift.data_ptr(), N,
                Ef.data_ptr(), ldE, Ea.data_ptr(), ldA, float(inflation), maxc,
                _lib.ptr(maxc_elem), _lib.OAKG_DEVICE_PTRS)
        ctx.check(rc)
        return Ea
    mixed = np.asarray(Ef).dtype == np.float32
    sdt = np.float32 if mixed else np.float64
    Ef = np.asfortranarray(Ef, dtype=sdt)
    yo = _host_vec(yo); rmse2 = _host_vec(rmse2); Hshift = _host_vec(Hshift, sdt)
    ex = _host_vec(excluded, np.uint8)
    rp = np.ascontiguousarray(rp, np.int64); co = np.ascontiguousarray(co, np.int32)
    va = np.ascontiguousarray(va, np.float64)
    z, o = g.structs(yo, rmse2, ex)
    h = Csr(m, n, len(va), rp.ctypes.data, co.ctypes.data, va.ctypes.data)
    Ea = np.zeros((n, N), order="F", dtype=sdt)
    fn = L.oakg_ensemble_analysis_f32 if mixed else L.oakg_ensemble_analysis
    rc = fn(ctx.handle, C.byref(z), C.byref(o), C.byref(h), Hshift.ctypes.data, N,
            Ef.ctypes.data, n, Ea.ctypes.data, n, float(inflation), maxc,
            _lib.ptr(maxc_elem), 0)
    ctx.check(rc)
    return Ea


REPORT_FIELDS = ("ingrid", "rms_yo-Hx", "bias_yo-Hx", "rms_invsqrtR_yo-Hx",
                 "projection_yo-Hx_into_HSf")


def _ranges(ranges, m):
    rp = np.ascontiguousarray([0, m] if ranges is None else ranges, dtype=np.int64)
    if rp.ndim != 1 or len(rp) < 2:
        raise ValueError("ranges must hold at least two offsets")
    return rp


def report(yo_Hx, HS, rmse2, excluded=None, ranges=None, ctx=None):
    """Statistics of OAK's report (assimilation.F90:3584-3628) on the GPU.

    yo_Hx (m), HS (m x N), rmse2 = the diagonal of R, excluded = OAK's
    exclude_obs (R is then wrapped as DCDCovar, assimilation.F90:3086-3092).
    ranges: observation offsets [nranges+1] (the per-variable reports of
    :3410-3430); default one range of all m.  Returns an array (nranges, 5)
    with the columns of REPORT_FIELDS."""
    ctx = ctx or default_context()
    L = _lib.load()
    m, N = HS.shape
    if _is_dev(HS):
        import torch
        ld = _check_dev_matrix(HS, m, N, "HS") if m > 0 else 1
        rp = torch.as_tensor(_ranges(ranges, m), device=HS.device)
        out = torch.empty((len(rp) - 1, 5), dtype=torch.float64, device=HS.device)
        rc = L.oakg_report(ctx.handle, m, N, _lib.ptr(yo_Hx), HS.data_ptr(), max(ld, 1),
                           _lib.ptr(rmse2), _lib.ptr(excluded), len(rp) - 1, rp.data_ptr(),
                           out.data_ptr(), _lib.OAKG_DEVICE_PTRS)
        ctx.check(rc)
        return out
    HS = np.asfortranarray(HS, dtype=np.float64)
    d = _host_vec(yo_Hx); r2 = _host_vec(rmse2); ex = _host_vec(excluded, np.uint8)
    rp = _ranges(ranges, m)
    out = np.zeros((len(rp) - 1, 5))
    rc = L.oakg_report(ctx.handle, m, N, _lib.ptr(d), HS.ctypes.data, max(m, 1), _lib.ptr(r2),
                       _lib.ptr(ex), len(rp) - 1, rp.ctypes.data, out.ctypes.data, 0)
    ctx.check(rc)
    return out


def ensemble_report(H, Hshift, yo, rmse2, E, excluded=None, ranges=None, ctx=None):
    """The forecast (E = Ef) or analysis (E = Ea) report of Assim's ensemble
    path (assimilation.F90:3389-3391, 3424-3429): HE = H E + Hshift,
    Hx = mean(HE), HS = (HE - Hx)/sqrt(N-1), d = yo - Hx.  H = (row_ptr, col,
    val) as for ensemble_analysis; a float32 ensemble uses the FP32 entry
    point.  Returns an array (nranges, 5) with the columns of REPORT_FIELDS."""
    ctx = ctx or default_context()
    L = _lib.load()
    rp_, co, va = H
    m = len(yo)
    n, N = E.shape
    if _is_dev(E):
        import torch
        ldE = _check_dev_matrix(E, n, N, "E")
        o = Obs(m, None, None, None, None, _lib.ptr(yo), _lib.ptr(rmse2), _lib.ptr(excluded))
        h = Csr(m, n, len(va), rp_.data_ptr(), co.data_ptr(), va.data_ptr())
        if E.dtype == torch.float64:
            fn = L.oakg_ensemble_report
        elif E.dtype == torch.float32:
            fn = L.oakg_ensemble_report_f32
            Hshift = Hshift.to(torch.float32).contiguous()
        else:
            raise TypeError("E must be float64 or float32")
        rr = torch.as_tensor(_ranges(ranges, m), device=E.device)
        out = torch.empty((len(rr) - 1, 5), dtype=torch.float64, device=E.device)
        rc = fn(ctx.handle, C.byref(o), C.byref(h), Hshift.data_ptr(), N, E.data_ptr(), ldE,
                len(rr) - 1, rr.data_ptr(), out.data_ptr(), _lib.OAKG_DEVICE_PTRS)
        ctx.check(rc)
        return out
    mixed = np.asarray(E).dtype == np.float32
    sdt = np.float32 if mixed else np.float64
    E = np.asfortranarray(E, dtype=sdt)
    yo = _host_vec(yo); rmse2 = _host_vec(rmse2); Hshift = _host_vec(Hshift, sdt)
    ex = _host_vec(excluded, np.uint8)
    rp_ = np.ascontiguousarray(rp_, np.int64); co = np.ascontiguousarray(co, np.int32)
    va = np.ascontiguousarray(va, np.float64)
    o = Obs(m, None, None, None, None, yo.ctypes.data, rmse2.ctypes.data, _lib.ptr(ex))
    h = Csr(m, n, len(va), rp_.ctypes.data, co.ctypes.data, va.ctypes.data)
    rr = _ranges(ranges, m)
    out = np.zeros((len(rr) - 1, 5))
    fn = L.oakg_ensemble_report_f32 if mixed else L.oakg_ensemble_report
    rc = fn(ctx.handle, C.byref(o), C.byref(h), Hshift.ctypes.data, N, E.ctypes.data, n,
            len(rr) - 1, rr.ctypes.data, out.ctypes.data, 0)
    ctx.check(rc)
    return out


def _fortran_e(x, w=15, d=7):
    """Fortran's Ew.d edit descriptor (0.ddddddd E+xx), right-justified."""
    if not np.isfinite(x):
        return ("NaN" if np.isnan(x) else ("-Infinity" if x < 0 else "Infinity")).rjust(w)
    if x == 0:
        return ("0." + "0" * d + "E+00").rjust(w)
    mant, ex = ("%.*e" % (d - 1, abs(x))).split("e")
    body = "0." + mant.replace(".", "") + "E%+03d" % (int(ex) + 1)
    return (("-" if x < 0 else "") + body).rjust(w)


def write_report(unit, prefix, mjd, stats):
    """The log lines of report (assimilation.F90:3606-3620) for one row of
    report()/ensemble_report() output, written to the text stream ``unit``."""
    unit.write("%smjd                      %15.4f\n" % (prefix, mjd))
    names = ("rms_yo-Hx                ", "bias_yo-Hx               ",
             "rms_invsqrtR_yo-Hx       ", "projection_yo-Hx_into_HSf")
    for name, v in zip(names, list(stats)[1:5]):
        unit.write("%s%s%s\n" % (prefix, name, _fortran_e(float(v))))
