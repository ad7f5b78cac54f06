"""ctypes front-end of the CPU oracle (oracle/oak_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg, as the checker.  The product path
(``oak_amd``) never imports this module.

All arrays are numpy float64 in Fortran (column-major) order, indices 0-based.
"""
import ctypes as C
import glob
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# OAK_ORACLE_LIB: another build of the same sources (the sanitised one of
# tests/test_oracle_sanitizers.py)
_LIB_PATH = os.environ.get("OAK_ORACLE_LIB") or os.path.join(_HERE, "liboak_oracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class OracleError(RuntimeError):
    pass


def build():
    """Compile the oracle with its Makefile (gcc)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE, "-B", "liboak_oracle.so"])


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        if os.environ.get("OAK_ORACLE_LIB"):
            raise OracleError("OAK_ORACLE_LIB names a missing file: %s" % _LIB_PATH)
        build()
    L = C.CDLL(_LIB_PATH)
    L.oko_distance.restype = C.c_double
    L.oko_distance.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    L.oko_locfun.restype = C.c_double
    L.oko_locfun.argtypes = [C.c_double]
    L.oko_select.restype = C.c_int64
    L.oko_symeig.restype = C.c_int
    L.oko_analysis_increment.restype = C.c_int
    L.oko_loc_analysis_csr.restype = C.c_int
    L.oko_loc_analysis_geo.restype = C.c_int
    L.oko_ensemble_analysis.restype = C.c_int
    L.oko_partition.restype = C.c_int
    L.oko_load_lapack.restype = C.c_int
    L.oko_load_lapack.argtypes = [C.c_char_p, C.c_char_p]
    L.oko_set_backend.restype = C.c_int
    L.oko_set_backend.argtypes = [C.c_int]
    _lib = L
    return L


def load_lapack():
    """Route the oracle's eigensolve/GEMMs through LAPACK dsyev and BLAS dgemm
    (the routines OAK calls, matoper_inc.F90:435,986-1005), taken from the
    OpenBLAS that scipy ships.  Returns True when LAPACK is active."""
    L = lib()
    import scipy
    root = os.path.dirname(os.path.dirname(scipy.__file__))
    for path in sorted(glob.glob(os.path.join(root, "scipy.libs", "libscipy_openblas*.so"))):
        for prefix in (b"scipy_", b""):
            if L.oko_load_lapack(path.encode(), prefix) == 0:
                return L.oko_set_backend(1) == 0
    return False


def set_backend(b):
    """0 = own Jacobi eigensolver and loop GEMMs, 1 = LAPACK/BLAS."""
    if lib().oko_set_backend(int(b)) != 0:
        raise OracleError("LAPACK backend not loaded")


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_vp)


def distance(metric, p0, p1):
    """assimilation.F90:3635-3672; p0 = observation, p1 = zone position."""
    return lib().oko_distance(int(metric), float(p0[0]), float(p0[1]), float(p1[0]), float(p1[1]))


def locfun(r):
    return lib().oko_locfun(float(r))


def select(metric, loctype, zpos, L, cutoff, ox, oy, oz=None, ot=None):
    """selectObservations for one zone (assimilation.F90:3683-3771).
    Returns (weight[m], relevant[m] bool, count)."""
    m = len(ox)
    zp = np.zeros(4)
    zp[: len(zpos)] = zpos
    ox = _f(ox); oy = _f(oy)
    oz = _f(oz if oz is not None else np.zeros(m))
    ot = _f(ot if ot is not None else np.zeros(m))
    w = np.zeros(m)
    rel = np.zeros(m, dtype=np.uint8)
    cnt = lib().oko_select(C.c_int(metric), C.c_int(loctype), _ptr(zp), C.c_double(L),
                           C.c_double(cutoff), C.c_int64(m), _ptr(ox), _ptr(oy), _ptr(oz),
                           _ptr(ot), _ptr(w), _ptr(rel))
    return w, rel.astype(bool), int(cnt)


def symeig(A):
    """dsyev('V','L') semantics: ascending eigenvalues, vectors in columns."""
    A = _f(A)
    n = A.shape[0]
    E = np.zeros(n)
    V = np.zeros((n, n), order="F")
    rc = lib().oko_symeig(C.c_int(n), _ptr(A), C.c_int(n), _ptr(E), _ptr(V))
    if rc:
        raise OracleError("symeig rc=%d" % rc)
    return E, V


def perp_space(w):
    w = _f(w)
    n = len(w)
    H = np.zeros((n, n - 1), order="F")
    lib().oko_perp_space(C.c_int(n), _ptr(w), _ptr(H))
    return H


def rotate_vector(w, v):
    w = _f(w); v = _f(v)
    n = len(w)
    Om = np.zeros((n, n), order="F")
    lib().oko_rotate_vector(C.c_int(n), _ptr(w), _ptr(v), _ptr(Om))
    return Om


def analysis_increment(Hxf, yo, Sf, HSf, sig2, w=None, dex=None):
    """analysisIncrement (rrsqrt.F90:100-190) with R = diag(sig2) tapered by w.
    Returns (xa_xf, Sa, ampl)."""
    Sf = _f(Sf); HSf = _f(HSf)
    p, N = HSf.shape
    s = Sf.shape[0]
    w = _f(np.ones(p) if w is None else w)
    sig2 = _f(sig2); Hxf = _f(Hxf); yo = _f(yo)
    dexa = None if dex is None else _f(dex)
    xa_xf = np.zeros(s)
    Sa = np.zeros((s, N), order="F")
    ampl = np.zeros(N)
    rc = lib().oko_analysis_increment(C.c_int64(p), C.c_int(N), C.c_int64(s), _ptr(Hxf), _ptr(yo),
                                      _ptr(w), _ptr(sig2), _ptr(dexa), _ptr(Sf), C.c_int64(s),
                                      _ptr(HSf), C.c_int64(p), _ptr(xa_xf), _ptr(Sa), C.c_int64(s),
                                      _ptr(ampl))
    if rc:
        raise OracleError("analysis_increment rc=%d" % rc)
    return xa_xf, Sa, ampl


def loc_analysis_csr(zone_ptr, sel_ptr, sel_idx, sel_w, xf, Hxf, yo, sig2, Sf, HSf,
                     dex=None, localise_obs=True, want_ampl=False, nthreads=0):
    """locAnalysis (rrsqrt.F90:258-466) with the selection given as CSR."""
    zone_ptr = np.ascontiguousarray(zone_ptr, dtype=np.int64)
    sel_ptr = np.ascontiguousarray(sel_ptr, dtype=np.int64)
    sel_idx = np.ascontiguousarray(sel_idx, dtype=np.int32)
    sel_w = _f(sel_w)
    Sf = _f(Sf); HSf = _f(HSf)
    n, N = Sf.shape
    m = HSf.shape[0]
    nz = len(zone_ptr) - 1
    xa = np.zeros(n)
    Sa = np.zeros((n, N), order="F")
    ampl = np.zeros((N, nz), order="F") if want_ampl else None
    dexa = None if dex is None else _f(dex)
    rc = lib().oko_loc_analysis_csr(
        C.c_int(nz), _ptr(zone_ptr), C.c_int(N), C.c_int64(m), _ptr(sel_ptr), _ptr(sel_idx),
        _ptr(sel_w), _ptr(_f(xf)), _ptr(_f(Hxf)), _ptr(_f(yo)), _ptr(_f(sig2)), _ptr(dexa),
        _ptr(Sf), C.c_int64(n), _ptr(HSf), C.c_int64(m), _ptr(xa), _ptr(Sa), C.c_int64(n),
        _ptr(ampl), C.c_int(1 if localise_obs else 0), C.c_int(nthreads))
    if rc:
        raise OracleError("loc_analysis_csr rc=%d" % rc)
    return (xa, Sa, ampl) if want_ampl else (xa, Sa)


def _geo_args(g):
    m = len(g["ox"])
    oz = g.get("oz"); ot = g.get("ot")
    return (_f(g["zpos"]), _f(g["L"]), _f(g["cutoff"]), _f(g["ox"]), _f(g["oy"]),
            _f(oz if oz is not None else np.zeros(m)), _f(ot if ot is not None else np.zeros(m)))


def loc_analysis_geo(geo, xf, Hxf, yo, sig2, Sf, HSf, dex=None, localise_obs=True,
                     zones=None, want_ampl=False, nthreads=0):
    """locAnalysis with OAK's selectObservations evaluated per zone.

    geo: dict with zone_ptr (nz+1), zpos (nz x 4: x,y,z,t of each zone's first
    element), L, cutoff (nz), metric, loctype, ox, oy[, oz, ot] (m).
    zones: optional (z0, z1) range to analyse (others keep xa = xf, Sa = Sf).
    Returns (xa, Sa[, ampl], counts)."""
    zone_ptr = np.ascontiguousarray(geo["zone_ptr"], dtype=np.int64)
    nz = len(zone_ptr) - 1
    z0, z1 = zones if zones is not None else (0, nz)
    zpos = np.ascontiguousarray(geo["zpos"], dtype=np.float64)
    _, L, cut, ox, oy, oz, ot = _geo_args(geo)
    Sf = _f(Sf); HSf = _f(HSf)
    n, N = Sf.shape
    m = HSf.shape[0]
    xa = np.zeros(n)
    Sa = np.zeros((n, N), order="F")
    ampl = np.zeros((N, nz), order="F") if want_ampl else None
    counts = np.zeros(nz, dtype=np.int64)
    dexa = None if dex is None else _f(dex)
    rc = lib().oko_loc_analysis_geo(
        C.c_int(nz), _ptr(zone_ptr), C.c_int(z0), C.c_int(z1), C.c_int(N), C.c_int64(m),
        C.c_int(geo["metric"]), C.c_int(geo["loctype"]), _ptr(zpos), _ptr(L), _ptr(cut),
        _ptr(ox), _ptr(oy), _ptr(oz), _ptr(ot), _ptr(_f(xf)), _ptr(_f(Hxf)), _ptr(_f(yo)),
        _ptr(_f(sig2)), _ptr(dexa), _ptr(Sf), C.c_int64(n), _ptr(HSf), C.c_int64(m), _ptr(xa),
        _ptr(Sa), C.c_int64(n), _ptr(ampl), C.c_int(1 if localise_obs else 0),
        C.c_int(nthreads), _ptr(counts))
    if rc:
        raise OracleError("loc_analysis_geo rc=%d" % rc)
    if want_ampl:
        return xa, Sa, ampl, counts
    return xa, Sa, counts


def obsoper(Hi, Hj, Hs, m, x):
    """COO H·x (matoper_inc.F90:220-242)."""
    Hi = np.ascontiguousarray(Hi, dtype=np.int32)
    Hj = np.ascontiguousarray(Hj, dtype=np.int32)
    Hs = _f(Hs)
    x = _f(x)
    Hx = np.zeros(m)
    lib().oko_obsoper(C.c_int64(m), C.c_int64(len(Hs)), _ptr(Hi), _ptr(Hj), _ptr(Hs), _ptr(x),
                      _ptr(Hx))
    return Hx


def ensemble_analysis(geo, H, Hshift, yo, sig2, E, dex=None, inflation=1.0,
                      max_correction=np.inf, zones=None, nthreads=0):
    """Assim's ensemble path with the local scheme (assimilation.F90:3094-3135,
    3299-3357, 3558-3563).  H = (Hi, Hj, Hs) COO, 0-based, columns in
    zone-permuted order.  max_correction: a scalar, or OAK's per-element
    maxCorrection array (n, permuted).  Returns (Ea, xa, counts)."""
    zone_ptr = np.ascontiguousarray(geo["zone_ptr"], dtype=np.int64)
    nz = len(zone_ptr) - 1
    z0, z1 = zones if zones is not None else (0, nz)
    zpos = np.ascontiguousarray(geo["zpos"], dtype=np.float64)
    _, L, cut, ox, oy, oz, ot = _geo_args(geo)
    Hi = np.ascontiguousarray(H[0], dtype=np.int32)
    Hj = np.ascontiguousarray(H[1], dtype=np.int32)
    Hs = _f(H[2])
    E = _f(E)
    n, N = E.shape
    m = len(yo)
    Ea = np.zeros((n, N), order="F")
    xa = np.zeros(n)
    counts = np.zeros(nz, dtype=np.int64)
    dexa = None if dex is None else _f(dex)
    if np.ndim(max_correction) == 0:
        mcv = None
    else:
        mcv = _f(max_correction)
        max_correction = np.inf
        if len(mcv) != n:
            raise OracleError("max_correction: one value per state element")
    rc = lib().oko_ensemble_analysis(
        C.c_int(nz), _ptr(zone_ptr), C.c_int(z0), C.c_int(z1), C.c_int(N), C.c_int64(m),
        C.c_int(geo["metric"]), C.c_int(geo["loctype"]), _ptr(zpos), _ptr(L), _ptr(cut),
        _ptr(ox), _ptr(oy), _ptr(oz), _ptr(ot), C.c_int64(len(Hs)), _ptr(Hi), _ptr(Hj), _ptr(Hs),
        _ptr(_f(Hshift)), _ptr(_f(yo)), _ptr(_f(sig2)), _ptr(dexa), _ptr(E),
        C.c_double(inflation), C.c_double(max_correction), _ptr(mcv), _ptr(Ea), _ptr(xa),
        C.c_int(nthreads),
        _ptr(counts))
    if rc:
        raise OracleError("ensemble_analysis rc=%d" % rc)
    return Ea, xa, counts


def report(d, HS, sig2, dex=None):
    """report (assimilation.F90:3584-3628) for innovations d (m), HS (m x N)
    and R = diag(sig2), optionally wrapped by the exclusion mask dex
    (0 = excluded).  Returns [ingrid, rms, bias, rms_invsqrtR, projection]."""
    d = _f(d)
    HS = np.asfortranarray(HS, dtype=np.float64)
    m, N = HS.shape
    out = np.zeros(5)
    lib().oko_report(C.c_int64(m), C.c_int(N), _ptr(d), _ptr(HS), C.c_int64(max(m, 1)),
                     _ptr(_f(sig2)), _ptr(None if dex is None else _f(dex)), _ptr(out))
    return out


def ensemble_report(H, Hshift, yo, sig2, E, dex=None, ranges=None):
    """The forecast (E = Ef) or analysis (E = Ea) report of Assim's ensemble
    path (assimilation.F90:3112-3134, 3325-3357, 3389-3391, 3424-3429).
    H = (Hi, Hj, Hs) COO, 0-based.  ranges: observation offsets [r+1]
    (default: one range of all m).  Returns an array (nranges, 5)."""
    Hi = np.ascontiguousarray(H[0], dtype=np.int32)
    Hj = np.ascontiguousarray(H[1], dtype=np.int32)
    Hs = _f(H[2])
    E = _f(E)
    n, N = E.shape
    m = len(yo)
    rp = np.ascontiguousarray([0, m] if ranges is None else ranges, dtype=np.int64)
    nr = len(rp) - 1
    out = np.zeros((nr, 5))
    rc = lib().oko_ensemble_report(
        C.c_int64(n), C.c_int(N), C.c_int64(m), C.c_int64(len(Hs)), _ptr(Hi), _ptr(Hj),
        _ptr(Hs), _ptr(_f(Hshift)), _ptr(_f(yo)), _ptr(_f(sig2)),
        _ptr(None if dex is None else _f(dex)), _ptr(E), C.c_int(nr), _ptr(rp), _ptr(out))
    if rc:
        raise OracleError("ensemble_report rc=%d" % rc)
    return out


def partition(part):
    """initPartition (assimilation.F90:578-641) after the gap relabelling of
    :408-412.  Returns (zone_size, zone_index, inv_zone_index), 0-based."""
    part = np.ascontiguousarray(part, dtype=np.int64)
    n = len(part)
    zs = np.zeros(n, dtype=np.int64)
    zi = np.zeros(n, dtype=np.int64)
    izi = np.zeros(n, dtype=np.int64)
    nz = lib().oko_partition(C.c_int64(n), _ptr(part), _ptr(zs), _ptr(zi), _ptr(izi))
    return zs[:nz].copy(), zi, izi


# --- observation operator from coordinates (oak_oracle_interp.c) ---

def _factorial(n):
    f = 1
    for i in range(2, n + 1):
        f *= i
    return f


def tetrahedra(n):
    """The hyper-tetrahedral split of the unit cube (ndgrid.F90:357-435):
    array (2^n, n+1, n! 2^(n-1)), column-major as OAK's g%tetrahedron."""
    nb = _factorial(n) * 2 ** (n - 1)
    tc = np.zeros((2 ** n, n + 1, nb), order="F")
    if lib().oko_tetrahedra(C.c_int(n), _ptr(tc)) != 0:
        raise OracleError("tetrahedra: n out of range")
    return tc


def _grid_arrays(coords):
    """coords: list of n arrays of the grid's shape (Fortran order: the first
    index fastest), the node coordinates per dimension."""
    n = len(coords)
    shape = np.ones(4, dtype=np.int64)
    shape[:n] = coords[0].shape
    flat = [np.ascontiguousarray(np.asarray(c, dtype=np.float64).ravel(order="F"))
            for c in coords]
    return n, shape, flat


def cinterp(coords, xi, mask=None):
    """cinterp (ndgrid.F90:1183-1257) at one point: (node linear indices
    (0-based, first index fastest), weights) of the 2^n cell corners, or
    None when xi is outside the grid or next to a masked node."""
    n, shape, flat = _grid_arrays(coords)
    cp = (C.c_void_p * 4)(*([a.ctypes.data for a in flat] + [None] * (4 - n)))
    mk = None if mask is None else np.ascontiguousarray(np.asarray(mask, np.uint8).ravel(order="F"))
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    cidx = np.zeros(16, dtype=np.int64)
    coeff = np.zeros(16)
    L = lib()
    L.oko_cinterp.restype = C.c_int
    r = L.oko_cinterp(C.c_int(n), _ptr(shape), cp, _ptr(mk), _ptr(xi), _ptr(cidx), _ptr(coeff))
    if r < 0:
        raise OracleError("cinterp rc=%d" % -r)
    if r == 0:
        return None
    return cidx[:r].copy(), coeff[:r].copy()


def obs_operator(grids, obs_var, obs_pos):
    """genObservationOper + packSparseMatrix (assimilation.F90:2471-2656,
    4132-4208) for one model variable per observation.

    grids: list of dicts with 'coords' (n arrays of the grid shape, Fortran
    order), optional 'mask' (True = land) and 'state' (node -> 0-based state
    column, -1 = not in the state; default: the node index).
    obs_var: [m] int (-1 = unknown variable); obs_pos: list of up to 4 arrays
    [m] (x, y, z, t).  Returns (row_ptr, col, val, valid) with H in CSR by
    observation."""
    nv = len(grids)
    ndim = np.zeros(max(nv, 1), dtype=np.int32)
    shape = np.ones((max(nv, 1), 4), dtype=np.int64)
    keep = []
    cptr = (C.c_void_p * (4 * max(nv, 1)))()
    mptr = (C.c_void_p * max(nv, 1))()
    sptr = (C.c_void_p * max(nv, 1))()
    twon_max = 1
    for v, g in enumerate(grids):
        n, sh, flat = _grid_arrays(g["coords"])
        ndim[v] = n
        shape[v] = sh
        twon_max = max(twon_max, 2 ** n)
        keep += flat
        for d in range(n):
            cptr[4 * v + d] = flat[d].ctypes.data
        if g.get("mask") is not None:
            mk = np.ascontiguousarray(np.asarray(g["mask"], np.uint8).ravel(order="F"))
            keep.append(mk)
            mptr[v] = mk.ctypes.data
        if g.get("state") is not None:
            st = np.ascontiguousarray(np.asarray(g["state"], np.int64).ravel(order="F"))
            keep.append(st)
            sptr[v] = st.ctypes.data
    obs_var = np.ascontiguousarray(obs_var, dtype=np.int32)
    m = len(obs_var)
    pos = [np.ascontiguousarray(p, dtype=np.float64) for p in obs_pos]
    pp = [C.c_void_p(p.ctypes.data) for p in pos] + [None] * (4 - len(pos))
    cap = max(1, m * twon_max)
    row_ptr = np.zeros(m + 1, dtype=np.int64)
    col = np.zeros(cap, dtype=np.int32)
    val = np.zeros(cap)
    valid = np.zeros(max(m, 1), dtype=np.uint8)
    L = lib()
    L.oko_obs_operator.restype = C.c_int
    rc = L.oko_obs_operator(C.c_int(nv), _ptr(ndim), _ptr(np.ascontiguousarray(shape.ravel())),
                            cptr, mptr, sptr, C.c_int64(m), _ptr(obs_var), pp[0], pp[1], pp[2],
                            pp[3], C.c_int64(cap), _ptr(row_ptr), _ptr(col), _ptr(val),
                            _ptr(valid))
    if rc:
        raise OracleError("obs_operator rc=%d" % rc)
    nz = int(row_ptr[-1])
    return row_ptr, col[:nz].copy(), val[:nz].copy(), valid[:m].astype(bool)
