"""ctypes bindings of the C ABI in include/oakg.h (oak_amd/liboakg.so).

The library is the product: every analysis runs in its HIP kernels.  Loading
fails loudly when the library is missing or when no GPU is visible; there is
no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboakg.so")

OAKG_OK, OAKG_EARG, OAKG_ENAN, OAKG_EHIP, OAKG_ENOMEM, OAKG_EUNSUP, OAKG_ECONV = range(7)
OAKG_FP64, OAKG_MIXED = 0, 1
OAKG_DEVICE_PTRS = 1
OAKG_MAX_N = 256
OAKG_NOALLOC_SCRATCH_BYTES = 256 << 20
OAKG_EIG_JACOBI, OAKG_EIG_QL = 0, 1

# Names every build of the library exports (checked by the CPU test suite).
EXPORTS = [
    "oakg_version", "oakg_create", "oakg_destroy", "oakg_last_error", "oakg_set_stream",
    "oakg_set_timing", "oakg_set_eigensolver", "oakg_get_stats", "oakg_get_counts", "oakg_local_analysis",
    "oakg_local_analysis_csr", "oakg_select_observations", "oakg_ensemble_analysis",
    "oakg_ensemble_analysis_f32", "oakg_local_nll", "oakg_report", "oakg_ensemble_report",
    "oakg_ensemble_report_f32", "oakg_coo_to_csr", "oakg_partition",
    "oakg_local_analysis_csr_noalloc", "oakg_obs_operator", "oakg_gather_rows",
    "oakg_scatter_rows", "oakg_set_tuning", "oakg_get_tuning",
]
OAKG_TUNE_SLOTS, OAKG_TUNE_QL_ZONES, OAKG_TUNE_GRAPH = 1, 2, 3
OAKG_TUNE_GRAPH_REPLAYS, OAKG_TUNE_GRAPH_CAPTURES, OAKG_TUNE_ORDER = 4, 5, 6
OAKG_TUNE_TRIDIAG_SPLIT, OAKG_TUNE_DUAL, OAKG_TUNE_PERSIST = 7, 8, 9


class OakError(RuntimeError):
    """Raised for a non-zero return code of the C ABI."""

    def __init__(self, code, msg):
        super().__init__("oakg error %d: %s" % (code, msg))
        self.code = code


class NaNError(OakError):
    """NaN in the analysis amplitudes (OAK stops with ERROR_STOP, rrsqrt.F90:145-149)."""


_vp = C.c_void_p


class Zones(C.Structure):
    _fields_ = [("nzones", C.c_int64), ("zone_ptr", _vp), ("zpos", _vp), ("corr_len", _vp),
                ("max_len", _vp), ("metric", C.c_int32), ("loctype", C.c_int32)]


class Obs(C.Structure):
    _fields_ = [("m", C.c_int64), ("x", _vp), ("y", _vp), ("z", _vp), ("t", _vp), ("yo", _vp),
                ("rmse2", _vp), ("excluded", _vp)]


class Csr(C.Structure):
    _fields_ = [("m", C.c_int64), ("n", C.c_int64), ("nnz", C.c_int64), ("row_ptr", _vp),
                ("col", _vp), ("val", _vp)]


class Grid(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("shape", C.c_int64 * 4), ("coord", _vp * 4),
                ("masked", _vp), ("state_index", _vp)]


class Stats(C.Structure):
    _fields_ = [("zones", C.c_int64), ("zones_analysed", C.c_int64),
                ("obs_selected", C.c_int64), ("pmax", C.c_int64), ("sweeps_max", C.c_int32),
                ("eig_fallbacks", C.c_int32), ("ms_total", C.c_double), ("ms_prep", C.c_double),
                ("ms_select", C.c_double), ("ms_gram", C.c_double), ("ms_eig", C.c_double),
                ("ms_tmat", C.c_double), ("ms_update", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load():
    """Load liboakg.so (no GPU needed for loading)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("oak_amd/liboakg.so is not built: run __graft_entry__.build() "
                      "(or make -C oak_amd/csrc)")
    # PyTorch ships its own HIP runtime with the same soname (libamdhip64.so.7).
    # Loading torch first makes liboakg.so bind to that copy, so one process
    # holds a single HIP runtime and torch tensors can be passed as device
    # pointers.  Without torch the system ROCm runtime is used.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    L.oakg_last_error.restype = C.c_char_p
    L.oakg_last_error.argtypes = [_vp]
    for name in EXPORTS:
        if not hasattr(L, name):
            raise OSError("liboakg.so does not export %s" % name)
    L.oakg_create.argtypes = [C.POINTER(_vp), C.c_int, C.c_int]
    L.oakg_destroy.argtypes = [_vp]
    L.oakg_set_stream.argtypes = [_vp, _vp]
    L.oakg_set_timing.argtypes = [_vp, C.c_int]
    L.oakg_set_eigensolver.argtypes = [_vp, C.c_int]
    L.oakg_get_stats.argtypes = [_vp, C.POINTER(Stats)]
    L.oakg_set_tuning.argtypes = [_vp, C.c_int, C.c_int]
    L.oakg_get_tuning.argtypes = [_vp, C.c_int, C.POINTER(C.c_int64)]
    L.oakg_get_counts.argtypes = [_vp, _vp, C.c_int64]
    L.oakg_local_analysis.argtypes = [
        _vp, C.POINTER(Zones), C.POINTER(Obs), C.c_int32, _vp, _vp, _vp, C.c_int64, _vp,
        C.c_int64, _vp, _vp, C.c_int64, _vp, C.c_int32, C.c_uint32]
    L.oakg_local_analysis_csr.argtypes = [
        _vp, C.c_int64, _vp, _vp, _vp, _vp, C.POINTER(Obs), C.c_int32, _vp, _vp, _vp,
        C.c_int64, _vp, C.c_int64, _vp, _vp, C.c_int64, _vp, C.c_int32, C.c_uint32]
    L.oakg_local_analysis_csr_noalloc.argtypes = L.oakg_local_analysis_csr.argtypes
    L.oakg_select_observations.argtypes = [
        _vp, C.POINTER(Zones), C.POINTER(Obs), C.c_int32, _vp, _vp, _vp, C.c_int64,
        C.POINTER(C.c_int64), C.c_uint32]
    L.oakg_ensemble_analysis.argtypes = [
        _vp, C.POINTER(Zones), C.POINTER(Obs), C.POINTER(Csr), _vp, C.c_int32, _vp, C.c_int64,
        _vp, C.c_int64, C.c_double, C.c_double, _vp, C.c_uint32]
    L.oakg_ensemble_analysis_f32.argtypes = L.oakg_ensemble_analysis.argtypes
    L.oakg_local_nll.argtypes = [
        _vp, C.POINTER(Zones), C.POINTER(Obs), C.c_int32, _vp, _vp, C.c_int64, C.c_double, _vp,
        _vp, _vp, _vp, _vp, C.c_uint32]
    L.oakg_report.argtypes = [
        _vp, C.c_int64, C.c_int32, _vp, _vp, C.c_int64, _vp, _vp, C.c_int32, _vp, _vp, C.c_uint32]
    L.oakg_ensemble_report.argtypes = [
        _vp, C.POINTER(Obs), C.POINTER(Csr), _vp, C.c_int32, _vp, C.c_int64, C.c_int32, _vp, _vp,
        C.c_uint32]
    L.oakg_ensemble_report_f32.argtypes = L.oakg_ensemble_report.argtypes
    L.oakg_coo_to_csr.argtypes = [C.c_int64, C.c_int64, _vp, _vp, _vp, _vp, _vp, _vp]
    L.oakg_gather_rows.argtypes = [_vp, C.c_int64, C.c_int32, _vp, C.c_int64, _vp, C.c_int64, _vp,
                                   C.c_int64, C.c_int32, C.c_uint32]
    L.oakg_scatter_rows.argtypes = L.oakg_gather_rows.argtypes
    L.oakg_obs_operator.argtypes = [
        _vp, C.c_int32, C.POINTER(Grid), C.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
        C.c_int64, _vp, C.POINTER(C.c_int64), C.c_uint32]
    L.oakg_partition.restype = C.c_int64
    L.oakg_partition.argtypes = [C.c_int64, _vp, _vp, _vp, _vp]
    _lib = L
    return L


def ptr(a):
    """Address of a numpy array or a torch tensor (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


class Context:
    """Owns an oakg_ctx (device buffers of one GPU)."""

    def __init__(self, device=0, precision=OAKG_FP64):
        L = load()
        h = _vp()
        rc = L.oakg_create(C.byref(h), int(device), int(precision))
        if rc != OAKG_OK:
            raise OakError(rc, "oakg_create failed (is a GPU visible?)")
        self._h = h
        self.device = device
        self.precision = precision

    @property
    def handle(self):
        return self._h

    def last_error(self):
        return load().oakg_last_error(self._h).decode(errors="replace")

    def check(self, rc):
        if rc == OAKG_OK:
            return
        msg = self.last_error()
        if rc == OAKG_ENAN:
            raise NaNError(rc, msg)
        raise OakError(rc, msg)

    def set_timing(self, on=True):
        self.check(load().oakg_set_timing(self._h, 1 if on else 0))

    def set_eigensolver(self, method):
        """OAKG_EIG_QL (default) or OAKG_EIG_JACOBI."""
        self.check(load().oakg_set_eigensolver(self._h, int(method)))

    def set_tuning(self, slots=None, ql_zones=None, graph=None, order=None, tri_split=None,
                   dual=None, persist=None):
        """Zone chunks in flight, zones per QL wave, HIP-graph replay of the
        zone loop, zones ordered by observation count, the N > 64 Householder
        reduction as two kernels, zones with fewer than `dual` observations
        solved in observation space, the N > 64 Gram as a persistent kernel
        (oakg_set_tuning)."""
        L = load()
        for key, v in ((OAKG_TUNE_SLOTS, slots), (OAKG_TUNE_QL_ZONES, ql_zones),
                       (OAKG_TUNE_GRAPH, graph), (OAKG_TUNE_ORDER, order),
                       (OAKG_TUNE_TRIDIAG_SPLIT, tri_split), (OAKG_TUNE_DUAL, dual),
                       (OAKG_TUNE_PERSIST, persist)):
            if v is not None:
                self.check(L.oakg_set_tuning(self._h, key, int(v)))

    def tuning(self):
        L = load()
        out = {}
        for name, key in (("slots", OAKG_TUNE_SLOTS), ("ql_zones", OAKG_TUNE_QL_ZONES),
                          ("graph", OAKG_TUNE_GRAPH), ("order", OAKG_TUNE_ORDER),
                          ("tri_split", OAKG_TUNE_TRIDIAG_SPLIT), ("dual", OAKG_TUNE_DUAL),
                          ("persist", OAKG_TUNE_PERSIST),
                          ("graph_replays", OAKG_TUNE_GRAPH_REPLAYS),
                          ("graph_captures", OAKG_TUNE_GRAPH_CAPTURES)):
            v = C.c_int64(0)
            self.check(L.oakg_get_tuning(self._h, key, C.byref(v)))
            out[name] = v.value
        return out

    def set_stream(self, stream_handle):
        self.check(load().oakg_set_stream(self._h, stream_handle))

    def stats(self):
        s = Stats()
        self.check(load().oakg_get_stats(self._h, C.byref(s)))
        return s.as_dict()

    def counts(self, nzones):
        out = np.zeros(nzones, dtype=np.int64)
        self.check(load().oakg_get_counts(self._h, out.ctypes.data, nzones))
        return out

    def close(self):
        if getattr(self, "_h", None):
            load().oakg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
