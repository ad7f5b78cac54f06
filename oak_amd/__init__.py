"""oak_amd: OAK's local ensemble square-root analysis on MI355X (gfx950).

The analysis runs in hand-written HIP kernels behind the C ABI of
include/oakg.h (oak_amd/liboakg.so).  The Python modules mirror OAK's
interfaces for that path:

* ``oak_amd.rrsqrt``        locAnalysis / analysis (rrsqrt.F90)
* ``oak_amd.assimilation``  selection geometry, oak_assim's ensemble path,
                            SparseMatrix -> CSR, initPartition (assimilation.F90)
* ``oak_amd.covariance``    DiagCovar (covariance.F90)
* ``oak_amd.synth``         synthetic inputs of the benchmark configurations
"""
from ._lib import (OAKG_DEVICE_PTRS, OAKG_FP64, OAKG_MIXED, Context, NaNError, OakError,
                   LIB_PATH, load)

__all__ = ["Context", "OakError", "NaNError", "load", "LIB_PATH", "OAKG_FP64", "OAKG_MIXED",
           "OAKG_DEVICE_PTRS"]
