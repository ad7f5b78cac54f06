"""Observation error covariance: the diagonal case of OAK's Covar family.

``DiagCovar`` mirrors covariance.F90:70-79 and 382-460 (R = diag(D)); the
analysis applies it as OAK's DCDCovar does (covariance.F90:612-619): with the
taper weights w of a zone, R_loc^-1 x = w * (w * x) / D.
"""
import numpy as np


class DiagCovar:
    """R = diag(D), D = rmse^2 > 0 (assimilation.F90:2120-2127)."""

    def __init__(self, D):
        self.D = np.ascontiguousarray(D, dtype=np.float64)
        self.n = self.D.size

    def full(self):
        return np.diag(self.D)

    def mtimes(self, x):
        return self.D * x

    def mldivide(self, x):
        return x / self.D

    def pack(self, mask):
        return DiagCovar(self.D[np.asarray(mask, bool)])
