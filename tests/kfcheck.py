"""Dense textbook Kalman-filter checks and the deterministic inputs of OAK's
test/test_rrsqrt.F90, written independently of the oracle's ensemble-space
formulas (numpy FP64 inverses, no eigendecomposition)."""
import numpy as np


def rrsqrt_inputs():
    """test/test_rrsqrt.F90:283-315: n=10, m=5, N=12, Ef = sin(3 i^2),
    H = reshape(1..50), y = 1..5, R = 2 I."""
    m, n, N = 5, 10, 12
    i = np.arange(1, n * N + 1, dtype=np.float64)
    Ef = np.sin(3.0 * i * i).reshape((n, N), order="F")
    H = np.arange(1, m * n + 1, dtype=np.float64).reshape((m, n), order="F")
    y = np.arange(1, m + 1, dtype=np.float64)
    xf = Ef.sum(axis=1) / N
    Sf = (Ef - xf[:, None]) / np.sqrt(N - 1.0)
    sig2 = np.full(m, 2.0)
    return dict(m=m, n=n, N=N, Ef=Ef, H=H, y=y, xf=xf, Sf=Sf, sig2=sig2)


def global_kf(xf, Sf, H, yo, R):
    """test/test_rrsqrt.F90:47-56: K = Pf H^T (H Pf H^T + R)^-1."""
    Pf = Sf @ Sf.T
    K = Pf @ H.T @ np.linalg.inv(H @ Pf @ H.T + R)
    Pa = Pf - K @ H @ Pf
    xa = xf + K @ (yo - H @ xf)
    return xa, Pa


def local_kf_rows(xf, Sf, H, yo, sig2, rows, iloc, wloc):
    """Zone check of test/test_rrsqrt.F90:193-226, as a Kalman gain: the zone
    rows of xa and Pa with R_loc^-1 = R^-1 o (w w^T) on the selected obs.
    Uses the gain form (no inverse of Pf) so that rank-deficient ensembles
    work too."""
    Pf = Sf @ Sf.T
    Hl = H[iloc, :]
    Rinv = np.diag(1.0 / sig2[iloc]) * np.outer(wloc, wloc)
    # (Hl Pf Hl^T + R_loc)^-1 = R_loc^-1 (I + Hl Pf Hl^T R_loc^-1)^-1, valid when
    # some weights vanish (R_loc^-1 singular).
    p = len(iloc)
    G = np.linalg.inv(np.eye(p) + Hl @ Pf @ Hl.T @ Rinv)
    K = Pf @ Hl.T @ Rinv @ G
    xa = xf + K @ (yo[iloc] - Hl @ xf)
    Pa = Pf - K @ Hl @ Pf
    return xa[rows], Pa[np.ix_(rows, rows)]


def rel_err(a, b):
    a = np.asarray(a); b = np.asarray(b)
    d = np.max(np.abs(a - b)) if a.size else 0.0
    s = np.max(np.abs(b)) if b.size else 0.0
    return d / s if s > 0 else d
