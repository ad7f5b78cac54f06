"""Error measures of an analysed ensemble against a reference analysis.

The mixed-precision path is judged on the analysis mean and on the analysis
spread (anomalies), each relative to the size of what the analysis changes
or produces:

  mean_err   = max |xa - xa_ref| / max |xa_ref - xf|
  spread_err = max |Aa - Aa_ref| / max |Aa_ref|,   Aa = Ea - mean(Ea)

with xf the forecast mean.  Both are computed in float64 (on the device for
torch tensors, column blocks at a time).
"""
import numpy as np


def _errors_np(Ea, Ea_ref, Ef):
    Ea, Ea_ref, Ef = (np.asarray(a, dtype=np.float64) for a in (Ea, Ea_ref, Ef))
    xf = Ef.mean(axis=1)
    xa = Ea.mean(axis=1)
    xr = Ea_ref.mean(axis=1)
    inc = np.abs(xr - xf).max()
    da = np.abs((Ea - xa[:, None]) - (Ea_ref - xr[:, None])).max()
    ar = np.abs(Ea_ref - xr[:, None]).max()
    return inc, np.abs(xa - xr).max(), da, ar


def _errors_torch(Ea, Ea_ref, Ef):
    import torch
    f = lambda a: a.to(torch.float64)
    xf = f(Ef).mean(dim=1) if Ef.dtype == torch.float64 else _mean_blocks(Ef)
    xa = _mean_blocks(Ea)
    xr = _mean_blocks(Ea_ref)
    inc = (xr - xf).abs().max().item()
    dmean = (xa - xr).abs().max().item()
    da = ar = 0.0
    for k in range(Ea.shape[1]):
        a = f(Ea[:, k]) - xa
        r = f(Ea_ref[:, k]) - xr
        da = max(da, (a - r).abs().max().item())
        ar = max(ar, r.abs().max().item())
    return inc, dmean, da, ar


def _mean_blocks(A):
    import torch
    s = torch.zeros(A.shape[0], dtype=torch.float64, device=A.device)
    for k in range(A.shape[1]):
        s += A[:, k].to(torch.float64)
    return s / A.shape[1]


def ensemble_errors(Ea, Ea_ref, Ef):
    if hasattr(Ea, "detach"):
        inc, dmean, da, ar = _errors_torch(Ea, Ea_ref, Ef)
    else:
        inc, dmean, da, ar = _errors_np(Ea, Ea_ref, Ef)
    tiny = np.finfo(float).tiny
    return {"mean_err": float(dmean / max(inc, tiny)), "spread_err": float(da / max(ar, tiny))}
