"""Innovation likelihood of the local analysis, differentiated by the
hand-written adjoint kernel of oakg_local_nll, behind torch.autograd.

For each zone z, with its local observations (taper weights w = exp(-(d/L)^2)
inside the fixed cutoff), innovations d = yo - Hxf, observation anomalies S
(HSf rows) and inflation a:

    J = sum_z 1/2 [ d^T D_z^-1 d + log det D_z ],  D_z = a^2 S S^T + diag(sigma^2 / w^2)

the negative log-likelihood (up to a constant) of the local innovations
under the localized ensemble covariance: the quantity whose minimum over
(log sigma, L, a) tunes OAK's analysis hyperparameters.  OAK itself has no
such objective; it reuses the analysis' own per-zone solve (rrsqrt.F90:
135-142), so its value and gradient cost about one analysis.
"""
from dataclasses import dataclass
import ctypes as C

import numpy as np

from . import _lib
from .assimilation import Geometry, _is_dev
from .rrsqrt import default_context


@dataclass
class NLLProblem:
    """Inputs of the likelihood that stay fixed while the hyperparameters move.

    All arrays on the same device (numpy on the host, or torch tensors on the
    GPU): geometry (its corr_len is replaced by L at each evaluation; its
    max_len, the cutoffs, stays), Hxf [m], yo [m], HSf [m, N] column-major."""
    geo: Geometry
    Hxf: object
    yo: object
    HSf: object
    N: int

    @property
    def m(self):
        return len(self.yo)


def problem_from_ensemble(P, device=None):
    """NLLProblem of a synthetic problem (oak_amd.synth): obs-space mean and
    anomalies of the ensemble (assimilation.F90:3112-3134)."""
    HE = np.zeros((P.m, P.N))
    np.add.at(HE, P.Hi, P.Hs[:, None] * P.E[P.Hj, :])
    Hxf = HE.mean(axis=1)
    HSf = np.asfortranarray((HE - Hxf[:, None]) / np.sqrt(P.N - 1))
    geo = Geometry(P.zone_ptr, P.zpos, P.corr_len, P.max_len, P.ox, P.oy, metric=P.metric,
                   loctype=P.loctype)
    if device is None:
        return NLLProblem(geo.host_arrays(), Hxf, np.ascontiguousarray(P.yo, dtype=np.float64),
                          HSf, P.N)
    import torch
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(device, dt)
    g = Geometry(t(P.zone_ptr, torch.int64), t(P.zpos), t(P.corr_len), t(P.max_len), t(P.ox),
                 t(P.oy), metric=P.metric, loctype=P.loctype)
    return NLLProblem(g, t(Hxf), t(P.yo), t(HSf.T).t(), P.N)


def evaluate(prob, sigma2, corr_len, inflation, ctx=None, per_zone=False):
    """J and its gradient through the C ABI.

    sigma2 [m] (observation error variances), corr_len [nzones] (L per zone),
    inflation (float).  Returns (J, dJ/dlog sigma [m], dJ/dL_z [nzones],
    dJ/dinflation[, J_z]) as numpy (host problem) or torch tensors (device)."""
    ctx = ctx or default_context()
    L = _lib.load()
    g = prob.geo
    nz = g.nzones
    m = prob.m
    dev = _is_dev(prob.yo)
    geo = Geometry(g.zone_ptr, g.zpos, corr_len, g.max_len, g.ox, g.oy, g.oz, g.ot, g.metric,
                   g.loctype)
    if dev:
        import torch
        kw = dict(dtype=torch.float64, device=prob.yo.device)
        out = torch.zeros(2, **kw)
        gth = torch.zeros(m, **kw)
        gL = torch.zeros(nz, **kw)
        Jz = torch.zeros(nz, **kw) if per_zone else None
        ld = prob.HSf.stride(1) if m > 0 else 1
        z, o = geo.structs(prob.yo, sigma2)
        rc = L.oakg_local_nll(ctx.handle, C.byref(z), C.byref(o), prob.N, prob.Hxf.data_ptr(),
                              prob.HSf.data_ptr(), ld, float(inflation), out.data_ptr(),
                              gth.data_ptr(), gL.data_ptr(), out.data_ptr() + 8,
                              Jz.data_ptr() if per_zone else None, _lib.OAKG_DEVICE_PTRS)
        ctx.check(rc)
        res = (out[0], gth, gL, out[1])
    else:
        sigma2 = np.ascontiguousarray(sigma2, dtype=np.float64)
        corr_len = np.ascontiguousarray(corr_len, dtype=np.float64)
        geo = Geometry(g.zone_ptr, g.zpos, corr_len, g.max_len, g.ox, g.oy, g.oz, g.ot, g.metric,
                       g.loctype)
        out = np.zeros(2)
        gth = np.zeros(m)
        gL = np.zeros(nz)
        Jz = np.zeros(nz) if per_zone else None
        HSf = np.asfortranarray(prob.HSf, dtype=np.float64)
        z, o = geo.structs(prob.yo, sigma2)
        rc = L.oakg_local_nll(ctx.handle, C.byref(z), C.byref(o), prob.N,
                              np.ascontiguousarray(prob.Hxf).ctypes.data, HSf.ctypes.data, m,
                              float(inflation), out.ctypes.data, gth.ctypes.data, gL.ctypes.data,
                              out.ctypes.data + 8, Jz.ctypes.data if per_zone else None, 0)
        ctx.check(rc)
        res = (float(out[0]), gth, gL, float(out[1]))
    return res + ((Jz,) if per_zone else ())


def _autograd():
    import torch

    class InnovationNLL(torch.autograd.Function):
        """J(log sigma, log L, log inflation) with a global L and inflation;
        the backward pass returns the gradients the adjoint kernel computed
        in the forward pass."""

        @staticmethod
        def forward(fctx, log_sigma, log_L, log_infl, prob, octx):
            sigma2 = torch.exp(2.0 * log_sigma.detach()).contiguous()
            Lv = float(torch.exp(log_L.detach()))
            av = float(torch.exp(log_infl.detach()))
            corr = torch.full((prob.geo.nzones,), Lv, dtype=torch.float64,
                              device=sigma2.device)
            J, gth, gLz, ga = evaluate(prob, sigma2, corr, av, octx)
            fctx.save_for_backward(gth, (gLz.sum() * Lv).reshape(()), (ga * av).reshape(()))
            return J.clone()

        @staticmethod
        def backward(fctx, gout):
            gth, gl, ga = fctx.saved_tensors
            return gout * gth, gout * gl, gout * ga, None, None

    return InnovationNLL


_fn = None


def innovation_nll(log_sigma, log_L, log_infl, prob, ctx=None):
    """Differentiable J (torch): log_sigma [m] and the scalars log L and
    log inflation on the problem's device."""
    global _fn
    if _fn is None:
        _fn = _autograd()
    return _fn.apply(log_sigma, log_L, log_infl, prob, ctx or default_context())


def tune(prob, log_sigma0, L0, infl0, steps=50, lr=0.05, ctx=None, callback=None,
         classes=None, fix_L=False):
    """Adam on (log sigma, log L, log inflation) minimising J; returns the
    tuned per-observation log sigma, L, the inflation and the history of J.

    classes: optional integer class of each observation; sigma is then one
    parameter per class (log_sigma0 per class), which is what a single
    innovation vector can identify: a free sigma per observation fits each
    residual.  fix_L keeps the localisation length at L0."""
    import torch
    dev = prob.yo.device
    ls = torch.nn.Parameter(torch.as_tensor(log_sigma0, dtype=torch.float64, device=dev).clone())
    lL = torch.nn.Parameter(torch.tensor(float(np.log(L0)), dtype=torch.float64, device=dev))
    la = torch.nn.Parameter(torch.tensor(float(np.log(infl0)), dtype=torch.float64, device=dev))
    cls = None
    if classes is not None:
        cls = torch.as_tensor(np.asarray(classes), dtype=torch.int64, device=dev)
    expand = (lambda p: p[cls]) if cls is not None else (lambda p: p)
    opt = torch.optim.Adam([ls, la] if fix_L else [ls, lL, la], lr=lr)
    hist = []
    for it in range(steps):
        opt.zero_grad()
        J = innovation_nll(expand(ls), lL, la, prob, ctx)
        J.backward()
        hist.append(float(J.detach()))
        if callback:
            callback(it, float(J.detach()), ls, lL, la)
        opt.step()
    with torch.no_grad():
        J = innovation_nll(expand(ls), lL, la, prob, ctx)
        hist.append(float(J))
        out = expand(ls).detach()
    return out, float(torch.exp(lL.detach())), float(torch.exp(la.detach())), hist
