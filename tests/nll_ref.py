"""FP64 reference of the local innovation likelihood, written from its
definition with the dense p x p innovation covariance of each zone:

    J = sum_z 1/2 [ d^T D^-1 d + log det D ],  D = a^2 S S^T + diag(sigma^2 / w^2)

(selection by the oracle; w = exp(-(dist/L)^2) evaluated in torch so that
autograd gives the reference gradient).  Independent of the ensemble-space
evaluation and of the hand-written adjoint of the library."""
import numpy as np
import torch

from oracle import oracle as O


def selection(P, prob_hsf, cutoff=None):
    """Per zone: (obs indices, distances) inside the cutoff (oracle selection)."""
    out = []
    for z in range(P.nzones):
        w, rel, cnt = O.select(P.metric, P.loctype, P.zpos[z], P.corr_len[z],
                               P.max_len[z] if cutoff is None else cutoff, P.ox, P.oy)
        idx = np.nonzero(rel)[0]
        d = np.sqrt((P.ox[idx] - P.zpos[z, 0]) ** 2 + (P.oy[idx] - P.zpos[z, 1]) ** 2)
        out.append((idx, d))
    return out


def nll_reference(sel, HSf, Hxf, yo, log_sigma, L, infl):
    """J (torch scalar, differentiable in log_sigma [m], L and infl)."""
    HSf = torch.as_tensor(HSf, dtype=torch.float64)
    innov = torch.as_tensor(yo - Hxf, dtype=torch.float64)
    sig2 = torch.exp(2.0 * log_sigma)
    J = torch.zeros((), dtype=torch.float64)
    for idx, dist in sel:
        if len(idx) == 0:
            continue
        ii = torch.as_tensor(idx)
        w = torch.exp(-(torch.as_tensor(dist) / L) ** 2)
        S = infl * HSf[ii]
        D = S @ S.T + torch.diag(sig2[ii] / (w * w))
        d = innov[ii]
        J = J + 0.5 * (d @ torch.linalg.solve(D, d) + torch.logdet(D))
    return J
