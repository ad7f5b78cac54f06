"""CPU checks of the likelihood reference: the dense p x p evaluation agrees
with the ensemble-space one (Woodbury: d^T D^-1 d = d^T R^-1 d - b^T M^-1 b,
det D = det R det M with M = I + S^T R^-1 S), which is the form the GPU
adjoint evaluates; and its autograd gradient agrees with central differences."""
import numpy as np
import torch

from tests.nll_ref import nll_reference, selection


def _small():
    from oak_amd import synth
    from oak_amd.nll import problem_from_ensemble
    P = synth.config3(nx=12, N=8, m=40, seed_cfg=31)
    prob = problem_from_ensemble(P)
    return P, prob


def test_dense_matches_ensemble_space():
    P, prob = _small()
    sel = selection(P, prob.HSf)
    ls = torch.as_tensor(0.5 * np.log(P.rmse2))
    L, a = 4.0, 1.2
    J = float(nll_reference(sel, prob.HSf, prob.Hxf, prob.yo, ls, torch.tensor(L), torch.tensor(a)))
    Je = 0.0
    for idx, dist in sel:
        if len(idx) == 0:
            continue
        w = np.exp(-(dist / L) ** 2)
        r = w * w / P.rmse2[idx]
        S = a * prob.HSf[idx]
        d = prob.yo[idx] - prob.Hxf[idx]
        M = np.eye(P.N) + S.T @ (r[:, None] * S)
        b = S.T @ (r * d)
        Je += 0.5 * (np.sum(r * d * d) - b @ np.linalg.solve(M, b) + np.sum(np.log(1.0 / r))
                     + np.linalg.slogdet(M)[1])
    # D holds sigma^2/w^2 up to ~1e9 for the observations near the cutoff: the
    # dense solve and log-determinant round to ~1e-10 of J; a dropped or
    # wrong term would show at 1e-3 or more
    assert abs(J - Je) <= 1e-9 * abs(Je)


def test_reference_gradient_vs_central_differences():
    P, prob = _small()
    sel = selection(P, prob.HSf)
    ls = torch.as_tensor(0.5 * np.log(P.rmse2)).clone().requires_grad_(True)
    L = torch.tensor(4.0, dtype=torch.float64, requires_grad=True)
    a = torch.tensor(1.2, dtype=torch.float64, requires_grad=True)
    J = nll_reference(sel, prob.HSf, prob.Hxf, prob.yo, ls, L, a)
    J.backward()
    h = 1e-6
    f = lambda ls_, L_, a_: float(nll_reference(sel, prob.HSf, prob.Hxf, prob.yo, ls_, L_, a_))
    base = ls.detach()
    for l in (0, 7, 23):
        e = torch.zeros_like(base); e[l] = h
        fd = (f(base + e, L.detach(), a.detach()) - f(base - e, L.detach(), a.detach())) / (2 * h)
        assert abs(fd - float(ls.grad[l])) <= 1e-6 * max(1.0, abs(fd)), l
    fdL = (f(base, L.detach() + h, a.detach()) - f(base, L.detach() - h, a.detach())) / (2 * h)
    fda = (f(base, L.detach(), a.detach() + h) - f(base, L.detach(), a.detach() - h)) / (2 * h)
    assert abs(fdL - float(L.grad)) <= 1e-6 * max(1.0, abs(fdL))
    assert abs(fda - float(a.grad)) <= 1e-6 * max(1.0, abs(fda))
