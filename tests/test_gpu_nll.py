"""Innovation likelihood and its hand-written adjoint on the GPU
(oakg_local_nll) against the dense FP64 reference (tests/nll_ref.py), central
finite differences of the GPU value (h = 1e-6) and a short Adam tuning.
Needs an MI355X."""
import numpy as np
import pytest
import torch

from tests.nll_ref import nll_reference, selection

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    from oak_amd import synth
    from oak_amd.nll import problem_from_ensemble
    P = synth.config3(nx=24, N=12, m=120, seed_cfg=33)
    return P, problem_from_ensemble(P), problem_from_ensemble(P, device="cuda:0")


def _ref(P, prob, ls, L, a):
    sel = selection(P, prob.HSf)
    lsr = torch.as_tensor(ls).clone().requires_grad_(True)
    Lr = torch.tensor(L, dtype=torch.float64, requires_grad=True)
    ar = torch.tensor(a, dtype=torch.float64, requires_grad=True)
    J = nll_reference(sel, prob.HSf, prob.Hxf, prob.yo, lsr, Lr, ar)
    J.backward()
    return float(J.detach()), lsr.grad.numpy(), float(Lr.grad), float(ar.grad)


@pytest.mark.parametrize("L,a", [(5.0, 1.0), (3.0, 1.3)])
def test_nll_value_and_gradient_vs_reference(small, L, a):
    from oak_amd.nll import evaluate
    P, prob, _ = small
    ls = 0.5 * np.log(P.rmse2) + 0.3
    J, gth, gLz, ga = evaluate(prob, np.exp(2 * ls), np.full(P.nzones, L), a)
    Jr, gthr, gLr, gar = _ref(P, prob, ls, L, a)
    assert abs(J - Jr) <= 1e-9 * abs(Jr)
    scale = np.abs(gthr).max()
    assert np.abs(gth - gthr).max() <= 1e-8 * scale
    assert abs(gLz.sum() - gLr) <= 1e-8 * max(1.0, abs(gLr))
    assert abs(ga - gar) <= 1e-8 * max(1.0, abs(gar))


def test_nll_gradient_vs_central_differences(small):
    from oak_amd.nll import evaluate
    P, prob, _ = small
    nz = P.nzones
    ls = 0.5 * np.log(P.rmse2)
    L, a, h = 4.0, 1.1, 1e-6
    f = lambda ls_, L_, a_: evaluate(prob, np.exp(2 * ls_), np.full(nz, L_), a_)[0]
    J, gth, gLz, ga = evaluate(prob, np.exp(2 * ls), np.full(nz, L), a)
    tol = 1e-6 * max(1.0, np.abs(gth).max())
    for l in (0, 1, 57, 118):
        e = np.zeros_like(ls); e[l] = h
        fd = (f(ls + e, L, a) - f(ls - e, L, a)) / (2 * h)
        assert abs(fd - gth[l]) <= tol, (l, fd, gth[l])
    fdL = (f(ls, L + h, a) - f(ls, L - h, a)) / (2 * h)
    fda = (f(ls, L, a + h) - f(ls, L, a - h)) / (2 * h)
    assert abs(fdL - gLz.sum()) <= 1e-6 * max(1.0, abs(fdL))
    assert abs(fda - ga) <= 1e-6 * max(1.0, abs(fda))


def test_nll_device_matches_host(small):
    from oak_amd.nll import evaluate
    P, prob, dprob = small
    nz = P.nzones
    ls = 0.5 * np.log(P.rmse2)
    J, gth, gLz, ga = evaluate(prob, np.exp(2 * ls), np.full(nz, 4.0), 1.1)
    d = dprob.yo.device
    Jd, gthd, gLzd, gad = evaluate(dprob, torch.as_tensor(np.exp(2 * ls), device=d),
                                   torch.full((nz,), 4.0, dtype=torch.float64, device=d), 1.1)
    assert float(Jd) == J and float(gad) == ga
    np.testing.assert_array_equal(gLzd.cpu().numpy(), gLz)
    np.testing.assert_allclose(gthd.cpu().numpy(), gth, rtol=1e-13, atol=1e-13 * np.abs(gth).max())


def test_autograd_function_and_adam_decrease(small):
    from oak_amd.nll import innovation_nll, tune
    P, _, dprob = small
    d = dprob.yo.device
    ls = torch.full((P.m,), float(np.log(0.3)), dtype=torch.float64, device=d, requires_grad=True)
    lL = torch.tensor(float(np.log(3.0)), dtype=torch.float64, device=d, requires_grad=True)
    la = torch.tensor(0.0, dtype=torch.float64, device=d, requires_grad=True)
    J = innovation_nll(ls, lL, la, dprob)
    J.backward()
    assert torch.isfinite(ls.grad).all() and torch.isfinite(lL.grad) and torch.isfinite(la.grad)
    _, L1, a1, hist = tune(dprob, ls.detach(), 3.0, 1.0, steps=10, lr=0.05)
    assert hist[-1] < hist[0]
