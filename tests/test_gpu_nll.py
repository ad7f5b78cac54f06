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


def test_tuning_recovers_generating_sigma_and_inflation():
    """One zone holding all m observations (w = 1), innovations drawn from the
    model itself, d ~ N(0, a^2 S S^T + sigma^2 I): Adam on a tied log sigma
    and the log inflation (L fixed) finds the generating values, sigma to the
    accuracy m = 2000 residuals give (a few %), the inflation to that of N = 32
    directions."""
    import torch
    from oak_amd.assimilation import Geometry
    from oak_amd.nll import NLLProblem, tune
    g = np.random.default_rng(41)
    m, N, sig, a_true = 2000, 32, 0.5, 1.3
    S = g.standard_normal((m, N)) / np.sqrt(N - 1)
    d = a_true * S @ g.standard_normal(N) + sig * g.standard_normal(m)
    Hxf = g.standard_normal(m)
    dev = torch.device("cuda:0")
    t = lambda x, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(x)).to(dev, dt)
    ox, oy = g.uniform(-1, 1, m), g.uniform(-1, 1, m)
    geo = Geometry(t(np.array([0, 1]), torch.int64), t(np.zeros((1, 2))), t(np.array([1e6])),
                   t(np.array([1e9])), t(ox), t(oy))
    prob = NLLProblem(geo, t(Hxf), t(Hxf + d), t(S.T).t(), N)
    ls, L1, a1, hist = tune(prob, torch.tensor([float(np.log(1.5))], dtype=torch.float64),
                            1e6, 0.6, steps=300, lr=0.05, classes=np.zeros(m, np.int64),
                            fix_L=True)
    s1 = float(torch.exp(ls[0]))
    assert hist[-1] < hist[0]
    assert abs(s1 / sig - 1.0) < 0.08, s1
    assert abs(a1 / a_true - 1.0) < 0.5, a1
    assert L1 == pytest.approx(1e6)
