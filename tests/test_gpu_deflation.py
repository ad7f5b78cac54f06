"""The QL deflation tests of k_tql (tridiag.hip) on the zones they change:
N = 128 with zones of p < N observations (N - p unit eigenvalues of M, where
the FP64 path deflates rounding-level off-diagonals by eps ||T|| and the
mixed path everywhere) and strongly observed ones (sigma_SSH = 0.03, ||M||
up to about 1e5), all through the ensemble-space eigensolver (no
observation-space bucket): FP64 against the oracle (LAPACK dsyev) at 1e-10,
the mixed path at 1e-4, and the two eigensolvers of the library (QL, Jacobi)
against each other.  Needs an MI355X."""
import numpy as np
import pytest

from oak_amd.errors import ensemble_errors
from oracle import oracle as O
from tests.kfcheck import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    from oak_amd import synth
    O.load_lapack()
    # 192 columns, local counts 15..175: about 50 with p < 64, 100 with
    # 64 <= p < 128, 35 with p >= 128
    P = synth.column_ocean(16, 12, 1, 128, 250, 4, seed_cfg=910, hlen=3, vlen=1, L=3.0,
                           cutoff=5.0, sig_ssh=0.03)
    geo = dict(zone_ptr=P.zone_ptr, zpos=P.zpos, L=P.corr_len, cutoff=P.max_len, metric=0,
               loctype=1, ox=P.ox, oy=P.oy)
    Eo, _, cnt = O.ensemble_analysis(geo, (P.Hi, P.Hj, P.Hs), P.Hshift, P.yo, P.rmse2, P.E,
                                     inflation=1.03, max_correction=0.5)
    yield P, Eo, cnt
    O.set_backend(0)


def _run(P, E, eig=None, ctx_tune=None):
    from oak_amd import assimilation as A
    from oak_amd._lib import Context
    geo = A.Geometry(P.zone_ptr, P.zpos, P.corr_len, P.max_len, P.ox, P.oy, metric=0, loctype=1)
    H = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    ctx = Context(0)
    ctx.set_tuning(dual=0, **(ctx_tune or {}))
    if eig is not None:
        ctx.set_eigensolver(eig)
    Ea = A.ensemble_analysis(geo, H, P.Hshift, P.yo, P.rmse2, E, inflation=1.03,
                             max_correction=0.5, ctx=ctx)
    st = ctx.stats()
    ctx.close()
    return Ea, st


def test_workload_has_rank_deficient_and_strong_zones(problem):
    P, _, cnt = problem
    assert ((cnt >= 64) & (cnt < 128)).sum() >= 50 and (cnt >= 128).sum() >= 20
    assert (cnt < 64).sum() >= 20


@pytest.mark.parametrize("tune", [dict(), dict(order=1, tri_split=1, persist=2)])
def test_fp64_vs_oracle(problem, tune):
    P, Eo, _ = problem
    Ea, st = _run(P, P.E, ctx_tune=tune)
    assert st["eig_fallbacks"] == 0
    xf = P.E.mean(axis=1)
    assert rel_err(Ea - xf[:, None], Eo - xf[:, None]) < 1e-10


def test_mixed_vs_oracle(problem):
    P, Eo, _ = problem
    E32 = np.asfortranarray(P.E, dtype=np.float32)
    Ea, st = _run(P, E32)
    assert st["eig_fallbacks"] == 0
    e = ensemble_errors(Ea, Eo, P.E)
    assert e["mean_err"] < 1e-4 and e["spread_err"] < 1e-4, e


def test_ql_vs_jacobi(problem):
    """The Jacobi eigensolver has no deflation test of this kind: the two
    agree to the FP64 bound."""
    from oak_amd._lib import OAKG_EIG_JACOBI, OAKG_EIG_QL
    P, _, _ = problem
    Eq, _ = _run(P, P.E, eig=OAKG_EIG_QL)
    Ej, _ = _run(P, P.E, eig=OAKG_EIG_JACOBI)
    xf = P.E.mean(axis=1)
    assert rel_err(Eq - xf[:, None], Ej - xf[:, None]) < 1e-10
